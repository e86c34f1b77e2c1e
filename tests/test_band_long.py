"""Long reads (1024+ bases) on the band that follows the alignment, in FP32
(the default, rescaled every 16 rows; four pairs of a read per wavefront, a
read's last one to three pairs one per wavefront) and in FP64 (rescaled every
64 rows).

The moving-band kernel against the oracle's restatement of the same band
(bvo_band_follow_forward: likelihood, touch mass, dropped share and the
accept decision), and the end-to-end result against the full-matrix FP64
oracle at |delta log10 L| <= 1e-3.  Inputs: the long-read generator (reads of
2-10 kb with 5-10 % errors, 60 % of them 1 bp indels) and planted drifts.
"""
import numpy as np
import pytest

from biovoltron_amd import band, pairhmm, synth

pytestmark = pytest.mark.gpu

TOL = 1e-3
FORM_FP64 = 1
FORM_BAND = {64: 4, 32: 5}  # the moving band's form per arithmetic
# the kernel against the oracle's FP64 moving band: FP64 rounding, or FP32
# rounding over reads of up to 10 kb
BAND_EPS = {64: 1e-8, 32: 2e-4}
VERIFY_ABOVE = -28.0 - 120.0 * np.log10(2.0) - 0.05
# the moving band's rules: moves that would drop more than 1e-9 of a row's
# mass are held; the touch mass counts the insertion runs leaving the bottom
RULES = dict(hold_eps=1e-9, irun=2)


def say(*a):
    print("[band-long]", *a, flush=True)


def long_sample(regions=1, seed=None):
    b = synth.generate(2, seed=seed, count=regions)
    sel = []
    for g in range(b.n_regions):
        reads, haps = b.region(g)
        for rd in reads:
            for h in haps:
                sel.append((rd, h))
    return b, sel


def run_plan(pb, bands=None, model=None, bits=32):
    with pairhmm.Plan(pb) as plan:
        plan.set_long_precision(bits)
        assert plan.long_precision() == bits
        if bands is not None:
            band.set_plan_bands(plan, bands)
        if model is not None:
            band.set_plan_model(plan, model)
        plan.run()
        out, form = plan.results()
        st = plan.stats()
        used = band.plan_bands(plan) if (bands is not None or model is not None) else None
    return out, form, st, used


def check_against_oracle(oracle, sel, out, form, used, tol, bits):
    n_acc = n_rej = 0
    worst = 0.0
    for k, (rd, hap) in enumerate(sel):
        b0, b1 = int(used[k, 0]), int(used[k, 1])
        if b0 > b1:
            assert form[k] in (0, FORM_FP64), k
            continue
        o = b0 - (64 - (b1 - b0 + 1)) // 2
        v, ratio, lost, _ = oracle.band_follow(rd, hap, o, **RULES)
        if form[k] == FORM_BAND[bits]:
            n_acc += 1
            worst = max(worst, abs(out[k] - v))
            assert ratio < 10 * tol and v < VERIFY_ABOVE + 1e-6, k
        else:
            n_rej += 1
            assert ratio > 0.1 * tol or v >= VERIFY_ABOVE - 1e-6, k
    return n_acc, n_rej, worst


@pytest.mark.parametrize("bits", [32, 64])
def test_long_band_matches_oracle_and_full_matrix(oracle, bits):
    """Learned mode on long-read regions: start bands from the seeds, the
    moving band for every long pair that has one.  Accepted pairs equal the
    oracle's FP64 moving band to the kernel's rounding; rejected ones took
    the full matrix; every result is within 1e-3 of the full-matrix FP64
    oracle."""
    b, sel = long_sample(regions=2)
    pb = pairhmm.PackedBatch.from_synth(b)
    say("long pairs", len(sel), "cells", b.cells)
    full, ff, _, _ = run_plan(pb)
    model = band.BandModel.load(band.DEFAULT_MODEL)
    out, form, st, used = run_plan(pb, model=model, bits=bits)
    tol = pairhmm.default_options().band_tol
    n_acc, n_rej, worst = check_against_oracle(oracle, sel, out, form, used, tol, bits)
    err = float(np.abs(out - full).max())
    say(f"{bits}-bit band: accepted {n_acc}, rejected {n_rej}, max |dlog10| vs oracle band {worst:.3g}, "
        f"vs full {err:.3g}, long band pairs {st.long_band_pairs}, fp64 cells {st.fp64_cells}")
    assert n_acc == st.long_band_pairs == int((form == FORM_BAND[bits]).sum())
    assert n_acc >= 0.9 * len(sel)
    assert worst <= BAND_EPS[bits]
    assert err <= TOL
    # the banded pairs computed far fewer FP64 cells than the full matrix
    assert st.fp64_cells < 0.1 * b.cells
    # a few against the full-matrix oracle directly
    for k in range(0, len(sel), max(1, len(sel) // 6)):
        rd, hap = sel[k]
        ex = oracle.log10_exact(rd, hap, 0)
        say(f"pair {k}: R={len(rd['seq'])} H={len(hap)} form {form[k]} gpu {out[k]:.6f} exact {ex:.6f}")
        assert abs(out[k] - ex) <= TOL


@pytest.mark.parametrize("bits", [32, 64])
def test_long_band_misplaced_start_falls_back(bits):
    """Start bands 300 diagonals away from the alignment: the moving band
    cannot reach it, the touch mass says so, and the full matrix decides."""
    b, sel = long_sample(regions=1)
    pb = pairhmm.PackedBatch.from_synth(b)
    full, ff, _, _ = run_plan(pb)
    model = band.BandModel.load(band.DEFAULT_MODEL)
    _, _, _, used = run_plan(pb, model=model)
    good = used[:, 0] <= used[:, 1]
    bands = used.copy()
    bands[good] += 300
    out, form, st, _ = run_plan(pb, bands=bands, bits=bits)
    say(bits, "misplaced", int(good.sum()), "rejected", int((form[good] != FORM_BAND[bits]).sum()))
    assert (form[good] != FORM_BAND[bits]).mean() >= 0.9
    rej = form != FORM_BAND[bits]
    assert np.array_equal(out[rej], full[rej])
    assert np.abs(out - full).max() <= TOL


def _drift_pair(rng, hap, pieces, q=30):
    seq = b"".join(hap[a:b] for a, b in pieces)
    R = len(seq)
    rd = dict(seq=seq, q=bytes([q]) * R, i=bytes([45]) * R, d=bytes([45]) * R, c=bytes([10]) * R)
    return rd


@pytest.mark.parametrize("bits", [32, 64])
def test_long_band_follows_planted_drift(oracle, bits):
    """Reads whose alignment drifts by 60-100 diagonals (20-base deletions
    and insertions), started 12 diagonals off centre: accepted, equal to the
    oracle's moving band (to the kernel's rounding) and to the full matrix."""
    rng = np.random.default_rng(11)
    hap = bytes(rng.choice(list(b"ACGT"), 3000).astype(np.uint8))
    ins = bytes(rng.choice(list(b"ACGT"), 20).astype(np.uint8))
    reads = [
        _drift_pair(rng, hap, [(100, 700), (720, 1300), (1320, 1900), (1920, 2500)], q=20),
        _drift_pair(rng, hap, [(50, 900), (920, 1700), (1720, 2600), (2620, 2900)], q=20),
    ]
    r3 = _drift_pair(rng, hap, [(300, 1500)], q=10)
    r3 = dict(r3, seq=r3["seq"][:600] + ins + r3["seq"][600:], q=r3["q"] + bytes([10]) * 20,
              i=r3["i"] + bytes([45]) * 20, d=r3["d"] + bytes([45]) * 20, c=r3["c"] + bytes([10]) * 20)
    reads.append(r3)
    starts = [100, 50, 300]
    sel = [(rd, hap) for rd in reads]
    pb = pairhmm.PackedBatch.from_regions([([rd], [hap]) for rd in reads])
    bands = np.array([[s - 32 + 12, s + 31 + 12] for s in starts], np.int32)
    out, form, st, used = run_plan(pb, bands=bands, bits=bits)
    for k, (rd, h) in enumerate(sel):
        v, ratio, lost, offs = oracle.band_follow(rd, h, int(bands[k, 0]), **RULES)
        ex = oracle.log10_exact(rd, h, 0)
        say(f"drift read {k}: R={len(rd['seq'])} form {form[k]} gpu {out[k]:.9f} band {v:.9f} "
            f"exact {ex:.9f} ratio {ratio:.2e} lost {lost:.2e} final offset {offs[-1]}")
        assert form[k] == FORM_BAND[bits]
        assert abs(out[k] - v) <= BAND_EPS[bits]
        assert abs(out[k] - ex) <= 1e-6 + BAND_EPS[bits]


@pytest.mark.parametrize("bits", [32, 64])
def test_long_band_haplotypes_drifting_apart(oracle, bits):
    """One long read against six haplotypes whose alignments drift in
    different directions (20-30 base insertions or deletions in the
    haplotype at different places, a longer prefix): in FP32 four of the pairs share a
    wavefront, each band moving on its own, and the other two run one per
    wavefront.  Every pair is accepted and equals the oracle's moving band
    (to the kernel's rounding) and the full-matrix oracle."""
    rng = np.random.default_rng(23)
    base = bytes(rng.choice(list(b"ACGT"), 3200).astype(np.uint8))
    extra = bytes(rng.choice(list(b"ACGT"), 20).astype(np.uint8))
    # base quality 10: log10 L of -135 to -182, out of FP32's reach (a pair
    # FP32 computes takes the FP32 full-matrix result, as in the reference).
    # The read skips 15 bases at row 1200 (+15 diagonals on every haplotype
    # that keeps them); the haplotypes add or drop 20-30 bases elsewhere.
    rd = _drift_pair(rng, base, [(200, 1400), (1415, 2700)], q=10)
    haps = [
        base,                                                # +15 at row 1200
        base[:900] + extra + base[900:],                     # +20 at row 700, then +15
        base[:1800] + base[1830:],                           # +15, then -30 at row 1600
        extra + base,                                        # starts at 220
        base[:500] + extra + base[500:1000] + base[1020:],   # +20 at row 300, -20 at row 800
        base[:2000] + base[2020:],                           # +15, then -20 at row 1800
    ]
    starts = [200, 200, 200, 220, 200, 200]
    sel = [(rd, h) for h in haps]
    pb = pairhmm.PackedBatch.from_regions([([rd], haps)])
    bands = np.array([[s - 32 + 6, s + 31 + 6] for s in starts], np.int32)
    out, form, st, used = run_plan(pb, bands=bands, bits=bits)
    for k, (r_, h) in enumerate(sel):
        v, ratio, lost, offs = oracle.band_follow(r_, h, int(bands[k, 0]), **RULES)
        ex = oracle.log10_exact(r_, h, 0)
        say(f"hap {k}: R={len(r_['seq'])} H={len(h)} form {form[k]} gpu {out[k]:.9f} band {v:.9f} "
            f"exact {ex:.9f} ratio {ratio:.2e} offsets {offs[0]} -> {offs[-1]}")
        assert form[k] == FORM_BAND[bits]
        assert abs(out[k] - v) <= BAND_EPS[bits]
        assert abs(out[k] - ex) <= 1e-6 + BAND_EPS[bits]
    assert st.long_band_pairs == len(sel)
