"""Banded PairHMM, band labels and the band MLP.

CPU: labels from the product's forward-backward vs the oracle's posteriors.
GPU: banded evaluation vs the full matrix, fallback on bad bands, MLP
gradients vs a torch FP32 reference, training, and the learned-vs-heuristic
ablation.
"""
import numpy as np
import pytest

from conftest import load_pairs
from biovoltron_amd import band, pairhmm, synth


def say(*a):
    print("[band]", *a, flush=True)


def fixture_batch(name="workload1.npz", step=1):
    pairs, d = load_pairs(name)
    sel = pairs[::step]
    return sel, pairhmm.PackedBatch.from_regions([([r], [h]) for r, h in sel])


def read_groups(step, n_haps=4, seed=5):
    """Fixture pairs as one read against n_haps haplotypes each: the pair's
    own and n_haps - 1 copies with 2 % of their bases substituted, one
    region per read.  Returns the pairs in output order and the batch."""
    pairs, _ = load_pairs("workload1.npz")
    rng = np.random.default_rng(seed)
    regions, flat = [], []
    for rd, hap in pairs[::step]:
        h = np.frombuffer(hap, np.uint8)
        haps = [hap]
        for _ in range(n_haps - 1):
            m = h.copy()
            k = rng.random(m.size) < 0.02
            m[k] = rng.choice(list(b"ACGT"), int(k.sum()))
            haps.append(m.tobytes())
        regions.append(([rd], haps))
        flat += [(rd, x) for x in haps]
    return flat, pairhmm.PackedBatch.from_regions(regions)


def oracle_span(oracle, rd, hap, frac=band.LABEL_MASS):
    R = len(rd["seq"])
    _, mass, _ = oracle.band_posterior(rd, hap)
    need = mass.sum() - (1.0 - frac)   # at most 1 - frac expected visits outside
    best, lo, win = (0, len(mass) - 1), 0, 0.0
    for hi in range(len(mass)):
        win += mass[hi]
        while lo < hi and win - mass[lo] >= need:
            win -= mass[lo]
            lo += 1
        if win >= need and hi - lo < best[1] - best[0]:
            best = (lo, hi)
    return best[0] - (R - 1), best[1] - (R - 1)


def test_labels_match_oracle_posterior(oracle):
    sel, pb = fixture_batch(step=7)
    lab = band.labels(pb)
    for k, (rd, hap) in enumerate(sel):
        assert tuple(lab[k]) == oracle_span(oracle, rd, hap)


def test_heuristic_bands_are_centred_on_the_anchor():
    X = np.zeros((3, 36), np.float32)
    X[:, 1] = [300 / 512, 300 / 512, 2000 / 512]
    X[0, 20] = 1.0
    X[2, 20] = 1.0
    b = band.heuristic_bands(X, np.array([7, 7, 7], np.int32))
    assert tuple(b[0]) == (7 - 32, 7 + 31)
    assert b[1, 0] > b[1, 1] and b[2, 0] > b[2, 1]   # no seed hits / haplotype too long


def test_targets_roundtrip():
    lab = np.array([[10, 13], [-5, -5]], np.int32)
    anchor = np.array([8, -4], np.int32)
    T = band.targets(lab, anchor)
    assert np.allclose(T[:, 0] * 16 + anchor, lab.mean(1))
    assert np.allclose(2 ** T[:, 1], lab[:, 1] - lab[:, 0] + 1)


# ---- GPU ----------------------------------------------------------------------

def run_plan(pb, bands=None, model=None, shape=0, mid=None, **kw):
    with pairhmm.Plan(pb) as plan:
        plan.set_band_shape(shape)
        if mid is not None:
            plan.set_band_mid(mid)
        if bands is not None:
            band.set_plan_bands(plan, bands)
        if model is not None:
            band.set_plan_model(plan, model, **kw)
        plan.run()
        out, form = plan.results()
        st = plan.stats()
        used = band.plan_bands(plan) if (bands is not None or model is not None) else None
    return out, form, st, used


def small_pairs(n, seed, rmax=20, hmax=30):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        R = int(rng.integers(1, rmax + 1))
        H = int(rng.integers(1, hmax + 1))
        hap = bytes(rng.choice(list(b"ACGT"), H).astype(np.uint8))
        seq = bytes(rng.choice(list(b"ACGTN"), R, p=[.24, .24, .24, .24, .04]).astype(np.uint8))
        rd = dict(seq=seq, q=bytes(rng.integers(5, 41, R).astype(np.uint8)),
                  i=bytes([45]) * R, d=bytes([45]) * R, c=bytes([10]) * R)
        out.append((rd, hap))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("rmax,hmax", [(10, 20), (20, 40)])
def test_covering_band_matches_full(rmax, hmax):
    """Bands that hold the whole matrix (R + H + 1 diagonals: up to 31, the
    narrow two-pairs-per-wavefront kernel, or up to 61, the wide one) cut
    nothing: every pair is accepted and equals the full matrix up to FP32
    reassociation of the deletion scan."""
    sel = small_pairs(301, rmax + hmax, rmax, hmax)
    pb = pairhmm.PackedBatch.from_regions([([r], [h]) for r, h in sel])
    say("covering band", len(sel))
    full, ff, _, _ = run_plan(pb)
    bands = np.array([[-len(r["seq"]), len(h)] for r, h in sel], np.int32)
    out, form, st, _ = run_plan(pb, bands=bands)
    err = np.abs(out - full)
    say("max |dlog10|", float(err.max()))
    assert np.all(form == 2)
    assert st.fallback_pairs == 0 and st.band_pairs == len(sel)
    assert err.max() <= 2e-6


@pytest.mark.gpu
def test_label_bands_match_full(oracle):
    """64 diagonals centred on each pair's 99.999 % posterior span."""
    sel, pb = fixture_batch()
    lab = band.labels(pb)
    say("label bands", len(sel))
    full, ff, _, _ = run_plan(pb)
    out, form, st, _ = run_plan(pb, bands=lab)
    err = np.abs(out - full)
    say("max |dlog10| band vs full", float(err.max()), "band", st.band_pairs, "full pass",
        st.fallback_pairs, "cells", st.cells_swept)
    assert err.max() <= 1e-4
    assert st.fallback_pairs <= len(sel) // 50
    # fewer cells than the full matrix
    cells_full = sum(len(r["seq"]) * len(h) for r, h in sel)
    assert st.cells_swept < 0.5 * cells_full


@pytest.mark.gpu
def test_misplaced_bands_fall_back_to_full():
    """Bands far from the alignment put their mass on the edge zones: the
    banded kernel rejects them and the full matrix gives the result."""
    sel, pb = fixture_batch(step=4)
    lab = band.labels(pb)
    bands = (lab + 40).astype(np.int32)
    say("misplaced bands", len(sel))
    full, ff, _, _ = run_plan(pb)
    out, form, st, _ = run_plan(pb, bands=bands)
    say("rejected", int((form == 3).sum()), "of", len(sel))
    assert (form == 3).sum() >= 0.9 * len(sel)
    fb = form == 3
    # rejected pairs run on the full matrix: the same values up to FP32
    # rounding (the fallback list and the plain pass use different kernels)
    assert np.abs(out[fb] - full[fb]).max() <= 2e-5
    assert np.abs(out - full).max() <= 1e-3


@pytest.mark.gpu
def test_ineligible_bands_run_full():
    """Bands wider than the banded kernel, empty bands, and reads whose gap
    continuation varies along the read run on the full matrix."""
    sel = small_pairs(40, 2)
    rd, hap = sel[0]
    rd = dict(rd, c=bytes([10 + (k % 2) for k in range(len(rd["seq"]))]))
    sel[0] = (rd, hap)
    pb = pairhmm.PackedBatch.from_regions([([r], [h]) for r, h in sel])
    full, _, _, _ = run_plan(pb)
    bands = np.array([[-len(r["seq"]), len(h)] for r, h in sel], np.int32)
    bands[1] = (-100, 100)   # 201 diagonals
    bands[2] = (1, 0)        # no band
    out, form, st, _ = run_plan(pb, bands=bands)
    assert form[0] == 0 and form[1] == 0 and form[2] == 0
    # on the full matrix: the plain pass's values up to FP32 rounding
    assert np.abs(out[:3] - full[:3]).max() <= 2e-5
    assert np.all(form[3:] == 2)


def torch_loss_grad(params, X, T, w):
    import torch
    p = torch.tensor(params, dtype=torch.float32, requires_grad=True)
    o = 0
    W1 = p[o:o + 64 * 36].view(64, 36); o += 64 * 36
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 64 * 64].view(64, 64); o += 64 * 64
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 128].view(2, 64); o += 128
    b3 = p[o:o + 2]
    x = torch.tensor(X)
    t = torch.tensor(T)
    h1 = torch.relu(x @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    e = y - t
    loss = 0.5 * (w[0] * e[:, 0] ** 2 + w[1] * e[:, 1] ** 2).mean()
    loss.backward()
    return float(loss), p.grad.numpy()


SEED_K, MAX_OCC, END_TOL, FEAT_MAX_HAP = 12, 4, 8, 1024


def host_features(read, hap):
    """The seed features of one pair restated on the host (band_kernels.hip,
    band_features_kernel): lengths, quality summary, the lower median of the
    seeds' first-hit offsets (the anchor), the 32-bin offset histogram over
    every occurrence, and the read span of the seeds hitting near the anchor.
    Returns (x [36], anchor, ends) and whether a seed had more occurrences than
    the kernel counts (their order then depends on the table layout)."""
    seq, q = read["seq"], np.frombuffer(read["q"], np.uint8) & 127
    R, H = len(seq), len(hap)
    x = np.zeros(36, np.float32)
    x[0], x[1] = np.float32(R) / 256, np.float32(H) / 512
    x[2] = np.float32(np.float32(q.sum() / R) / np.float32(40))
    x[3] = np.float32(q.min()) / np.float32(40)
    nseeds = min(64, R - SEED_K + 1) if R >= SEED_K else 0
    stride = max(1, (R - SEED_K + 1 + 63) // 64) if nseeds else 1
    seeds = [p for p in (l * stride for l in range(nseeds))
             if p + SEED_K <= R and b"N" not in seq[p:p + SEED_K]]
    where = {}
    if 12 <= H <= FEAT_MAX_HAP:
        for j in range(H - SEED_K + 1):
            kmer = hap[j:j + SEED_K]
            if b"N" not in kmer:
                where.setdefault(kmer, []).append(j)
    hits, occs, ambiguous = {}, [], False
    for p in seeds:
        js = where.get(seq[p:p + SEED_K], [])
        if js:
            hits[p] = min(js) - p
            ambiguous |= len(js) > MAX_OCC
            occs += [j - p for j in js]
    anchor = int((H - R) / 2)
    if hits:
        anchor = sorted(hits.values())[(len(hits) - 1) // 2]
    cnt = np.zeros(32, np.int64)
    for o in occs:
        b = (o - anchor + 64) >> 2
        if 0 <= b < 32:
            cnt[b] += 1
    inv = np.float32(1) / np.float32(len(seeds)) if seeds else np.float32(0)
    x[4:] = cnt.astype(np.float32) * inv
    near = [p for p, o in hits.items() if abs(o - anchor) <= END_TOL]
    ends = (min(near), max(near) + SEED_K) if near else (-1, -1)
    return x, anchor, ends, ambiguous


@pytest.mark.gpu
def test_features_match_host_restatement():
    """Device seed features (hash-table lookups, wave-wide median, histogram
    and end reductions) against the host restatement, pair by pair."""
    b = synth.generate(1, count=3)
    plan = pairhmm.Plan(pairhmm.PackedBatch.from_synth(b))
    try:
        X, anc, ends = band.plan_features(plan, with_ends=True)
    finally:
        plan.close()
    o, checked, skipped = 0, 0, 0
    for g in range(b.n_regions):
        reads, haps = b.region(g)
        for rd in reads:
            for hp in haps:
                x, a, e, amb = host_features(rd, hp)
                if amb:
                    skipped += 1
                else:
                    assert anc[o] == a, (g, o)
                    assert tuple(ends[o]) == e, (g, o)
                    assert np.allclose(X[o], x, rtol=1e-6, atol=1e-7), (g, o, X[o] - x)
                    checked += 1
                o += 1
    say("features checked", checked, "skipped (seeds with > 4 occurrences)", skipped)
    assert o == len(X) and checked > 0.99 * o


@pytest.mark.gpu
def test_mlp_gradients_match_torch_reference():
    rng = np.random.default_rng(4)
    X = rng.random((300, 36)).astype(np.float32)
    T = rng.normal(size=(300, 2)).astype(np.float32)
    m = band.BandModel(seed=11)
    p = m.params()
    loss, grad = m.loss_grad(X, T)
    rl, rg = torch_loss_grad(p, X, T, band.LOSS_W)
    say("loss", loss, rl, "max grad diff", float(np.abs(grad - rg).max()), float(np.abs(rg).max()))
    assert loss == pytest.approx(rl, rel=1e-4)
    assert np.allclose(grad, rg, rtol=1e-3, atol=1e-5 * np.abs(rg).max())
    Y = m.predict(X)
    import torch
    with torch.no_grad():
        o = 0
        pt = torch.tensor(p)
        W1 = pt[o:o + 2304].view(64, 36); b1 = pt[2304:2368]
        W2 = pt[2368:6464].view(64, 64); b2 = pt[6464:6528]
        W3 = pt[6528:6656].view(2, 64); b3 = pt[6656:6658]
        ref = (torch.relu(torch.relu(torch.tensor(X) @ W1.T + b1) @ W2.T + b2) @ W3.T + b3).numpy()
    assert np.allclose(Y, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(5)
    X = rng.random((4096, 36)).astype(np.float32)
    T = np.stack([X[:, 4:8].sum(1) - X[:, 0], np.log2(1 + 8 * X[:, 2])], 1).astype(np.float32)
    m = band.BandModel(seed=3)
    losses = m.fit(X, T, epochs=30, batch=256, lr=3e-3)
    say("losses", losses[0], losses[-1])
    assert losses[-1] < 0.1 * losses[0]


@pytest.mark.gpu
def test_learned_band_against_heuristic_band(tmp_path):
    """Ablation: the trained network against the same banded kernel with the
    band centred on the seed anchor and no width prediction."""
    data = band.build_training_set(n_pairs=30_000)
    say("training set", data[0].shape, data[4])
    model, rep = band.train(epochs=25, data=data, out=tmp_path / "m.npz")
    say("train losses", rep.losses[0], rep.losses[-1])
    assert rep.losses[-1] < 0.5 * rep.losses[0]
    ev = synth.generate(1, seed=synth.SEED_BASE + 999, count=30)
    pb = pairhmm.PackedBatch.from_synth(ev)
    full, _, stf, _ = run_plan(pb)
    res = {}
    with pairhmm.Plan(pb) as plan:
        X, anchor, ends = band.plan_features(plan, with_ends=True)
    for name, kw in (("learned", dict(model=model)),
                     ("rules", dict(bands=band.rule_bands(X, anchor, ends))),
                     ("rules_narrow", dict(bands=band.rule_bands(X, anchor, ends, narrow=True))),
                     ("heuristic64", dict(bands=band.heuristic_bands(X, anchor, 64))),
                     ("heuristic32", dict(bands=band.heuristic_bands(X, anchor, 32)))):
        out, form, st, used = run_plan(pb, **kw)
        err = float(np.abs(out - full).max())
        narrow = int(((used[:, 1] - used[:, 0] + 1 == 32) & (form == 2)).sum())
        res[name] = (err, st.fallback_pairs, st.cells_swept, narrow)
        say(name, "max |dlog10|", err, "band", st.band_pairs, "narrow", narrow, "full pass",
            st.fallback_pairs, "cells", st.cells_swept, "of", stf.cells_swept)
    # Always-narrow bands are fast but can miss alignments that leave the
    # band unseen (a read end matching nowhere, an inexact repeat copy); the
    # learned band stays within tolerance.
    assert res["learned"][0] <= 1e-3
    for name, (err, fb, cells, nar) in res.items():
        assert cells < 0.5 * stf.cells_swept, name
    # What the network adds to the same seed rules: its width prediction puts
    # the pairs it is sure of on the narrow kernel (fewer cells than the rules
    # alone with wide bands), where the rules alone would send every pair
    # that passes their seed tests (rules_narrow, which is not safe at scale).
    assert res["learned"][3] > 0
    assert res["learned"][2] < res["rules"][2]
    assert res["learned"][3] < res["rules_narrow"][3]
    loaded = band.BandModel.load(tmp_path / "m.npz")
    assert np.array_equal(loaded.params(), model.params())
    # Adam's moments and step count are saved, so training can resume
    m1, m2, step = model.state()
    l1, l2, lstep = loaded.state()
    assert step == lstep > 0
    assert np.array_equal(m1, l1) and np.array_equal(m2, l2)
    # and the library-format file beside the .npz holds the same parameters
    assert np.array_equal(band.read_model_file(tmp_path / "m.bin"), model.params())


@pytest.mark.gpu
def test_shipped_bin_model_runs_like_npz_model():
    """The model file the C++ drop-in loads (models/band_mlp.bin, read and
    checked by the library) gives the same bands and bit-identical results
    as the trained .npz on a plan of short-read regions."""
    ev = synth.generate(1, seed=synth.SEED_BASE + 998, count=3)
    pb = pairhmm.PackedBatch.from_synth(ev)
    a_out, a_form, a_st, a_used = run_plan(pb, model=band.BandModel.load(band.DEFAULT_MODEL))
    b_out, b_form, b_st, b_used = run_plan(pb, model=band.BandModel.load_bin())
    say("pairs", pb.n_pairs, "band", a_st.band_pairs, "full pass", a_st.fallback_pairs)
    assert a_st.band_pairs > 0
    assert np.array_equal(a_used, b_used)
    assert np.array_equal(a_form, b_form)
    assert a_out.tobytes() == b_out.tobytes()


SLIDE = 8  # the wide band's start slide (diagonals)
HOLD = 1e-9  # a move that drops more of its row's mass is held


def slide_schedule(R, H, o_lo, width):
    """Per-row offsets the wide and mid bands' start slide asks for: up to
    SLIDE diagonals higher on rows 0-1, one lower per row down to o_lo;
    capped where the matrix ends.  The narrow band does not slide."""
    P = SLIDE if width >= 48 else 0
    pt = min(P, max(0, (H - 1) - (o_lo + width - 1)))
    off = [o_lo + max(0, pt + 1 - r) for r in range(1, R + 1)]
    return np.array([off[0]] + off, np.int32)


def oracle_band(oracle, rd, hap, o_lo, width):
    """The oracle's restatement of the banded kernels: the slide (held by the
    hold rule) and the insertion runs that leave through the bottom edge; the
    mid band's edge zones are three diagonals wide, the others' four."""
    R, H = len(rd["seq"]), len(hap)
    edge = 3 if width == 48 else 4
    v, ratio, _, _ = oracle.band_follow(rd, hap, o_lo, width, edge, max_move=1, hold_eps=HOLD, irun=2,
                                        sched=slide_schedule(R, H, o_lo, width))
    return v, ratio


def test_slide_schedule_shape():
    s = slide_schedule(100, 400, 50, 64)
    assert s[0] == s[1] == 58 and s[9] == 50 and s[100] == 50
    assert np.abs(np.diff(s)).max() == 1
    # a band already past the matrix's top does not slide
    assert (slide_schedule(20, 30, -22, 64) == -22).all()
    assert (slide_schedule(100, 400, 50, 32) == 50).all()
    s = slide_schedule(100, 400, 50, 48)
    assert s[0] == 58 and s[100] == 50


@pytest.mark.gpu
@pytest.mark.parametrize("width,shape", [(64, 4), (64, 1), (48, 4), (32, 4)])
def test_band_kernel_matches_oracle_band_forward(oracle, width, shape):
    """The banded kernels against the oracle's FP64 restatement of the same
    band (likelihood and the touch-mass decision), on bands placed around and
    across the alignment; the wide band with its start slide, in both shapes
    (four pairs of one read per wavefront, or one), and the mid band (48
    diagonals, three per lane, edge zones of three).  Each read runs against
    four haplotypes, so the quad kernels take whole groups."""
    sel, pb = read_groups(step=2)
    lab = band.labels(pb)
    rng = np.random.default_rng(width)
    centre = (lab[:, 0] + lab[:, 1]) // 2 + rng.integers(-width // 2, width // 2 + 1, len(sel))
    bands = np.stack([centre - width // 2, centre - width // 2 + width - 1], 1).astype(np.int32)
    out, form, st, _ = run_plan(pb, bands=bands, shape=shape, mid=width == 48)
    tol = pairhmm.default_options().band_tol
    n_acc = n_rej = 0
    worst = 0.0
    for k, (rd, hap) in enumerate(sel):
        v, ratio = oracle_band(oracle, rd, hap, int(bands[k, 0]), width)
        if form[k] == 2:
            n_acc += 1
            worst = max(worst, abs(out[k] - v))
            assert ratio < 10 * tol, k
        elif form[k] == 3:
            n_rej += 1
            assert ratio > 0.1 * tol, k
    say(f"width {width}, shape {shape}: accepted {n_acc}, rejected {n_rej}, "
        f"max |dlog10| vs oracle band {worst:.3g}")
    assert n_acc > 50 and n_rej > 10
    assert worst <= 2e-5


@pytest.mark.gpu
def test_mid_band_on_off():
    """Bands of 48 diagonals set by the caller, centred on the posterior span:
    with the mid band on they run as 48-diagonal bands (fewer cells swept);
    with it off (the default) or with the one-pair shape, as 64-diagonal
    bands with the same centre.  Every result within 1e-3 of the full
    matrix.  Learned bands never take 48 diagonals, mid band on or off."""
    sel, pb = read_groups(step=4)
    lab = band.labels(pb)
    centre = (lab[:, 0] + lab[:, 1]) // 2
    bands = np.stack([centre - 24, centre + 23], 1).astype(np.int32)
    full, _, _, _ = run_plan(pb)
    res = {}
    for name, shape, mid in (("mid", 4, True), ("no_mid", 4, None), ("one_pair", 1, True)):
        out, form, st, used = run_plan(pb, bands=bands, shape=shape, mid=mid)
        assert np.array_equal(used, bands), name
        err = float(np.abs(out - full).max())
        res[name] = (int(st.cells_swept), int((form == 2).sum()), err)
        say(name, "cells", st.cells_swept, "banded", int((form == 2).sum()), "of", len(sel),
            "max |dlog10|", err)
        assert err <= 1e-3, name
    assert res["mid"][1] > 0.9 * len(sel)
    assert res["mid"][0] < 0.9 * res["no_mid"][0] and res["mid"][0] < 0.9 * res["one_pair"][0]
    model = band.BandModel.load(band.DEFAULT_MODEL)
    for mid in (True, False):
        _, _, _, used = run_plan(pb, model=model, mid=mid)
        w = used[:, 1] - used[:, 0] + 1
        assert not (w == 48).any(), mid


@pytest.mark.gpu
@pytest.mark.parametrize("n_haps", [5, 6, 7])
def test_quad_fill_of_leftover_pairs(n_haps):
    """A read's last one to three wide pairs either fill one more four-pair
    wavefront (padding slots idle) or run one pair per wavefront: with
    64-diagonal caller bands, every fill setting (1-4) and the one-pair shape
    give the same forms on all but a few pairs at the accept threshold and
    the same likelihoods to FP32 rounding, all within 1e-3 of the full
    matrix."""
    sel, pb = read_groups(step=8, n_haps=n_haps)
    lab = band.labels(pb)
    centre = (lab[:, 0] + lab[:, 1]) // 2
    bands = np.stack([centre - 32, centre + 31], 1).astype(np.int32)
    full, _, _, _ = run_plan(pb)
    base = None
    for shape, fill in ((4, 4), (4, 3), (4, 2), (4, 1), (1, 4)):
        with pairhmm.Plan(pb) as plan:
            plan.set_band_shape(shape)
            plan.set_quad_fill(fill)
            assert plan.quad_fill() == fill
            band.set_plan_bands(plan, bands)
            plan.run()
            out, form = plan.results()
        err = float(np.abs(out - full).max())
        say(f"{n_haps} haplotypes, shape {shape}, fill {fill}: banded {int((form == 2).sum())} of {len(sel)}, "
            f"max |dlog10| vs full {err:.3g}")
        assert err <= 1e-3
        assert (form == 2).mean() > 0.9
        if base is None:
            base = (out, form)
            continue
        same = form == base[1]
        assert same.mean() >= 0.999
        both = same & (form == 2)
        assert float(np.abs(out[both] - base[0][both]).max()) <= 1e-4
    with pairhmm.Plan(pb) as plan:
        for bad in (0, 5):
            with pytest.raises(pairhmm.PairHMMError):
                plan.set_quad_fill(bad)


@pytest.mark.gpu
def test_fp64_band_rescue_of_underflowing_short_pairs(oracle):
    """Short-read pairs whose FP32 band sum falls below 1e-28 (beyond FP32's
    reach: the reference re-runs them in FP64) take the FP64 moving band
    (form 4) when their band passed the touch test: equal to the FP64
    full-matrix oracle, and counted as FP64 and banded pairs."""
    b = synth.generate(1, seed=synth.SEED_BASE + 555, count=300)
    pb = pairhmm.PackedBatch.from_synth(b)
    full, ff, _, _ = run_plan(pb)
    out, form, st, _ = run_plan(pb, model=band.BandModel.load(band.DEFAULT_MODEL))
    rr, hh = b.pair_table()
    resc = np.flatnonzero(form == 4)
    say("pairs", pb.n_pairs, "FP64 in the full leg", int((ff == 1).sum()), "on the FP64 band", resc.size)
    assert resc.size > 0
    assert (ff[resc] == 1).all() and (full[resc] < -64.1).all()
    assert st.fp64_pairs >= resc.size and st.band_pairs >= resc.size
    assert np.abs(out - full).max() <= 1e-3
    rng = np.random.default_rng(5)
    worst = 0.0
    for k in rng.choice(resc, size=min(40, resc.size), replace=False):
        r, h = int(rr[k]), int(hh[k])
        a, e = int(b.read_off[r]), int(b.read_off[r + 1])
        rd = dict(seq=b.read_bases[a:e].tobytes(), q=b.read_qual[a:e].tobytes(), i=b.read_ins[a:e].tobytes(),
                  d=b.read_del[a:e].tobytes(), c=b.read_gcp[a:e].tobytes())
        hap = b.hap_bases[int(b.hap_off[h]):int(b.hap_off[h + 1])].tobytes()
        worst = max(worst, abs(out[k] - oracle.log10_exact(rd, hap, 0)))
    say("max |dlog10| of the FP64 band rescue vs the FP64 oracle", worst)
    assert worst <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("learned", [False, True])
def test_graph_replays_match_eager_runs(learned):
    """bv_plan_run replays one captured graph; every replay, with eager runs
    of another plan and host synchronisations in between, equals the eager
    run of the same plan."""
    ev = synth.generate(1, seed=synth.SEED_BASE + 77, count=20)
    pb = pairhmm.PackedBatch.from_synth(ev)
    model = band.BandModel.load(band.DEFAULT_MODEL) if learned else None
    with pairhmm.Plan(pb, pairhmm.default_options(use_graph=0)) as eager, \
            pairhmm.Plan(pb, pairhmm.default_options(use_graph=1)) as graph:
        if model is not None:
            band.set_plan_model(eager, model)
            band.set_plan_model(graph, model)
        eager.run()
        ref, ref_form = eager.results()
        for k in range(4):
            graph.run()
            out, form = graph.results()
            assert np.array_equal(out, ref) and np.array_equal(form, ref_form), k
            eager.run()
            eager.results()
        if learned:
            # a new band setup rebuilds the graph
            band.set_plan_model(graph, model, margin=12)
            band.set_plan_model(eager, model, margin=12)
            eager.run()
            ref2, _ = eager.results()
            graph.run()
            out2, _ = graph.results()
            assert np.array_equal(out2, ref2)
    say("graph replays equal eager runs, learned =", learned)


@pytest.mark.gpu
def test_graph_with_long_reads_on_second_stream():
    """With long reads and a band model, the long-read passes run on a
    second stream joined back before the run ends (a fork and join in the
    captured graph).  Graph replays, with eager runs of another plan in
    between, equal the eager run; long pairs took the moving band."""
    ev = synth.generate(4, count=24)
    pb = pairhmm.PackedBatch.from_synth(ev)
    R = np.diff(pb.arrays["read_off"])
    assert (R >= 1024).any() and (R < 1024).any()
    model = band.BandModel.load(band.DEFAULT_MODEL)
    with pairhmm.Plan(pb, pairhmm.default_options(use_graph=0)) as eager, \
            pairhmm.Plan(pb, pairhmm.default_options(use_graph=1)) as graph:
        band.set_plan_model(eager, model)
        band.set_plan_model(graph, model)
        eager.run()
        ref, ref_form = eager.results()
        # long pairs on the moving band (FP32, or FP64); short pairs whose
        # FP32 band underflowed take the FP64 band too (form 4)
        rr, _ = ev.pair_table()
        long_pair = np.diff(ev.read_off)[rr] >= 1024
        on_band = np.isin(ref_form, (4, 5)) & long_pair
        say("long-read pairs on the moving band", int(on_band.sum()), "of", len(ref),
            "; short pairs on the FP64 band rescue", int(((ref_form == 4) & ~long_pair).sum()))
        assert on_band.sum() > 0
        for k in range(3):
            graph.run()
            out, form = graph.results()
            assert np.array_equal(out, ref) and np.array_equal(form, ref_form), k
            eager.run()
            eager.results()
        assert graph.stats().long_band_pairs == int(on_band.sum())


@pytest.mark.gpu
def test_autotune_changes_speed_not_results():
    """Grids change only the speed (bit-identical results); the wide band's
    two shapes give the same forms and agree to FP32 rounding (the D scan
    sums in another order), and so does whatever shape the autotuner picks.
    (With the mid band off: the one-pair shape has none, so the two shapes
    then run the same bands.)"""
    ev = synth.generate(1, seed=synth.SEED_BASE + 78, count=40)
    pb = pairhmm.PackedBatch.from_synth(ev)
    with pairhmm.Plan(pb) as plan:
        band.set_plan_model(plan, band.BandModel.load(band.DEFAULT_MODEL))
        plan.set_band_mid(False)
        assert not plan.band_mid()
        res = {}
        for shape in (4, 1):
            plan.set_band_shape(shape)
            assert plan.band_shape() == shape
            plan.run()
            res[shape] = plan.results()
            # another grid for the same shape (not a multiple of the XCD
            # count: one shared claim counter): bit-identical
            plan.set_grids(dict(band=257))
            plan.run()
            o2, f2 = plan.results()
            assert np.array_equal(o2, res[shape][0]) and np.array_equal(f2, res[shape][1])
        (o4, f4), (o1, f1) = res[4], res[1]
        assert np.array_equal(f4, f1)
        banded = f4 == 2
        assert banded.sum() > 1000
        d = np.abs(o4 - o1)
        assert np.array_equal(o4[~banded], o1[~banded])
        say(f"shapes 4 vs 1: {banded.sum()} banded pairs, max |dlog10| {d.max():.3g}")
        assert d.max() <= 1e-5
        grids = plan.autotune(1)
        plan.run()
        out, form = plan.results()
    say("autotuned", grids)
    assert all(v >= 1 for v in grids.values()) and grids["wide_pairs"] in (1, 4)
    ref = res[grids["wide_pairs"]]
    assert np.array_equal(form, ref[1]) and np.abs(out - ref[0]).max() <= 1e-5


def edge_regions(seed):
    """Regions at the fused prologue's edges: a read against 80 haplotypes
    (more than one 64-pair piece), reads shorter than a seed, a haplotype
    over 1024 bases (striped full-matrix list), a read over 448 bases (no
    band), and a read whose gap continuation varies (no band)."""
    rng = np.random.default_rng(seed)

    def hap(n):
        return rng.choice(list(b"ACGT"), n).astype(np.uint8)

    def mutate(h, rate=0.01):
        h = h.copy()
        k = rng.random(h.size) < rate
        h[k] = rng.choice(list(b"ACGT"), int(k.sum()))
        return h

    def read(h, R, c=10):
        o = int(rng.integers(0, max(1, h.size - R + 1)))
        s = mutate(h[o:o + R])
        return dict(seq=s.tobytes(), q=rng.integers(20, 41, s.size).astype(np.uint8).tobytes(),
                    i=bytes([45]) * s.size, d=bytes([45]) * s.size,
                    c=(bytes([c]) * s.size if np.isscalar(c) else bytes(c[:s.size])))

    base = hap(400)
    haps80 = [mutate(base).tobytes() for _ in range(80)]
    many = ([read(np.frombuffer(haps80[k % 80], np.uint8), int(rng.integers(60, 200))) for k in range(6)], haps80)
    b2 = hap(300)
    short = ([read(b2, 8), read(b2, 11), read(b2, 12), read(b2, 150)], [mutate(b2).tobytes() for _ in range(3)])
    b3 = hap(1100)
    longhap = ([read(b3, 150), read(b3, 200)], [b3.tobytes(), mutate(b3[:900]).tobytes()])
    b4 = hap(700)
    vary = rng.integers(5, 30, 460).astype(np.uint8)
    tall = ([read(b4, 460), read(b4, 449), read(b4, 300, c=vary)], [mutate(b4).tobytes() for _ in range(5)])
    return [many, short, longhap, tall]


@pytest.mark.gpu
def test_fused_prologue_equals_separate_kernels():
    """The fused prologue (features, MLP, band rules and list split in one
    kernel) and the split one (features as compact records, then one kernel
    on) give the band table, forms and results of the separate kernels bit
    for bit.  The exception: a task of more than 64 haplotypes
    is split into 64-pair pieces, which may group its wide pairs four at a
    time differently, and the quad and one-pair kernels agree to FP32
    rounding there."""
    b = synth.generate(1, seed=synth.workload_seed(1) + 977, count=150)
    regions = [b.region(g) for g in range(b.n_regions)] + edge_regions(5)
    pb = pairhmm.PackedBatch.from_regions(regions)
    model = band.BandModel.load(band.DEFAULT_MODEL)
    res = {}
    for mode in (2, 1, 0):  # split (records), fused, separate
        with pairhmm.Plan(pb) as plan:
            band.set_plan_model(plan, model)
            plan.set_band_prologue(mode)
            say("prologue mode", mode)
            plan.run()
            out, form = plan.results()
            st = plan.stats()
            res[mode] = (out, form, band.plan_bands(plan), int(st.band_pairs), int(st.fallback_pairs))
    o0, f0, b0, n0, m0 = res[0]
    # pairs of the 80-haplotype region (the first edge region)
    sizes = [len(r) * len(h) for r, h in regions]
    p0 = int(np.sum(sizes[:b.n_regions]))
    many = np.zeros(o0.size, bool)
    many[p0:p0 + sizes[b.n_regions]] = True
    for mode in (2, 1):
        o1, f1, b1, n1, m1 = res[mode]
        assert np.array_equal(b1, b0), mode
        assert np.array_equal(f1, f0), mode
        assert (n1, m1) == (n0, m0), mode
        assert np.array_equal(o1[~many], o0[~many]), mode
        assert np.abs(o1[many] - o0[many]).max() <= 1e-6, mode
    # the split and fused prologues group the long task's pieces alike
    assert np.array_equal(res[2][0], res[1][0])
    # the edge cases took the lists they should
    assert (f0 == 2).sum() > 0.9 * b.n_pairs
    say("pairs", o0.size, "banded", int((f0 == 2).sum()), "full", int(np.isin(f0, (0, 1, 3)).sum()))
