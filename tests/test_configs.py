"""The learned-band and combined workloads at scale, against the oracle.

* Learned band on 10,000 fresh short-read regions (the generator of the
  batched workload with a seed disjoint from the band model's training pairs
  and from the bench's batch), with the committed model.
* The combined workload: 90 % short-read and 10 % long-read regions in one
  plan (learned band with full-matrix fallback, moving FP64 band for long
  reads, FP64 rescue).

Each run is checked over every pair against the same plan's full-matrix run
(|dlog10 L| <= 1e-3), and on a seeded sample of every result form against
the FP64 full-matrix oracle (bvo_log10_exact: renormalised per row, never
underflows) and the reference driver rule (bvo_log10_reference: FP32, FP64
re-run below 1e-28, log10; compared where the reference is finite).  The
reference's rule is pairhmm_avx.hpp:121-157.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import ORACLE_SO
from biovoltron_amd import band, pairhmm, synth

TOL = 1e-3
LONG = 1024  # reads this long run in FP64 first


def say(*a):
    print("[configs]", *a, flush=True)


class _PairOracle:
    """The oracle's per-pair functions on pointers into a batch's arrays."""

    def __init__(self, b):
        lib = C.CDLL(str(ORACLE_SO))
        lib.bvo_init()
        self.exact = lib.bvo_log10_exact
        self.exact.restype = C.c_double
        self.exact.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_int]
        self.ref = lib.bvo_log10_reference
        self.ref.restype = C.c_double
        self.ref.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.c_void_p]
        self.b = b

    def _args(self, r, h):
        b = self.b
        ro, ho = int(b.read_off[r]), int(b.hap_off[h])
        at = lambda a, o: a.ctypes.data + o  # noqa: E731
        return (int(b.read_off[r + 1]) - ro, int(b.hap_off[h + 1]) - ho, at(b.read_bases, ro),
                at(b.read_qual, ro), at(b.read_ins, ro), at(b.read_del, ro), at(b.read_gcp, ro),
                at(b.hap_bases, ho))

    def pair(self, rh):
        r, h = rh
        a = self._args(r, h)
        return self.exact(*a, 0), self.ref(*a, 1, None)


def check_forms(b, legs, n_short=300, n_long=4, seed=0):
    """Per leg and form: max |dlog10 L| against the FP64 oracle and the
    reference rule on a seeded sample; asserts both within TOL."""
    rr, hh = b.pair_table()
    rlen = np.diff(b.read_off)[rr]
    orc = _PairOracle(b)
    rng = np.random.default_rng(seed)
    seen = {}
    with ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0))))) as pool:
        for leg, (out, form) in legs.items():
            for code in np.unique(form):
                for long_reads in (False, True):
                    sel = np.flatnonzero((form == code) & ((rlen >= LONG) == long_reads))
                    if sel.size == 0:
                        continue
                    pick = np.sort(rng.choice(sel, size=min(sel.size, n_long if long_reads else n_short),
                                              replace=False))
                    res = np.array(list(pool.map(orc.pair, [(int(rr[p]), int(hh[p])) for p in pick])))
                    ex, ref = res[:, 0], res[:, 1]
                    d_ex = float(np.abs(out[pick] - ex).max())
                    fin = np.isfinite(ref)
                    d_ref = float(np.abs(out[pick][fin] - ref[fin]).max()) if fin.any() else None
                    key = f"{leg}/{pairhmm.FORM_NAMES[int(code)]}/{'long' if long_reads else 'short'}"
                    seen[key] = (int(sel.size), int(pick.size), d_ex, d_ref)
                    say(key, "pairs", sel.size, "sampled", pick.size, "max |d| vs FP64 oracle", f"{d_ex:.2e}",
                        "vs reference rule", "n/a (below its FP64 floor)" if d_ref is None else f"{d_ref:.2e}")
                    assert d_ex <= TOL, key
                    assert d_ref is None or d_ref <= TOL, key
    return seen


def run_legs(b, margin=10, max_width=68):
    """Full-matrix run, then the learned band on the same plan."""
    with pairhmm.Plan(pairhmm.PackedBatch.from_synth(b)) as plan:
        plan.run()
        full, ff = plan.results()
        band.set_plan_model(plan, band.BandModel.load(band.DEFAULT_MODEL), margin=margin, max_width=max_width)
        plan.run()
        lb, fb = plan.results()
        st = plan.stats()
    return (full, ff), (lb, fb), st


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_learned_band_on_fresh_short_read_regions():
    # fresh regions: seed 20261206, not the training set's (SEED_BASE + 301)
    # nor the bench's batch (SEED_BASE + 1)
    seed = synth.SEED_BASE + 3
    b = synth.generate(1, seed=seed, count=10_000)
    say("learned band: 10,000 regions, seed", seed, "pairs", b.n_pairs, f"cells {b.cells:.3e}")
    (full, ff), (lb, fb), st = run_legs(b)
    assert np.isfinite(full).all() and np.isfinite(lb).all()
    d = np.abs(lb - full)
    say("max |dlog10 L| learned band vs full matrix over all pairs", f"{d.max():.2e}",
        "banded", int((fb == 2).sum()), "rejected", int((fb == 3).sum()), "full list", int((fb == 0).sum()),
        "fp64", int((fb == 1).sum()))
    assert d.max() <= TOL
    # the band takes most pairs and the fallback is exercised
    assert (fb == 2).sum() > 0.9 * b.n_pairs
    assert (fb == 3).sum() > 0 and st.fallback_pairs > 0
    check_forms(b, {"full": (full, ff), "learned": (lb, fb)}, seed=3)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_combined_workload_against_full_and_oracle():
    regions = int(os.environ.get("BV_COMBINED_REGIONS", "100000"))
    b = synth.generate(4, count=regions)
    R = np.diff(b.read_off)
    say("combined workload:", b.n_regions, "regions, pairs", b.n_pairs, f"cells {b.cells:.3e}",
        "long reads", int((R >= LONG).sum()))
    (full, ff), (lb, fb), st = run_legs(b)
    assert np.isfinite(full).all() and np.isfinite(lb).all()
    d = np.abs(lb - full)
    rr, _ = b.pair_table()
    long_pair = R[rr] >= LONG
    say("max |dlog10 L| vs full matrix: short", f"{d[~long_pair].max():.2e}", "long", f"{d[long_pair].max():.2e}",
        "forms", {pairhmm.FORM_NAMES[int(k)]: int((fb == k).sum()) for k in np.unique(fb)})
    assert d.max() <= TOL
    # every path of the run is exercised: banded, rejected, FP64 rescue,
    # long reads on the moving band and in the full-matrix FP64 pass
    for code in (1, 2, 3, 5):  # 5: the long reads' moving band (FP32, the default)
        assert (fb == code).sum() > 0, pairhmm.FORM_NAMES[code]
    assert (ff == 1).sum() > 0 and ((ff == 1) & long_pair).sum() > 0
    assert ((fb == 1) & long_pair).sum() > 0  # long pairs the moving band left
    check_forms(b, {"full": (full, ff), "learned": (lb, fb)}, seed=4)
