"""Post-processing of the per-pair results: normalize_likelihoods and
filter_poorly_modeled_reads (pairhmm.hpp:83-114) and the genotyper's
marginal_likelihoods (genotyper.hpp:200-241).

The oracle's restatement (bvo_postprocess_region) is pinned by hand-worked
rows; the device kernel (bv_plan_postprocess) must equal it bit for bit on
the plan's own results."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT))

LOWEST = -np.finfo(np.float64).max


def post_oracle(oracle, L, read_len, normalize=True, hap_allele=None, n_alleles=0):
    """ctypes call of the oracle's restatement for one region."""
    lib = oracle.lib
    f = lib.bvo_postprocess_region
    f.restype = None
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                  C.c_void_p, C.c_void_p, C.c_void_p]
    L = np.ascontiguousarray(L, np.float64)
    n, m = L.shape
    rl = np.ascontiguousarray(read_len, np.int64)
    keep = np.zeros(n, np.uint8)
    Ln = np.empty_like(L)
    A = None
    ha = None
    if hap_allele is not None:
        ha = np.ascontiguousarray(hap_allele, np.int32)
        A = np.empty((n, n_alleles), np.float64)
    f(n, m, L.ctypes.data, rl.ctypes.data, int(normalize), None if ha is None else ha.ctypes.data,
      int(n_alleles), keep.ctypes.data, Ln.ctypes.data, None if A is None else A.ctypes.data)
    return keep, Ln, A


def test_post_oracle_hand_worked_rows(oracle):
    """Rows worked by hand.  Cap = best - 4.5; the filter keeps a read when
    its best is below min(2, ceil(0.02 len)) * -4: -4 for reads of 1-50
    bases, -8 from 51 bases (ceil(1.02) = 2); a best equal to the threshold
    is dropped (strict test)."""
    L = np.array([
        [-1.0, -7.0, -3.0],      # cap -5.5: [-1, -5.5, -3]; best -1 >= -8: dropped
        [-10.0, -20.0, -14.5],   # cap -14.5: [-10, -14.5, -14.5]; -10 < -8: kept
        [-8.0, -9.0, -30.0],     # best -8 is not below -8: dropped
        [-4.0, -4.0, -100.0],    # 50 bases: threshold -4, -4 not below: dropped
        [-4.5, -50.0, -5.0],     # 50 bases: -4.5 < -4: kept; cap -9: [-4.5, -9, -5]
        [-np.inf, -np.inf, -np.inf],  # below the FP64 floor: cap -inf, kept
    ])
    read_len = [150, 150, 100, 50, 50, 300]
    keep, Ln, A = post_oracle(oracle, L, read_len, True, [0, 1, 0], 3)
    assert keep.tolist() == [0, 1, 0, 0, 1, 1]
    np.testing.assert_array_equal(Ln[0], [-1.0, -5.5, -3.0])
    np.testing.assert_array_equal(Ln[1], [-10.0, -14.5, -14.5])
    np.testing.assert_array_equal(Ln[2], [-8.0, -9.0, -12.5])
    np.testing.assert_array_equal(Ln[4], [-4.5, -9.0, -5.0])
    assert np.isneginf(Ln[5]).all()
    # alleles: 0 <- haplotypes 0 and 2, 1 <- haplotype 1, 2 <- none (lowest)
    np.testing.assert_array_equal(A[:, 0], [-1.0, -10.0, -8.0, -4.0, -4.5, LOWEST])
    np.testing.assert_array_equal(A[:, 1], [-5.5, -14.5, -9.0, -4.0, -9.0, LOWEST])
    assert (A[:, 2] == LOWEST).all()
    # without normalisation the marginal takes the raw entries
    keep2, Ln2, A2 = post_oracle(oracle, L, read_len, False, [0, 1, 0], 2)
    np.testing.assert_array_equal(Ln2, L)
    np.testing.assert_array_equal(A2[:, 1], [-7.0, -20.0, -9.0, -4.0, -50.0, LOWEST])
    assert keep2.tolist() == keep.tolist()


def test_post_oracle_filter_threshold_by_length(oracle):
    """min(2, ceil(0.02 len)) * -4 for every length 1..400: -4 up to 50
    bases, -8 from 51 (0.02 is not exact in binary; 50 * 0.02 rounds to 1.0)."""
    lens = np.arange(1, 401)
    for best, want in ((-3.99, lambda n: False), (-4.01, lambda n: n <= 50),
                       (-7.99, lambda n: n <= 50), (-8.01, lambda n: True)):
        L = np.full((len(lens), 2), best)
        keep, _, _ = post_oracle(oracle, L, lens)
        assert keep.tolist() == [int(want(n)) for n in lens], best


def test_post_oracle_matches_python_mirror(oracle):
    """The oracle against the host mirror of the reference class
    (PairHMM.normalize_likelihoods / filter_poorly_modeled_reads) on random
    rows with ties and infinities."""
    from biovoltron_amd.pairhmm import PairHMM

    rng = np.random.default_rng(5)
    for _ in range(50):
        n, m = int(rng.integers(1, 30)), int(rng.integers(1, 12))
        L = -np.round(rng.exponential(6.0, (n, m)), 1)
        L[rng.random((n, m)) < 0.05] = -np.inf
        lens = rng.integers(1, 300, n)
        keep, Ln, _ = post_oracle(oracle, L, lens)
        rows = [list(r) for r in L]
        PairHMM.normalize_likelihoods(rows)

        class Rd:
            def __init__(self, k):
                self.seq = "A" * int(k)

        reads = [Rd(k) for k in lens]
        kept = [list(r) for r in rows]
        PairHMM.filter_poorly_modeled_reads(reads, kept)
        np.testing.assert_array_equal(Ln, np.array(rows))
        assert [r for r, k in zip(Ln.tolist(), keep) if k] == kept


def _plan_with_results(workload, count, seed_off=0):
    from biovoltron_amd import pairhmm, synth

    b = synth.generate(workload, count=count, first=seed_off)
    plan = pairhmm.Plan(pairhmm.PackedBatch.from_synth(b))
    plan.run()
    out, _ = plan.results()
    return b, plan, out


@pytest.mark.gpu
@pytest.mark.parametrize("workload,count", [(1, 40), (2, 2)])
def test_device_postprocess_equals_oracle(oracle, workload, count):
    """bv_plan_postprocess against the oracle's restatement, region by
    region, bit for bit: keep flags, normalised rows, allele maxima for a
    random haplotype -> allele map (some alleles without haplotypes)."""
    b, plan, out = _plan_with_results(workload, count)
    rng = np.random.default_rng(11 + workload)
    G = b.n_regions
    nh = np.diff(b.region_hap_off)
    n_alleles = np.array([int(rng.integers(1, k + 2)) for k in nh], np.int32)
    hap_allele = np.concatenate([rng.integers(0, n_alleles[g], nh[g]) for g in range(G)]).astype(np.int32)
    print(f"post: workload {workload}, {G} regions, {b.n_pairs} pairs", flush=True)
    for normalize in (True, False):
        keep, Ln, A = plan.postprocess(hap_allele, n_alleles, normalize=normalize, normalized=True)
        a0 = 0
        for g in range(G):
            r0, r1 = int(b.region_read_off[g]), int(b.region_read_off[g + 1])
            h0, h1 = int(b.region_hap_off[g]), int(b.region_hap_off[g + 1])
            p0 = int(b.region_pair_off[g])
            n, m = r1 - r0, h1 - h0
            Lg = out[p0:p0 + n * m].reshape(n, m)
            k_o, Ln_o, A_o = post_oracle(oracle, Lg, np.diff(b.read_off)[r0:r1], normalize,
                                         hap_allele[h0:h1], int(n_alleles[g]))
            np.testing.assert_array_equal(keep[r0:r1], k_o)
            np.testing.assert_array_equal(Ln[p0:p0 + n * m], Ln_o.ravel())
            na = int(n_alleles[g])
            np.testing.assert_array_equal(A[a0:a0 + n * na], A_o.ravel())
            a0 += n * na
        assert a0 == len(A)
    assert plan.stats().last_ms_post > 0.0
    # keep alone, no allele map
    k2, n2, A2 = plan.postprocess()
    assert n2 is None and A2 is None and np.array_equal(k2, keep)
    plan.close()


@pytest.mark.gpu
def test_device_postprocess_rejects_bad_maps():
    from biovoltron_amd import pairhmm

    b, plan, _ = _plan_with_results(1, 3)
    nh = np.diff(b.region_hap_off)
    na = np.full(b.n_regions, 2, np.int32)
    ha = np.zeros(int(nh.sum()), np.int32)
    ha[-1] = 2  # allele out of range for its region
    with pytest.raises(pairhmm.PairHMMError, match="maps to allele"):
        plan.postprocess(ha, na)
    with pytest.raises(pairhmm.PairHMMError, match="given together"):
        pairhmm._check(plan._lib.bv_plan_postprocess(plan._h, 1, ha.ctypes.data, None, None, None, None))
    with pytest.raises(pairhmm.PairHMMError, match="flags"):
        pairhmm._check(plan._lib.bv_plan_postprocess(plan._h, 6, None, None, None, None, None))
    plan.close()


@pytest.mark.gpu
def test_device_postprocess_empty_regions(oracle):
    """Regions without reads (with or without haplotypes) between regions
    with reads: the binary search over region offsets must skip them, and
    the allele blocks of read-less regions are empty."""
    from biovoltron_amd import pairhmm

    rng = np.random.default_rng(3)

    def rd(n, q=30):
        s = bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8))
        return dict(seq=s, q=bytes([q]) * n, i=bytes([45]) * n, d=bytes([45]) * n, c=bytes([10]) * n)

    def hap(n):
        return bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8))

    regions = [([], [hap(40)]), ([rd(30), rd(35)], [hap(60), hap(61), hap(59)]), ([], []),
               ([rd(20)], [hap(45)]), ([], [hap(50), hap(52)]), ([rd(25), rd(26), rd(60)], [hap(80), hap(81)])]
    pb = pairhmm.PackedBatch.from_regions(regions)
    plan = pairhmm.Plan(pb)
    plan.run()
    out, _ = plan.results()
    nh = np.diff(pb.arrays["region_hap_off"])
    na = np.array([max(1, k) for k in nh], np.int32)
    ha = np.concatenate([np.arange(k) % max(1, k) for k in nh]).astype(np.int32)
    keep, Ln, A = plan.postprocess(ha, na, normalized=True)
    rro = pb.arrays["region_read_off"]
    read_len = np.diff(pb.arrays["read_off"])
    p0 = a0 = 0
    for g, (reads, haps) in enumerate(regions):
        n, m = len(reads), len(haps)
        if n == 0:
            continue
        r0 = int(rro[g])
        k_o, Ln_o, A_o = post_oracle(oracle, out[p0:p0 + n * m].reshape(n, m), read_len[r0:r0 + n], True,
                                     ha[int(pb.arrays["region_hap_off"][g]):int(pb.arrays["region_hap_off"][g + 1])],
                                     int(na[g]))
        np.testing.assert_array_equal(keep[r0:r0 + n], k_o)
        np.testing.assert_array_equal(Ln[p0:p0 + n * m], Ln_o.ravel())
        np.testing.assert_array_equal(A[a0:a0 + n * int(na[g])], A_o.ravel())
        p0 += n * m
        a0 += n * int(na[g])
    assert p0 == len(out) and a0 == len(A)
    plan.close()


@pytest.mark.gpu
def test_device_postprocess_reads_without_pairs():
    """Reads that lie outside every region leave the plan without pairs; the
    keep flags must still be written (0 for every read), not left as the
    caller's buffer held them."""
    from biovoltron_amd import pairhmm

    n = 12
    seq = np.frombuffer(b"ACGTACGTAC" * 3, np.uint8)
    pb = pairhmm.PackedBatch([0, 0], [0, 1], [0, 10, 20, 30], seq, np.full(30, 30, np.uint8),
                             np.full(30, 45, np.uint8), np.full(30, 45, np.uint8), np.full(30, 10, np.uint8),
                             [0, n], np.frombuffer(b"ACGTACGTACGT", np.uint8))
    assert pb.n_pairs == 0 and pb.n_reads == 3
    plan = pairhmm.Plan(pb)
    plan.run()
    keep = np.full(pb.n_reads, 7, np.uint8)
    pairhmm._check(plan._lib.bv_plan_postprocess(plan._h, 1, None, None, keep.ctypes.data, None, None))
    np.testing.assert_array_equal(keep, np.zeros(3, np.uint8))
    plan.close()
