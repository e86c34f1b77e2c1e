"""The oracle's scalar model (bvo_scalar_log10) and post-processing
(bvo_postprocess_region) against the reference's own PairHMM
(include/biovoltron/algo/align/inexact_match/pairhmm.hpp:39-141), compiled
from its header into oracle/_ref/libref_scalar.so by `make -C oracle ref`
where the reference tree is present (CPU only).

* Raw likelihoods: regions of one haplotype (normalisation is then the
  identity) whose reads match it poorly, so the reference's inverted filter
  keeps them and its output shows the raw log10 values; random transition
  matrices (seven distinct entries, so a swapped index shows).
* Whole regions: several haplotypes and reads sampled from them with errors:
  the reference's compute_likelihoods against the oracle's per-pair scalar
  values followed by its normalise and filter (kept reads and every value).
* Post-processing alone: random matrices with ties and values placed on the
  normalisation cap and the filter thresholds, bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import ROOT
from test_post import post_oracle

REF_SCALAR = ROOT / "oracle" / "_ref" / "libref_scalar.so"
pytestmark = pytest.mark.skipif(not REF_SCALAR.exists(),
                                reason="oracle/_ref/libref_scalar.so is built only where the reference tree is")

ORIGINAL_DEFAULT = [0.9998, 0.0001, 0.0001, 0.9, 0.1, 0.9, 0.1]  # pairhmm.hpp:32-33


def ref_lib():
    lib = C.CDLL(str(REF_SCALAR))
    vp = C.c_void_p
    lib.ref_scalar_region.restype = C.c_int
    lib.ref_scalar_region.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp]
    lib.ref_postprocess.restype = C.c_int
    lib.ref_postprocess.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    return lib


def ref_region(lib, reads, haps, trans):
    ro = np.concatenate([[0], np.cumsum([len(r["seq"]) for r in reads])]).astype(np.int64)
    ho = np.concatenate([[0], np.cumsum([len(h) for h in haps])]).astype(np.int64)
    seq = np.frombuffer(b"".join(r["seq"] for r in reads), np.uint8).copy()
    q = np.frombuffer(b"".join(r["q"] for r in reads), np.uint8).copy()
    hb = np.frombuffer(b"".join(haps), np.uint8).copy()
    tr = np.ascontiguousarray(trans, np.float64)
    out = np.zeros((len(reads), len(haps)), np.float64)
    kept = np.zeros(len(reads), np.uint8)
    n = lib.ref_scalar_region(len(reads), ro.ctypes.data, seq.ctypes.data, q.ctypes.data, len(haps),
                              ho.ctypes.data, hb.ctypes.data, tr.ctypes.data, out.ctypes.data, kept.ctypes.data)
    return out[:n], kept.astype(bool)


def rand_read(rng, n, src=None, err=0.0, qlo=0, qhi=41):
    if src is not None and len(src) >= n:
        off = int(rng.integers(0, len(src) - n + 1))
        s = bytearray(src[off:off + n])
        for k in range(n):
            if rng.random() < err:
                s[k] = int(rng.choice(list(b"ACGTN")))
    else:
        s = bytearray(rng.choice(list(b"ACGTN"), n, p=[.24, .24, .24, .24, .04]).astype(np.uint8))
    q = rng.integers(qlo, qhi, n).astype(np.uint8)
    return dict(seq=bytes(s), q=q.tobytes(), i=bytes([45]) * n, d=bytes([45]) * n, c=bytes([10]) * n)


def rand_trans(rng):
    t = rng.uniform(0.01, 0.99, 7)
    return list(t)


@pytest.mark.parametrize("seed", range(4))
def test_scalar_raw_likelihoods_match_reference(oracle, seed):
    lib = ref_lib()
    rng = np.random.default_rng(seed)
    compared = 0
    for case in range(30):
        H = int(rng.integers(1, 80))
        hap = bytes(rng.choice(list(b"ACGTN"), H, p=[.24, .24, .24, .24, .04]).astype(np.uint8))
        reads = [rand_read(rng, int(rng.integers(1, 120))) for _ in range(int(rng.integers(1, 8)))]
        trans = ORIGINAL_DEFAULT if case % 3 == 0 else rand_trans(rng)
        out, kept = ref_region(lib, reads, [hap], trans)
        rows = iter(out)
        for rd, k in zip(reads, kept):
            if not k:
                continue
            want = next(rows)[0]
            got = oracle.scalar_log10(rd, hap, trans)
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (seed, case, len(rd["seq"]), H, got, want)
            compared += 1
    assert compared >= 40


@pytest.mark.parametrize("seed", range(3))
def test_scalar_regions_match_reference(oracle, seed):
    lib = ref_lib()
    rng = np.random.default_rng(100 + seed)
    kept_total = dropped_total = 0
    for case in range(6):
        base = bytes(rng.choice(list(b"ACGT"), int(rng.integers(40, 160))).astype(np.uint8))
        haps = [base]
        for _ in range(int(rng.integers(1, 5))):
            h = bytearray(base)
            for k in rng.integers(0, len(h), 3):
                h[k] = int(rng.choice(list(b"ACGT")))
            haps.append(bytes(h))
        reads = [rand_read(rng, int(rng.integers(10, 120)), src=haps[int(rng.integers(0, len(haps)))],
                           err=float(rng.choice([0.0, 0.02, 0.3])), qlo=5) for _ in range(int(rng.integers(2, 12)))]
        trans = ORIGINAL_DEFAULT if case % 2 == 0 else rand_trans(rng)
        want, kept_ref = ref_region(lib, reads, haps, trans)
        raw = np.array([[oracle.scalar_log10(rd, h, trans) for h in haps] for rd in reads])
        keep, Ln, _ = post_oracle(oracle, raw, [len(rd["seq"]) for rd in reads])
        assert np.array_equal(keep.astype(bool), kept_ref), (seed, case)
        got = Ln[keep.astype(bool)]
        assert got.shape == want.shape
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12), (seed, case, np.abs(got - want).max())
        kept_total += int(keep.sum())
        dropped_total += int((~keep.astype(bool)).sum())
    # both sides of the filter were exercised
    assert kept_total > 0 and dropped_total > 0


def test_postprocess_matches_reference_bit_for_bit(oracle):
    lib = ref_lib()
    rng = np.random.default_rng(7)
    for case in range(200):
        n, m = int(rng.integers(1, 9)), int(rng.integers(1, 7))
        L = rng.uniform(-30, 0, (n, m))
        lens = rng.integers(1, 400, n).astype(np.int64)
        for r in range(n):
            kind = rng.integers(0, 5)
            if kind == 0:  # best exactly on the filter threshold
                thr = min(2.0, np.ceil(lens[r] * 0.02)) * -4.0
                L[r, int(rng.integers(0, m))] = thr
                L[r] = np.minimum(L[r], thr)
            elif kind == 1 and m > 1:  # a value exactly on the cap
                L[r, 0] = L[r].max() - 4.5
            elif kind == 2:  # ties for the best
                L[r, :] = np.round(L[r], 0)
        out = np.zeros((n, m), np.float64)
        kept = np.zeros(n, np.uint8)
        k = lib.ref_postprocess(n, m, np.ascontiguousarray(L).ctypes.data, lens.ctypes.data, out.ctypes.data,
                                kept.ctypes.data)
        keep, Ln, _ = post_oracle(oracle, L, lens)
        assert np.array_equal(keep, kept), case
        assert Ln[keep.astype(bool)].tobytes() == out[:k].tobytes(), case
