"""bv_compute_full_prob_noalloc against bv_compute_full_prob (same kernels,
bit-identical scaled sums), through the C ABI with device buffers the test
allocates itself; two calls in a row, the second smaller, on the shared
static scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DevicePairs(C.Structure):
    _fields_ = [("n_pairs", C.c_int64), ("read_off", C.c_void_p), ("read_bases", C.c_void_p),
                ("read_qual", C.c_void_p), ("read_ins", C.c_void_p), ("read_del", C.c_void_p),
                ("read_gcp", C.c_void_p), ("hap_off", C.c_void_p), ("hap_bases", C.c_void_p)]


class _Hip:
    """The HIP runtime the library links (same soname, so the same loaded copy)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(max(1, arr.nbytes))) == 0
        self.bufs.append(p)
        if arr.nbytes:
            assert self.lib.hipMemcpy(p, C.c_void_p(arr.ctypes.data), C.c_size_t(arr.nbytes), 1) == 0
        return p.value

    def download(self, ptr: int, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(8 * n), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _testcases(seed, n_pairs):
    """Pairs of the 64 x 8 single-region generator, shuffled, plus short
    edge pairs (one base on either side, N on both sides)."""
    from biovoltron_amd import synth

    b = synth.generate(0, seed=seed)
    reads, haps = b.region(0)
    rng = np.random.default_rng(seed)
    pairs = [(r, h) for r in reads for h in haps]
    pick = rng.permutation(len(pairs))[:n_pairs]
    tcs = [dict(rs=pairs[k][0]["seq"], q=pairs[k][0]["q"], i=pairs[k][0]["i"], d=pairs[k][0]["d"],
                c=pairs[k][0]["c"], hap=pairs[k][1]) for k in pick]
    tcs.append(dict(rs=b"A", q=b"\x1e", i=b"\x2d", d=b"\x2d", c=b"\x0a", hap=b"A"))
    tcs.append(dict(rs=b"ANGT", q=b"\x1e\x14\x28\x0a", i=b"\x2d" * 4, d=b"\x2d" * 4, c=b"\x0a" * 4,
                    hap=b"ACNTTG"))
    return tcs


def _device_pairs(hip, tcs):
    ro = np.cumsum([0] + [len(t["rs"]) for t in tcs]).astype(np.int64)
    ho = np.cumsum([0] + [len(t["hap"]) for t in tcs]).astype(np.int64)
    cat = lambda k: np.frombuffer(b"".join(t[k] for t in tcs), np.uint8)  # noqa: E731
    return _DevicePairs(len(tcs), hip.upload(ro), hip.upload(cat("rs")), hip.upload(cat("q")),
                        hip.upload(cat("i")), hip.upload(cat("d")), hip.upload(cat("c")),
                        hip.upload(ho), hip.upload(cat("hap")))


def _noalloc(lib, hip, tcs, precision, max_hap_len):
    dp = _device_pairs(hip, tcs)
    out = hip.upload(np.full(len(tcs), -7.0))
    rc = lib.bv_compute_full_prob_noalloc(C.byref(dp), precision, max_hap_len, C.c_void_p(out), None)
    if rc != 0:
        raise RuntimeError(f"status {rc}: {lib.bv_pairhmm_last_error().decode()}")
    return hip.download(out, len(tcs))


def test_noalloc_matches_full_prob_on_two_calls():
    from biovoltron_amd import pairhmm

    lib = pairhmm.library()
    lib.bv_compute_full_prob_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    assert lib.bv_pairhmm_init(-1) == 0
    hip = _Hip()
    try:
        first = _testcases(11, 512)
        second = _testcases(12, 61)  # smaller, other pairs: the scratch holds the first call's rows
        for precision in (32, 64):
            for tcs in (first, second):
                hmax = max(len(t["hap"]) for t in tcs)
                got = _noalloc(lib, hip, tcs, precision, hmax)
                want = pairhmm.compute_full_prob(tcs, precision)
                print(f"noalloc precision {precision}: {len(tcs)} pairs, max H {hmax}", flush=True)
                np.testing.assert_array_equal(got, want)
                assert np.all(np.isfinite(got)) and np.all(got > 0)
        hip.free()
    finally:
        hip.free()


def test_noalloc_limits():
    from biovoltron_amd import pairhmm

    lib = pairhmm.library()
    lib.bv_compute_full_prob_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    assert lib.bv_pairhmm_init(-1) == 0
    hip = _Hip()
    try:
        tcs = _testcases(13, 20)
        # a haplotype longer than max_hap_len: NaN for that pair, the others unchanged
        hmax = max(len(t["hap"]) for t in tcs)
        long_k = int(np.argmax([len(t["hap"]) for t in tcs]))
        got = _noalloc(lib, hip, tcs, 32, hmax - 1)
        want = pairhmm.compute_full_prob(tcs, 32)
        assert np.isnan(got[long_k])
        ok = np.array([len(t["hap"]) <= hmax - 1 for t in tcs])
        np.testing.assert_array_equal(got[ok], want[ok])
        # one carry row beyond the scratch (FP64: 16 bytes per column)
        too_long = (32 << 20) // 16
        with pytest.raises(RuntimeError, match="scratch"):
            _noalloc(lib, hip, tcs, 64, too_long)
        with pytest.raises(RuntimeError, match="precision"):
            _noalloc(lib, hip, tcs, 16, hmax)
    finally:
        hip.free()
