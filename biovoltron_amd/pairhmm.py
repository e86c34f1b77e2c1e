"""Python host side of the MI355X PairHMM, mirroring the reference API.

Reference interfaces mirrored (paths relative to the reference tree):

* ``AvxPairHMM.compute_likelihoods(haplotypes, reads)`` --
  include/biovoltron/algo/align/inexact_match/pairhmm_avx.hpp:160-171
  (FP32 first, FP64 re-run below 1e-28, log10, then normalise and filter).
* ``PairHMM.compute_likelihoods(haplotypes, reads, trans)`` --
  include/biovoltron/algo/align/inexact_match/pairhmm.hpp:116-141 (constant
  transition matrix, FP64).
* ``PairHMM.normalize_likelihoods`` / ``filter_poorly_modeled_reads`` --
  pairhmm.hpp:83-114, reproduced as the reference behaves: a read is *kept*
  when its best likelihood is below the threshold.
* ``compute_full_prob(testcases, precision)`` -- the per-pair kernels
  compute_full_prob_avx{s,d} (lib/intel_native/pairhmm/avx-pairhmm-template.h:210).

All compute goes through libbvpairhmm.so (HIP, gfx950); there is no CPU
fallback.  Reads are duck-typed like SamRecord: ``.seq`` and ``.qual`` (str or
bytes) and ``insertion_gop()`` / ``deletion_gop()`` / ``overall_gcp()``;
haplotypes need ``.seq``.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from ._build import HIP_LIB

BV_OK = 0
MODE_AVX = 0
MODE_SCALAR = 1
FORM_NAMES = {0: "fp32", 1: "fp64", 2: "band_fp32", 3: "band_fallback", 4: "band_fp64", 5: "band_long_fp32"}


class PairHMMError(RuntimeError):
    pass


class _RegionBatch(C.Structure):
    _fields_ = [
        ("n_regions", C.c_int64), ("region_read_off", C.c_void_p), ("region_hap_off", C.c_void_p),
        ("n_reads", C.c_int64), ("read_off", C.c_void_p), ("read_bases", C.c_void_p),
        ("read_qual", C.c_void_p), ("read_ins", C.c_void_p), ("read_del", C.c_void_p),
        ("read_gcp", C.c_void_p), ("n_haps", C.c_int64), ("hap_off", C.c_void_p),
        ("hap_bases", C.c_void_p),
    ]


class Options(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("fp64_renorm", C.c_int32), ("force_fp64", C.c_int32),
        ("grid_f32", C.c_int32), ("grid_f64", C.c_int32), ("band_tol", C.c_float),
        ("band_narrow", C.c_int32), ("use_graph", C.c_int32), ("trans", C.c_double * 7),
    ]


class PlanStats(C.Structure):
    _fields_ = [
        ("n_pairs", C.c_int64), ("cells", C.c_int64), ("fp64_pairs", C.c_int64),
        ("fp64_cells", C.c_int64), ("tasks", C.c_int64), ("grid_f32", C.c_int32),
        ("grid_f64", C.c_int32), ("last_ms_f32", C.c_float), ("last_ms_f64", C.c_float),
        ("band_pairs", C.c_int64), ("fallback_pairs", C.c_int64), ("cells_swept", C.c_int64),
        ("last_ms_fallback", C.c_float), ("last_ms_band", C.c_float),
        ("long_band_pairs", C.c_int64), ("last_ms_long", C.c_float), ("last_ms_post", C.c_float),
        ("setup_ms_host", C.c_float), ("setup_ms_upload", C.c_float), ("setup_bytes", C.c_int64),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("compute_units", C.c_int32),
                ("blocks_per_cu_f32", C.c_int32), ("blocks_per_cu_f64", C.c_int32),
                ("arch", C.c_char * 64)]


class _Testcase(C.Structure):
    _fields_ = [("rslen", C.c_int), ("haplen", C.c_int), ("q", C.c_char_p), ("i", C.c_char_p),
                ("d", C.c_char_p), ("c", C.c_char_p), ("hap", C.c_char_p), ("rs", C.c_char_p)]


# Symbols declared in include/bv_pairhmm.h.
EXPORTS = (
    "bv_pairhmm_init", "bv_pairhmm_shutdown", "bv_pairhmm_last_error", "bv_pairhmm_device_info",
    "bv_pairhmm_tables", "bv_compute_full_prob", "bv_options_default", "bv_plan_create",
    "bv_plan_run", "bv_plan_time", "bv_plan_results", "bv_plan_stats_get", "bv_plan_destroy",
    "bv_plan_autotune", "bv_plan_set_grids", "bv_plan_set_band_shape", "bv_plan_band_shape",
    "bv_plan_set_band_prologue", "bv_plan_set_band_mid", "bv_plan_band_mid",
    "bv_plan_set_quad_fill", "bv_plan_quad_fill",
    "bv_plan_set_long_precision", "bv_plan_long_precision",
    "bv_plan_postprocess", "bv_sw_align", "bv_sw_release",
    "bv_pairhmm_batch",
)

BV_POST_NORMALIZE = 1

_lib = None


def library() -> C.CDLL:
    """Loads libbvpairhmm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not HIP_LIB.exists():
            raise PairHMMError(f"{HIP_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(HIP_LIB))
        vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
        lib.bv_pairhmm_init.argtypes = [i32]
        lib.bv_pairhmm_last_error.restype = C.c_char_p
        lib.bv_pairhmm_device_info.argtypes = [C.POINTER(DeviceInfo)]
        lib.bv_pairhmm_tables.argtypes = [vp, vp, vp, vp]
        lib.bv_compute_full_prob.argtypes = [vp, i32, i32, vp]
        lib.bv_options_default.argtypes = [C.POINTER(Options)]
        lib.bv_options_default.restype = None
        lib.bv_plan_create.argtypes = [C.POINTER(_RegionBatch), C.POINTER(Options), C.POINTER(vp)]
        lib.bv_plan_run.argtypes = [vp]
        lib.bv_plan_time.argtypes = [vp, i32, C.POINTER(C.c_float)]
        lib.bv_plan_results.argtypes = [vp, vp, vp]
        lib.bv_plan_stats_get.argtypes = [vp, C.POINTER(PlanStats)]
        lib.bv_plan_autotune.argtypes = [vp, i32, vp]
        lib.bv_plan_set_grids.argtypes = [vp, vp]
        lib.bv_plan_set_band_shape.argtypes = [vp, i32]
        lib.bv_plan_band_shape.argtypes = [vp, vp]
        lib.bv_plan_set_band_prologue.argtypes = [vp, i32]
        lib.bv_plan_set_band_mid.argtypes = [vp, i32]
        lib.bv_plan_band_mid.argtypes = [vp, vp]
        lib.bv_plan_set_quad_fill.argtypes = [vp, i32]
        lib.bv_plan_quad_fill.argtypes = [vp, vp]
        lib.bv_plan_set_long_precision.argtypes = [vp, i32]
        lib.bv_plan_long_precision.argtypes = [vp, vp]
        lib.bv_plan_postprocess.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        lib.bv_plan_destroy.argtypes = [vp]
        lib.bv_plan_destroy.restype = None
        lib.bv_pairhmm_batch.argtypes = [C.POINTER(_RegionBatch), C.POINTER(Options), vp, vp]
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != BV_OK:
        msg = library().bv_pairhmm_last_error().decode(errors="replace")
        raise PairHMMError(f"status {rc}: {msg}")


def default_options(**kw) -> Options:
    o = Options()
    library().bv_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "trans":
            for i, x in enumerate(v):
                o.trans[i] = float(x)
        elif k == "band_tol":
            o.band_tol = float(v)
        else:
            setattr(o, k, int(v))
    return o


def device_info() -> DeviceInfo:
    info = DeviceInfo()
    _check(library().bv_pairhmm_device_info(C.byref(info)))
    return info


def tables():
    """(ph2pr_f32[128], ph2pr_f64[128], mm_f32[8256], mm_f64[8256]) host copies."""
    pf = np.zeros(128, np.float32)
    pd = np.zeros(128, np.float64)
    mf = np.zeros(8256, np.float32)
    md = np.zeros(8256, np.float64)
    _check(library().bv_pairhmm_tables(pf.ctypes.data, pd.ctypes.data, mf.ctypes.data, md.ctypes.data))
    return pf, pd, mf, md


# ---- packed batches ---------------------------------------------------------

class PackedBatch:
    """Owns contiguous numpy arrays in the bv_region_batch layout."""

    def __init__(self, region_read_off, region_hap_off, read_off, read_bases, read_qual,
                 read_ins, read_del, read_gcp, hap_off, hap_bases):
        self.arrays = dict(
            region_read_off=np.ascontiguousarray(region_read_off, np.int64),
            region_hap_off=np.ascontiguousarray(region_hap_off, np.int64),
            read_off=np.ascontiguousarray(read_off, np.int64),
            read_bases=np.ascontiguousarray(read_bases, np.uint8),
            read_qual=np.ascontiguousarray(read_qual, np.uint8),
            read_ins=np.ascontiguousarray(read_ins, np.uint8),
            read_del=np.ascontiguousarray(read_del, np.uint8),
            read_gcp=np.ascontiguousarray(read_gcp, np.uint8),
            hap_off=np.ascontiguousarray(hap_off, np.int64),
            hap_bases=np.ascontiguousarray(hap_bases, np.uint8),
        )
        a = self.arrays
        self.n_regions = len(a["region_read_off"]) - 1
        self.n_reads = len(a["read_off"]) - 1
        self.n_haps = len(a["hap_off"]) - 1
        rr = np.diff(a["region_read_off"])
        hh = np.diff(a["region_hap_off"])
        self.n_pairs = int((rr * hh).sum()) if self.n_regions > 0 else 0

    @classmethod
    def from_synth(cls, b) -> "PackedBatch":
        return cls(b.region_read_off, b.region_hap_off, b.read_off, b.read_bases, b.read_qual,
                   b.read_ins, b.read_del, b.read_gcp, b.hap_off, b.hap_bases)

    @classmethod
    def from_regions(cls, regions) -> "PackedBatch":
        """regions: iterable of (reads, haps); a read is a dict with byte
        strings seq/q/i/d/c, a haplotype is a byte string."""
        rro, rho, ro, ho = [0], [0], [0], [0]
        rs, q, i, d, c, hp = [], [], [], [], [], []
        for reads, haps in regions:
            for r in reads:
                n = len(r["seq"])
                for key, dst in (("seq", rs), ("q", q), ("i", i), ("d", d), ("c", c)):
                    v = r[key]
                    if len(v) != n:
                        raise PairHMMError(f"read field {key} has length {len(v)}, seq has {n}")
                    dst.append(bytes(v))
                ro.append(ro[-1] + n)
            for h in haps:
                hp.append(bytes(h))
                ho.append(ho[-1] + len(h))
            rro.append(rro[-1] + len(reads))
            rho.append(rho[-1] + len(haps))

        def cat(parts):
            return np.frombuffer(b"".join(parts), np.uint8) if parts else np.zeros(0, np.uint8)
        return cls(rro, rho, ro, cat(rs), cat(q), cat(i), cat(d), cat(c), ho, cat(hp))

    def cstruct(self) -> _RegionBatch:
        a = self.arrays
        return _RegionBatch(
            self.n_regions, a["region_read_off"].ctypes.data, a["region_hap_off"].ctypes.data,
            self.n_reads, a["read_off"].ctypes.data, a["read_bases"].ctypes.data,
            a["read_qual"].ctypes.data, a["read_ins"].ctypes.data, a["read_del"].ctypes.data,
            a["read_gcp"].ctypes.data, self.n_haps, a["hap_off"].ctypes.data,
            a["hap_bases"].ctypes.data)


class Plan:
    """Device-resident batch (bv_plan_*): upload once, run many times."""

    def __init__(self, batch: PackedBatch, options: Options | None = None):
        lib = library()
        self._lib = lib
        self.batch = batch
        self.options = options or default_options()
        h = C.c_void_p()
        cb = batch.cstruct()
        _check(lib.bv_plan_create(C.byref(cb), C.byref(self.options), C.byref(h)))
        self._h = h

    def run(self) -> None:
        _check(self._lib.bv_plan_run(self._h))

    def time(self, iters: int) -> float:
        ms = C.c_float()
        _check(self._lib.bv_plan_time(self._h, int(iters), C.byref(ms)))
        return float(ms.value)

    def results(self):
        n = self.batch.n_pairs
        out = np.empty(n, np.float64)
        form = np.empty(n, np.uint8)
        _check(self._lib.bv_plan_results(self._h, out.ctypes.data, form.ctypes.data))
        return out, form

    def autotune(self, iters: int = 1) -> dict:
        """Picks the wide band's shape (pairs per wavefront) and the
        persistent grids (banded, FP32, FP64, long-read band) by timing runs
        of this plan; returns them."""
        g = (C.c_int32 * 4)()
        _check(self._lib.bv_plan_autotune(self._h, int(iters), g))
        return dict(band=g[0], fp32=g[1], fp64=g[2], long=g[3], wide_pairs=self.band_shape())

    def set_band_shape(self, wide_pairs: int) -> None:
        """Pairs per wavefront of the 64-diagonal banded kernel: 4 (default)
        or 1 (bv_plan_set_band_shape)."""
        _check(self._lib.bv_plan_set_band_shape(self._h, int(wide_pairs)))

    def set_band_prologue(self, mode: int) -> None:
        """With a band model: 2 (default) the features kernel writing
        compact records, then one kernel from them to the work lists; 1 one
        kernel from seed features to work lists; 0 the separate features,
        MLP, band-rule and dispatch kernels (bv_plan_set_band_prologue)."""
        _check(self._lib.bv_plan_set_band_prologue(self._h, int(mode)))

    def band_shape(self) -> int:
        v = C.c_int32()
        _check(self._lib.bv_plan_band_shape(self._h, C.byref(v)))
        return int(v.value)

    def set_band_mid(self, on: bool) -> None:
        """The quad shape's 48-diagonal mid band on or off (default): with it
        on, bands of 33-48 diagonals from set_plan_bands run as 48-diagonal
        bands, else as 64-diagonal ones (bv_plan_set_band_mid).  Learned
        bands never take 48 diagonals."""
        _check(self._lib.bv_plan_set_band_mid(self._h, 1 if on else 0))

    def set_long_precision(self, bits: int) -> None:
        """Arithmetic of the long reads' moving band: 32 (default, FP32
        rescaled every 16 rows, form 5) or 64 (FP64, form 4)
        (bv_plan_set_long_precision)."""
        _check(self._lib.bv_plan_set_long_precision(self._h, int(bits)))

    def long_precision(self) -> int:
        v = C.c_int32()
        _check(self._lib.bv_plan_long_precision(self._h, C.byref(v)))
        return int(v.value)

    def set_quad_fill(self, min_pairs: int) -> None:
        """The quad shape's leftovers: a read's last 1-3 wide pairs fill one
        more four-pair wavefront (empty slots idle) when at least min_pairs
        of them are left, else run one per wavefront; 4 never fills
        (bv_plan_set_quad_fill, default 3)."""
        _check(self._lib.bv_plan_set_quad_fill(self._h, int(min_pairs)))

    def quad_fill(self) -> int:
        v = C.c_int32()
        _check(self._lib.bv_plan_quad_fill(self._h, C.byref(v)))
        return int(v.value)

    def band_mid(self) -> bool:
        v = C.c_int32()
        _check(self._lib.bv_plan_band_mid(self._h, C.byref(v)))
        return bool(v.value)

    def set_grids(self, grids: dict) -> None:
        """Sets persistent grids (keys band, fp32, fp64, long; missing or 0
        keeps a grid), clamped to this plan's maxima."""
        g = (C.c_int32 * 4)(*(int(grids.get(k) or 0) for k in ("band", "fp32", "fp64", "long")))
        _check(self._lib.bv_plan_set_grids(self._h, g))

    def postprocess(self, hap_allele=None, n_alleles=None, normalize: bool = True,
                    keep: bool = True, normalized: bool = False):
        """Device post-processing of the last run (bv_plan_postprocess):
        normalize_likelihoods / filter_poorly_modeled_reads per read
        (pairhmm.hpp:83-114) and, given a haplotype -> allele map (global
        haplotype index -> allele index in its region) and the allele count
        per region, Genotyper::marginal_likelihoods (genotyper.hpp:222-241).
        Returns (keep[n_reads] or None, normalised log10 L[n_pairs] or None,
        allele likelihoods (region by region, reads x alleles) or None)."""
        b = self.batch
        k = np.zeros(b.n_reads, np.uint8) if keep else None
        nz = np.empty(b.n_pairs, np.float64) if normalized else None
        A = ha = na = None
        if hap_allele is not None:
            ha = np.ascontiguousarray(hap_allele, np.int32)
            na = np.ascontiguousarray(n_alleles, np.int32)
            rro = b.arrays["region_read_off"]
            A = np.empty(int((np.diff(rro) * na).sum()), np.float64)
        ptr = lambda x: None if x is None else x.ctypes.data  # noqa: E731
        _check(self._lib.bv_plan_postprocess(self._h, BV_POST_NORMALIZE if normalize else 0,
                                             ptr(ha), ptr(na), ptr(k), ptr(nz), ptr(A)))
        return k, nz, A

    def synchronize(self) -> None:
        """Waits for the plan's stream (no copies)."""
        _check(self._lib.bv_plan_results(self._h, None, None))

    def stats(self) -> PlanStats:
        s = PlanStats()
        _check(self._lib.bv_plan_stats_get(self._h, C.byref(s)))
        return s

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.bv_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def compute_batch(batch: PackedBatch, options: Options | None = None):
    """log10 likelihood and form per pair (region-major, read-major)."""
    with Plan(batch, options) as p:
        p.run()
        return p.results()


def compute_full_prob(testcases: Sequence[dict], precision: int = 32) -> np.ndarray:
    """Scaled sums S (not logs) as compute_full_prob_avx{s,d} returns them.
    Each testcase is a dict with byte strings rs, q, i, d, c, hap."""
    n = len(testcases)
    arr = (_Testcase * max(n, 1))()
    keep = []
    for k, t in enumerate(testcases):
        fields = {key: bytes(t[key]) for key in ("rs", "q", "i", "d", "c", "hap")}
        keep.append(fields)
        arr[k] = _Testcase(len(fields["rs"]), len(fields["hap"]), fields["q"], fields["i"],
                           fields["d"], fields["c"], fields["hap"], fields["rs"])
    out = np.zeros(n, np.float64)
    _check(library().bv_compute_full_prob(C.cast(arr, C.c_void_p), n, int(precision),
                                          out.ctypes.data))
    return out


# ---- reference-shaped class API --------------------------------------------

def _b(x) -> bytes:
    if isinstance(x, (bytes, bytearray, memoryview)):
        return bytes(x)
    if isinstance(x, str):
        return x.encode("latin-1")
    return bytes(x)


def _read_fields(rd) -> dict:
    seq = _b(rd.seq)
    n = len(seq)

    def field(name, default):
        f = getattr(rd, name, None)
        v = f() if callable(f) else f
        return _b(v)[:n] if v is not None else default
    return dict(seq=seq, q=_b(rd.qual), i=field("insertion_gop", None),
                d=field("deletion_gop", None), c=field("overall_gcp", None))


class PairHMM:
    """Scalar PairHMM (pairhmm.hpp:14-142): constant transitions, FP64."""

    # pairhmm.hpp:27-36
    TRISTATE_CORRECTION = 3.0
    MAXIMUM_BEST_ALT_LIKELIHOOD_DIFFERENCE = -4.5
    EXPECTED_ERROR_RATE_PER_BASE = 0.02
    LOG10_QUALITY_PER_BASE = -4.0
    MAXIMUM_EXPECTED_ERROR_PER_READ = 2.0
    ORIGINAL_DEFAULT = (0.9998, 0.0001, 0.0001, 0.9, 0.1, 0.9, 0.1)

    @staticmethod
    def normalize_likelihoods(log_likelihoods: list) -> None:
        """pairhmm.hpp:83-93: cap every entry at (row best - 4.5), in place."""
        for row in log_likelihoods:
            if not row:
                continue
            cap = max(row) + PairHMM.MAXIMUM_BEST_ALT_LIKELIHOOD_DIFFERENCE
            for j, v in enumerate(row):
                if v < cap:
                    row[j] = cap

    @staticmethod
    def filter_poorly_modeled_reads(reads: list, log_likelihoods: list) -> None:
        """pairhmm.hpp:95-114 as written: rows whose best likelihood is BELOW
        min(2, ceil(0.02*len)) * -4 are kept, the others dropped (reads too)."""
        keep_r, keep_l = [], []
        for rd, row in zip(reads, log_likelihoods):
            best = max(row)
            thr = min(PairHMM.MAXIMUM_EXPECTED_ERROR_PER_READ,
                      math.ceil(len(rd.seq) * PairHMM.EXPECTED_ERROR_RATE_PER_BASE)) \
                * PairHMM.LOG10_QUALITY_PER_BASE
            if best < thr:
                keep_r.append(rd)
                keep_l.append(row)
        reads[:] = keep_r
        log_likelihoods[:] = keep_l

    @staticmethod
    def _raw(haplotypes, reads, options: Options) -> list:
        rfields = [_read_fields(r) for r in reads]
        for f in rfields:
            n = len(f["seq"])
            for k in ("i", "d", "c"):
                if f[k] is None:
                    f[k] = b"\0" * n
        haps = [_b(h.seq) for h in haplotypes]
        if not reads:
            return []
        batch = PackedBatch.from_regions([(rfields, haps)])
        out, _ = compute_batch(batch, options)
        nh = len(haps)
        return [list(out[i * nh:(i + 1) * nh]) for i in range(len(reads))]

    def compute_likelihoods(self, haplotypes, reads: list, trans=None) -> list:
        opt = default_options(mode=MODE_SCALAR, fp64_renorm=0,
                              trans=trans if trans is not None else self.ORIGINAL_DEFAULT)
        ll = self._raw(haplotypes, reads, opt)
        self.normalize_likelihoods(ll)
        self.filter_poorly_modeled_reads(reads, ll)
        return ll


class AvxPairHMM(PairHMM):
    """AvxPairHMM (pairhmm_avx.hpp:24-172): per-base GOP/GCP transitions,
    FP32 with FP64 re-run below 1e-28.  Unlike the reference object it can be
    called more than once (the reference accumulates stale testcases)."""

    def __init__(self, fp64_renorm: bool = True):
        self.fp64_renorm = bool(fp64_renorm)

    def compute_likelihoods(self, haplotypes, reads: list) -> list:  # type: ignore[override]
        opt = default_options(mode=MODE_AVX, fp64_renorm=int(self.fp64_renorm))
        for r in reads:
            f = _read_fields(r)
            if f["i"] is None or f["d"] is None or f["c"] is None:
                raise PairHMMError("AvxPairHMM needs insertion_gop/deletion_gop/overall_gcp")
        ll = self._raw(haplotypes, reads, opt)
        self.normalize_likelihoods(ll)
        self.filter_poorly_modeled_reads(reads, ll)
        return ll


GpuPairHMM = AvxPairHMM
