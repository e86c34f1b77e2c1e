"""Smith-Waterman alignment of haplotypes to a reference window on the GPU
(bv_sw_align), with the interface of the reference's SmithWaterman
(include/biovoltron/algo/align/inexact_match/smithwaterman.hpp:17-255).

* ``SmithWaterman.align(ref, alt, params)`` -- one pair, returns
  ``(offset, cigar_string)`` as the reference's ``align`` does (:229-255).
* ``align_batch(pairs, params)`` -- many pairs in one launch, returns
  ``[(offset, [(op, length), ...]), ...]``.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .pairhmm import PairHMMError, _check, library


class _SwBatch(C.Structure):
    _fields_ = [("n_pairs", C.c_int64), ("ref_off", C.c_void_p), ("ref", C.c_void_p),
                ("alt_off", C.c_void_p), ("alt", C.c_void_p)]


def _lib():
    lib = library()
    if not getattr(lib, "_sw_ready", False):
        lib.bv_sw_align.argtypes = [C.POINTER(_SwBatch), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.POINTER(C.c_float)]
        lib._sw_ready = True
    return lib


def _bytes(s) -> bytes:
    return s.encode() if isinstance(s, str) else bytes(s)


def align_batch(pairs, params=None, timing: dict | None = None):
    """pairs: iterable of (ref, alt) (str or bytes).  Returns per pair
    (offset, [(op, length), ...]).  With `timing`, timing['kernel_ms'] gets
    the kernel time and timing['call_ms'] the library call's."""
    params = SmithWaterman.NEW_SW_PARAMETERS if params is None else params
    refs = [_bytes(r) for r, _ in pairs]
    alts = [_bytes(a) for _, a in pairs]
    n = len(refs)
    if n == 0:
        return []
    ro = np.zeros(n + 1, np.int64)
    ao = np.zeros(n + 1, np.int64)
    ro[1:] = np.cumsum([len(r) for r in refs])
    ao[1:] = np.cumsum([len(a) for a in alts])
    rb = np.frombuffer(b"".join(refs), np.uint8) if ro[-1] else np.zeros(1, np.uint8)
    ab = np.frombuffer(b"".join(alts), np.uint8) if ao[-1] else np.zeros(1, np.uint8)
    slots = np.zeros(n + 1, np.int64)
    slots[1:] = np.cumsum(np.diff(ro) + np.diff(ao) + 3)
    p = np.ascontiguousarray(params, np.int32)
    off = np.zeros(n, np.int32)
    cnt = np.zeros(n, np.int32)
    ops = np.zeros(max(1, int(slots[-1])), np.uint8)
    lens = np.zeros(max(1, int(slots[-1])), np.int32)
    b = _SwBatch(n, ro.ctypes.data, rb.ctypes.data, ao.ctypes.data, ab.ctypes.data)
    ms = C.c_float()
    t0 = time.perf_counter()
    _check(_lib().bv_sw_align(C.byref(b), p.ctypes.data, off.ctypes.data, cnt.ctypes.data, ops.ctypes.data,
                              lens.ctypes.data, C.byref(ms)))
    if timing is not None:
        timing["kernel_ms"] = float(ms.value)
        timing["call_ms"] = (time.perf_counter() - t0) * 1e3  # the C-ABI call: copies, kernels, results
    out = []
    for k in range(n):
        s = int(slots[k])
        out.append((int(off[k]), [(chr(ops[s + e]), int(lens[s + e])) for e in range(int(cnt[k]))]))
    return out


def release() -> None:
    """Frees the aligner's scratch that the library keeps between calls
    (bv_sw_release)."""
    lib = _lib()
    lib.bv_sw_release.restype = None
    lib.bv_sw_release()


def cigar_string(elements) -> str:
    """Cigar's string form (cigar.hpp:340-344): every element, zero-length
    ones included."""
    return "".join(f"{n}{op}" for op, n in elements)


class SmithWaterman:
    """smithwaterman.hpp:17-255 on the GPU."""

    ORIGINAL_DEFAULT = (3, -1, -4, -3)
    STANDARD_NGS = (25, -50, -110, -6)
    NEW_SW_PARAMETERS = (200, -150, -260, -11)
    ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS = (10, -15, -30, -5)
    MAX_MISMATCHES = 2

    @staticmethod
    def align(ref, alt, params=NEW_SW_PARAMETERS):
        if not ref or not alt:
            raise PairHMMError("SmithWaterman.align: empty sequence")
        off, el = align_batch([(ref, alt)], params)[0]
        return off, cigar_string(el)
