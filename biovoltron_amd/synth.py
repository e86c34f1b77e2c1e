"""Synthetic PairHMM regions (wraps libbv_synth.so; generator in csrc/synth.cpp).

`generate(workload, seed, first, count)` returns a `Batch` whose arrays are
numpy views of library-owned memory, released when the Batch is collected.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._build import SYNTH_LIB

# Fixed seeds of the named workloads (recorded so every number is reproducible).
SEED_BASE = 20261203


def workload_seed(workload: int) -> int:
    return SEED_BASE + workload


class _CBatch(C.Structure):
    _fields_ = [
        ("n_regions", C.c_int64), ("n_reads", C.c_int64), ("n_haps", C.c_int64),
        ("n_pairs", C.c_int64), ("read_bytes", C.c_int64), ("hap_bytes", C.c_int64),
        ("cells", C.c_int64),
        ("region_read_off", C.POINTER(C.c_int64)), ("region_hap_off", C.POINTER(C.c_int64)),
        ("region_pair_off", C.POINTER(C.c_int64)), ("region_workload", C.POINTER(C.c_uint8)),
        ("read_off", C.POINTER(C.c_int64)), ("read_bases", C.POINTER(C.c_uint8)),
        ("read_qual", C.POINTER(C.c_uint8)), ("read_ins", C.POINTER(C.c_uint8)),
        ("read_del", C.POINTER(C.c_uint8)), ("read_gcp", C.POINTER(C.c_uint8)),
        ("hap_off", C.POINTER(C.c_int64)), ("hap_bases", C.POINTER(C.c_uint8)),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not SYNTH_LIB.exists():
            raise RuntimeError(f"{SYNTH_LIB} is not built (run __graft_entry__.build())")
        lib = C.CDLL(str(SYNTH_LIB))
        lib.bv_synth_generate.restype = C.POINTER(_CBatch)
        lib.bv_synth_generate.argtypes = [C.c_int, C.c_uint64, C.c_int64, C.c_int64, C.c_int]
        lib.bv_synth_free.argtypes = [C.POINTER(_CBatch)]
        lib.bv_synth_free.restype = None
        _lib = lib
    return _lib


def _view(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class _Owner:
    def __init__(self, lib, ptr):
        self.lib, self.ptr = lib, ptr

    def __del__(self):
        try:
            self.lib.bv_synth_free(self.ptr)
        except Exception:
            pass


@dataclass
class Batch:
    """Regions in packed structure-of-arrays form.

    Pairs are numbered region by region, read-major inside a region:
    pair = region_pair_off[g] + r_local * n_haps(g) + h_local.
    """
    n_regions: int
    n_reads: int
    n_haps: int
    n_pairs: int
    cells: int
    region_read_off: np.ndarray
    region_hap_off: np.ndarray
    region_pair_off: np.ndarray
    region_workload: np.ndarray
    read_off: np.ndarray
    read_bases: np.ndarray
    read_qual: np.ndarray
    read_ins: np.ndarray
    read_del: np.ndarray
    read_gcp: np.ndarray
    hap_off: np.ndarray
    hap_bases: np.ndarray
    _owner: object = None

    def region(self, g: int):
        """(reads, haps) of region g as lists of dicts of byte strings."""
        r0, r1 = int(self.region_read_off[g]), int(self.region_read_off[g + 1])
        h0, h1 = int(self.region_hap_off[g]), int(self.region_hap_off[g + 1])
        reads = []
        for r in range(r0, r1):
            a, b = int(self.read_off[r]), int(self.read_off[r + 1])
            reads.append(dict(seq=self.read_bases[a:b].tobytes(), q=self.read_qual[a:b].tobytes(),
                              i=self.read_ins[a:b].tobytes(), d=self.read_del[a:b].tobytes(),
                              c=self.read_gcp[a:b].tobytes()))
        haps = [self.hap_bases[int(self.hap_off[h]):int(self.hap_off[h + 1])].tobytes()
                for h in range(h0, h1)]
        return reads, haps

    def pair_table(self) -> tuple[np.ndarray, np.ndarray]:
        """(read index, hap index) of every pair in output order."""
        rr, hh = [], []
        for g in range(self.n_regions):
            r0, r1 = int(self.region_read_off[g]), int(self.region_read_off[g + 1])
            h0, h1 = int(self.region_hap_off[g]), int(self.region_hap_off[g + 1])
            r = np.repeat(np.arange(r0, r1, dtype=np.int64), h1 - h0)
            h = np.tile(np.arange(h0, h1, dtype=np.int64), r1 - r0)
            rr.append(r)
            hh.append(h)
        if not rr:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        return np.concatenate(rr), np.concatenate(hh)


def generate(workload: int, seed: int | None = None, first: int = 0, count: int = 1,
             n_threads: int = 0) -> Batch:
    lib = _load()
    if seed is None:
        seed = workload_seed(workload)
    p = lib.bv_synth_generate(int(workload), int(seed), int(first), int(count), int(n_threads))
    if not p:
        raise MemoryError("bv_synth_generate failed")
    c = p.contents
    own = _Owner(lib, p)
    nr, nh, ng = c.n_reads, c.n_haps, c.n_regions
    return Batch(
        n_regions=ng, n_reads=nr, n_haps=nh, n_pairs=c.n_pairs, cells=c.cells,
        region_read_off=_view(c.region_read_off, ng + 1, np.int64),
        region_hap_off=_view(c.region_hap_off, ng + 1, np.int64),
        region_pair_off=_view(c.region_pair_off, ng + 1, np.int64),
        region_workload=_view(c.region_workload, ng, np.uint8),
        read_off=_view(c.read_off, nr + 1, np.int64),
        read_bases=_view(c.read_bases, c.read_bytes, np.uint8),
        read_qual=_view(c.read_qual, c.read_bytes, np.uint8),
        read_ins=_view(c.read_ins, c.read_bytes, np.uint8),
        read_del=_view(c.read_del, c.read_bytes, np.uint8),
        read_gcp=_view(c.read_gcp, c.read_bytes, np.uint8),
        hap_off=_view(c.hap_off, nh + 1, np.int64),
        hap_bases=_view(c.hap_bases, c.hap_bytes, np.uint8),
        _owner=own,
    )
