import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"
ORACLE_SO = ROOT / "oracle" / "liboracle_pairhmm.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref_pairhmm.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


class Oracle:
    """ctypes view of oracle/liboracle_pairhmm.so (the CPU checker)."""

    def __init__(self, path):
        lib = C.CDLL(str(path))
        P, I, D, F = C.c_char_p, C.c_int, C.c_double, C.c_float
        lib.bvo_prob_f32.restype = F
        lib.bvo_prob_f32.argtypes = [I, I] + [P] * 6 + [I]
        lib.bvo_prob_f64.restype = D
        lib.bvo_prob_f64.argtypes = [I, I] + [P] * 6 + [I]
        lib.bvo_log10_reference.restype = D
        lib.bvo_log10_reference.argtypes = [I, I] + [P] * 6 + [I, C.c_void_p]
        lib.bvo_log10_exact.restype = D
        lib.bvo_log10_exact.argtypes = [I, I] + [P] * 6 + [I]
        lib.bvo_scalar_log10.restype = D
        lib.bvo_scalar_log10.argtypes = [I, I, P, P, P, C.c_void_p]
        lib.bvo_band_posterior.restype = D
        lib.bvo_band_posterior.argtypes = [I, I] + [P] * 6 + [C.c_void_p, C.c_void_p]
        lib.bvo_band_forward.restype = D
        lib.bvo_band_forward.argtypes = [I, I] + [P] * 6 + [I, I, I, C.c_void_p]
        lib.bvo_band_follow_forward.restype = D
        lib.bvo_band_follow_forward.argtypes = [I, I] + [P] * 6 + [I] * 7 + [D, I] + [C.c_void_p] * 4
        lib.bvo_region_log10.restype = C.c_long
        lib.bvo_region_log10.argtypes = [I] + [C.c_void_p] * 6 + [I, C.c_void_p, C.c_void_p, I, C.c_void_p]
        for n in ("bvo_ph2pr_f32", "bvo_mm_f32"):
            getattr(lib, n).restype = F
        for n in ("bvo_ph2pr_f64", "bvo_mm_f64"):
            getattr(lib, n).restype = D
        lib.bvo_init()
        self.lib = lib

    @staticmethod
    def _args(rd, hap):
        return (len(rd["seq"]), len(hap), rd["seq"], rd["q"], rd["i"], rd["d"], rd["c"], hap)

    def prob_f32(self, rd, hap, ftz=1):
        return self.lib.bvo_prob_f32(*self._args(rd, hap), ftz)

    def prob_f64(self, rd, hap, ftz=1):
        return self.lib.bvo_prob_f64(*self._args(rd, hap), ftz)

    def log10_reference(self, rd, hap):
        r = C.c_int(0)
        v = self.lib.bvo_log10_reference(*self._args(rd, hap), 1, C.byref(r))
        return v, r.value

    def log10_exact(self, rd, hap, mode=0):
        return self.lib.bvo_log10_exact(*self._args(rd, hap), mode)

    def scalar_log10(self, rd, hap, trans):
        t = np.ascontiguousarray(trans, np.float64)
        return self.lib.bvo_scalar_log10(len(rd["seq"]), len(hap), rd["seq"], rd["q"], hap, t.ctypes.data)

    def band_forward(self, rd, hap, o_lo, width, edge=4):
        """(log10 banded likelihood, touch ratio) of the banded FP64 forward."""
        t = C.c_double(0.0)
        v = self.lib.bvo_band_forward(*self._args(rd, hap), int(o_lo), int(width), int(edge), C.byref(t))
        return v, t.value

    def band_follow(self, rd, hap, o_init, width=64, edge=4, every=16, max_move=4, end_rows=0,
                    end_bias=0, hold_eps=0.0, irun=0, sched=None):
        """(log10 banded likelihood, touch ratio, dropped share, per-row band
        offsets) of the FP64 forward whose band moves (towards the row mass
        centroid, or along sched)."""
        t, lost = C.c_double(0.0), C.c_double(0.0)
        offs = np.zeros(len(rd["seq"]) + 1, np.int32)
        sp = None
        if sched is not None:
            sched = np.ascontiguousarray(sched, np.int32)
            sp = sched.ctypes.data
        v = self.lib.bvo_band_follow_forward(*self._args(rd, hap), int(o_init), int(width), int(edge),
                                             int(every), int(max_move), int(end_rows), int(end_bias),
                                             float(hold_eps), int(irun), sp, C.byref(t), C.byref(lost),
                                             offs.ctypes.data)
        return v, t.value, lost.value, offs

    def band_posterior(self, rd, hap):
        R, H = len(rd["seq"]), len(hap)
        mass = np.zeros(R + H - 1)
        rows = np.zeros(R)
        ll = self.lib.bvo_band_posterior(*self._args(rd, hap), mass.ctypes.data, rows.ctypes.data)
        return ll, mass, rows


def _ensure_oracle():
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return Oracle(ORACLE_SO)


@pytest.fixture(scope="session")
def oracle():
    return _ensure_oracle()


@pytest.fixture(scope="session")
def reflib():
    """The reference kernels built from the reference tree, where present."""
    if not REF_SO.exists():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lib = C.CDLL(str(REF_SO))
    P, I = C.c_char_p, C.c_int
    lib.ref_prob_f32.restype = C.c_float
    lib.ref_prob_f32.argtypes = [I, I] + [P] * 6 + [I]
    lib.ref_prob_f64.restype = C.c_double
    lib.ref_prob_f64.argtypes = [I, I] + [P] * 6 + [I]
    return lib


def load_pairs(name):
    """(pairs, data) from a golden fixture: pairs is a list of (read dict, hap)."""
    d = dict(np.load(GOLDEN / name, allow_pickle=False))
    ro, ho = d["read_off"], d["hap_off"]
    pairs = []
    for k in range(len(ro) - 1):
        a, b = int(ro[k]), int(ro[k + 1])
        rd = dict(seq=d["read_bases"][a:b].tobytes(), q=d["read_qual"][a:b].tobytes(),
                  i=d["read_ins"][a:b].tobytes(), d=d["read_del"][a:b].tobytes(),
                  c=d["read_gcp"][a:b].tobytes())
        pairs.append((rd, d["hap_bases"][int(ho[k]):int(ho[k + 1])].tobytes()))
    return pairs, d


@pytest.fixture(scope="session")
def golden():
    return load_pairs
