"""In-tree build of the native libraries (no JIT cache, nothing installed).

libbvpairhmm.so  HIP kernels + the C ABI (hipcc, gfx950)
libbv_synth.so   synthetic region generator (g++, OpenMP)
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
INCLUDE = ROOT / "include"

HIP_LIB = PKG / "libbvpairhmm.so"
SYNTH_LIB = PKG / "libbv_synth.so"

HIP_SOURCES = ["pairhmm_kernels.hip", "pairhmm_stack.hip", "pairhmm_band.hip", "band_kernels.hip", "post_kernels.hip",
               "sw_kernels.hip", "pairhmm_host.cpp", "model_io.cpp"]

# Packed FP32 ops (v_pk_*_f32) are disabled for all device code: some of
# their forms have given wrong results on this platform without an error.
# hipcc's host pass warns that it does not know the feature.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            f"build failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_synth() -> Path:
    tmp = SYNTH_LIB.with_suffix(".so.tmp")
    _run(["g++", "-std=c++20", "-O2", "-fopenmp", "-fPIC", "-shared",
          "-o", str(tmp), str(CSRC / "synth.cpp")])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_hip() -> Path:
    tmp = HIP_LIB.with_suffix(".so.tmp")
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    _run([hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(tmp), *srcs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


CPP_TEST = ROOT / "tests" / "cpp" / "test_adapter"


def build_cpp_tests() -> Path:
    """The C++ adapter test program (links libbvpairhmm.so)."""
    tmp = CPP_TEST.with_suffix(".tmp")
    _run(["g++", "-std=c++20", "-O2", f"-I{INCLUDE}", "-o", str(tmp),
          str(CPP_TEST.with_suffix(".cpp")), f"-L{PKG}", "-lbvpairhmm",
          f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../biovoltron_amd"])
    os.replace(tmp, CPP_TEST)
    return CPP_TEST


SANITIZE = ROOT / "tests" / "cpp" / "sanitize"
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fopenmp"]


def build_sanitize() -> Path:
    """Host code under ASan + UBSan (g++/gcc only, no device code): the
    oracle's C restatement, the synthetic generator and the model file I/O,
    linked with the driver in tests/cpp/sanitize_main.cpp."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        objs = []
        for src, cc, std in ((ROOT / "oracle" / "pairhmm_oracle.c", "gcc", "-std=c11"),
                             (ROOT / "oracle" / "sw_oracle.c", "gcc", "-std=c11"),
                             (CSRC / "synth.cpp", "g++", "-std=c++20"),
                             (CSRC / "model_io.cpp", "g++", "-std=c++20"),
                             (ROOT / "tests" / "cpp" / "sanitize_main.cpp", "g++", "-std=c++20")):
            obj = Path(td) / (src.stem + ".o")
            extra = ["-ffp-contract=off"] if cc == "gcc" else []
            _run([cc, std, *SAN_FLAGS, *extra, f"-I{INCLUDE}", "-c", "-o", str(obj), str(src)])
            objs.append(str(obj))
        tmp = SANITIZE.with_suffix(".tmp")
        _run(["g++", *SAN_FLAGS, "-o", str(tmp), *objs, "-ldl", "-lm"])
        os.replace(tmp, SANITIZE)
    return SANITIZE


def build_oracle(with_reference: bool = True) -> None:
    """The CPU checker (test infrastructure); the reference build only where
    the reference tree exists."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    if with_reference and Path("/root/reference/lib/intel_native/pairhmm").is_dir():
        _run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"])


def build_all() -> None:
    build_synth()
    build_hip()
    build_cpp_tests()
    build_oracle()
    build_sanitize()


if __name__ == "__main__":
    build_all()
    print("built", HIP_LIB, SYNTH_LIB)
