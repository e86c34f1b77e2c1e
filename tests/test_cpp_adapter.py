"""The header-only C++ adapter (include/biovoltron_amd/gpu_pairhmm.hpp),
through the compiled test program tests/cpp/test_adapter."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = ROOT / "tests" / "cpp" / "test_adapter"


def run(mode):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    res = subprocess.run([str(BIN), mode], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def test_adapter_cpu_semantics():
    assert "cpu tests: ok" in run("cpu")


@pytest.mark.gpu
def test_adapter_gpu_matches_oracle(oracle):
    out = run("gpu").strip().splitlines()
    mat = np.array([[float(x) for x in line.split()] for line in out[:3]])
    base = "AATCGAAGGTCGTAAGGACACGGTTGAGCGTTCAGCGTTCATGTGAGTCCTCACCACTTATGGCTCCATAGCC"
    haps = [base, base[:30] + "T" + base[31:], base[5:]]
    reads = [(base[10:50], 35), (base[:50], 25), (base[20:50] + "GG", 20)]
    for i, (s, q) in enumerate(reads):
        n = len(s)
        rd = dict(seq=s.encode(), q=bytes([q]) * n, i=bytes([45]) * n, d=bytes([45]) * n,
                  c=bytes([10]) * n)
        for j, h in enumerate(haps):
            ref, _ = oracle.log10_reference(rd, h.encode())
            assert abs(mat[i, j] - ref) <= 1e-3
    assert out[3].startswith("kept") and out[4].startswith("scalar kept")
    # HaplotypeCaller batching: all windows in one run equal one call per window
    assert out[5].startswith("avx batch equal") and out[6].startswith("scalar batch equal"), out[5:]
