"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU): the
oracle's C restatement, the synthetic workload generator and the band model
file reader/writer, built by biovoltron_amd._build.build_sanitize() (part of
__graft_entry__.build()) into tests/cpp/sanitize and driven by
tests/cpp/sanitize_main.cpp over random and edge-case inputs.  A sanitizer
report aborts the program; the test wants a clean exit."""
import subprocess

import pytest

from conftest import ROOT

BIN = ROOT / "tests" / "cpp" / "sanitize"


def test_host_code_clean_under_asan_ubsan(tmp_path):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300, cwd=tmp_path,
                         env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
                              "OMP_NUM_THREADS": "2", "PATH": "/usr/bin:/bin"})
    assert res.returncode == 0, (res.stdout[-2000:], res.stderr[-4000:])
    assert res.stdout.strip().endswith("sanitize: ok")
