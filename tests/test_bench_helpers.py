"""Host-side helpers of bench.py: the committed rocprof summary parses into
the HBM traffic the roofline reports, and the Smith-Waterman pairs have the
haplotype caller's shape."""
from __future__ import annotations

import sys

from conftest import ROOT

sys.path.insert(0, str(ROOT))


def test_profiled_traffic_reads_the_committed_summary():
    import bench

    band = bench.profiled_traffic("pairhmm_workload4.md", "learned",
                                  ["band_quad_kernel<4>", "band_forward_kernel<64>", "band_forward_kernel<32>"])
    full = bench.profiled_traffic("pairhmm_workload4.md", "full", ["full-matrix FP32 pass (short reads, stacked)"])
    assert band is not None and full is not None
    # the three banded kernels together move over a hundred GB per run; the
    # stacked full-matrix FP32 pass keeps its carry row in LDS and moves
    # about a hundred GB (read bytes and per-read sums)
    assert 1e11 < band[0] < 1e12 and 1e10 < full[0] < 1e12
    assert bench.profiled_traffic("pairhmm_workload4.md", "learned", ["no_such_kernel"]) is None


def test_sw_pairs_shape():
    import bench
    from biovoltron_amd import synth

    b = synth.generate(4, count=40)
    pairs = bench.sw_pairs(b, 40)
    assert pairs
    nh = [int(b.region_hap_off[g + 1] - b.region_hap_off[g]) for g in range(40)]
    assert len(pairs) <= sum(k - 1 for k in nh if k > 1)
    assert all(1 <= len(r) <= 1024 and 1 <= len(a) <= 1024 for r, a in pairs)
