"""The C++ drop-in compiles against the reference's own types: a translation
unit that includes the reference's PairHMM header (which brings SamRecord<>,
Haplotype and QualityUtils, include/biovoltron/file_io/sam.hpp:186-259,
utility/haplotype/haplotype.hpp:11-26) beside include/biovoltron_amd/
gpu_pairhmm.hpp, calls every GpuPairHMM entry point with those types, and
checks that the results have the reference's result type.  CPU only: it
runs where the reference tree is present and needs no device (syntax and
type checking; the GPU tests run the same adapter on stand-in records)."""
import shutil
import subprocess
from pathlib import Path

import pytest

from conftest import ROOT

REF_INCLUDE = Path("/root/reference/include")
pytestmark = pytest.mark.skipif(not (REF_INCLUDE / "biovoltron").is_dir() or shutil.which("g++") is None,
                                reason="the reference tree or g++ is absent")

TU = r"""
#include <biovoltron/algo/align/inexact_match/pairhmm.hpp>
#include <biovoltron_amd/gpu_pairhmm.hpp>
#include <type_traits>
using namespace biovoltron;
int main() {
  std::vector<Haplotype> haps;
  std::vector<SamRecord<>> reads;
  const biovoltron_amd::GpuPairHMM pairhmm;
  auto ll = pairhmm.compute_likelihoods(haps, reads);
  auto ll2 = pairhmm.compute_likelihoods(haps, reads, PairHMM::ORIGINAL_DEFAULT);
  static_assert(std::is_same_v<decltype(ll), decltype(PairHMM{}.compute_likelihoods(haps, reads))>);
  static_assert(std::is_same_v<decltype(ll2), decltype(ll)>);
  std::vector<std::pair<const std::vector<Haplotype>*, std::vector<SamRecord<>>*>> windows;
  auto all = pairhmm.compute_likelihoods_batch(windows);
  static_assert(std::is_same_v<decltype(all), std::vector<decltype(ll)>>);
  auto all2 = pairhmm.compute_likelihoods_batch(windows, PairHMM::ORIGINAL_DEFAULT);
  static_assert(std::is_same_v<decltype(all2), decltype(all)>);
  const biovoltron_amd::GpuPairHMM learned{biovoltron_amd::LearnedBand{}};
  const biovoltron_amd::GpuPairHMM copy = learned;
  (void)copy.compute_likelihoods(haps, reads);
  return 0;
}
"""


def test_adapter_compiles_against_reference_types(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(TU)
    res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", f"-I{REF_INCLUDE}",
                          f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr[-3000:]
