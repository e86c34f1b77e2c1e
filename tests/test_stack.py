"""The stacked full-matrix kernel (reads of a region stacked into one tall
matrix, swept in 63-row stripes with the carry row in LDS) against the
oracle, on regions built to hit its boundaries: stripes that end reads in
many lanes (more than the per-stripe limit of read ends), reads of one base,
read lengths around the 63-row stripe, reads that start with an N or a
quality of 0 or 1 (they must start a stripe), haplotypes from 1 to 1024 bases,
and a region whose 1025-base haplotype sends its reads to the striped kernel.
"""
import numpy as np
import pytest

from biovoltron_amd import pairhmm

pytestmark = pytest.mark.gpu

TOL = 1e-3
TOL_FP32 = 1e-4


def say(*a):
    print("[stack]", *a, flush=True)


def read(rng, n, hap=None, start=None, q=None, err=0.02):
    if hap is not None and len(hap) >= n:
        off = int(rng.integers(0, len(hap) - n + 1))
        s = bytearray(hap[off:off + n])
        for k in range(n):
            if rng.random() < err:
                s[k] = int(rng.choice(list(b"ACGT")))
    else:
        s = bytearray(rng.choice(list(b"ACGT"), n).astype(np.uint8))
    quals = rng.integers(10, 41, n).astype(np.uint8)
    if start is not None:
        s[0] = start
    if q is not None:
        quals[0] = q
    if n > 4 and rng.random() < 0.2:
        s[int(rng.integers(1, n))] = ord("N")
    return dict(seq=bytes(s), q=quals.tobytes(), i=bytes([45]) * n, d=bytes([45]) * n, c=bytes([10]) * n)


def hapseq(rng, n):
    return bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8))


def regions(seed):
    rng = np.random.default_rng(seed)
    out = []
    # many short reads: up to 60 read ends in one stripe
    haps = [hapseq(rng, int(h)) for h in (40, 41, 1, 2)]
    out.append(([read(rng, int(n), haps[0]) for n in rng.integers(1, 6, 90)], haps))
    # lengths around the stripe height, one region each with a few haplotype lengths
    for lens, hl in (([62, 63, 64, 1], [63, 64, 65]), ([125, 126, 127, 2, 3], [200, 1024]),
                     ([1] * 70, [5, 30]), ([250, 249, 7, 188, 189, 190], [500, 600, 64])):
        haps = [hapseq(rng, h) for h in hl]
        out.append(([read(rng, n, haps[0]) for n in lens], haps))
    # reads starting with N or quality 0 / 1 after other reads (they start a stripe)
    haps = [hapseq(rng, 120), hapseq(rng, 119)]
    rs = []
    for k in range(12):
        n = int(rng.integers(3, 40))
        kind = k % 4
        rs.append(read(rng, n, haps[0], start=ord("N") if kind == 1 else None,
                       q=0 if kind == 2 else (1 if kind == 3 else None)))
    out.append((rs, haps))
    # a region with a haplotype too long for the stacked kernel
    haps = [hapseq(rng, 1025), hapseq(rng, 300)]
    out.append(([read(rng, n, haps[1]) for n in (50, 80, 1)], haps))
    # random regions
    for _ in range(20):
        haps = [hapseq(rng, int(h)) for h in rng.integers(1, 700, int(rng.integers(1, 6)))]
        rs = [read(rng, int(n), haps[0]) for n in rng.integers(1, 260, int(rng.integers(1, 30)))]
        out.append((rs, haps))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_stacked_regions_vs_oracle(oracle, seed):
    regs = regions(seed)
    pb = pairhmm.PackedBatch.from_regions(regs)
    say("regions", len(regs), "pairs", pb.n_pairs)
    with pairhmm.Plan(pb) as plan:
        plan.run()
        out, form = plan.results()
        st = plan.stats()
    cells = sum(len(r["seq"]) * len(h) for rs, hs in regs for r in rs for h in hs)
    # every pair runs in FP32 once; pairs rescued to FP64 run again there
    assert st.cells_swept == cells + st.fp64_cells
    k = 0
    worst = worst_exact = 0.0
    for rs, hs in regs:
        for rd in rs:
            for hap in hs:
                ref, _ = oracle.log10_reference(rd, hap)
                exact = oracle.log10_exact(rd, hap, 0)
                if np.isfinite(ref):
                    worst = max(worst, abs(out[k] - ref))
                    assert abs(out[k] - ref) <= TOL, (k, len(rd["seq"]), len(hap), out[k], ref)
                if form[k] == 0:
                    d = abs(out[k] - exact)
                    worst_exact = max(worst_exact, d)
                    assert d <= TOL_FP32 * max(1.0, abs(exact) / 10), (k, len(rd["seq"]), len(hap), out[k], exact)
                k += 1
    say("max |dlog10| vs reference rule", worst, "FP32 form vs FP64 oracle", worst_exact,
        "fp64 pairs", int((form == 1).sum()))
