"""The header-only C++ adapter (include/biovoltron_amd/gpu_pairhmm.hpp),
through the compiled test program tests/cpp/test_adapter."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT

BIN = ROOT / "tests" / "cpp" / "test_adapter"


def run(mode, *args):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    res = subprocess.run([str(BIN), mode, *args], capture_output=True, text=True, timeout=120)
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
    batch_lines = [ln for ln in out if " batch " in ln]
    assert batch_lines[0].startswith("avx batch equal") and batch_lines[1].startswith("scalar batch equal"), \
        batch_lines
    check_windows_against_oracle(oracle, out)


def parse_windows(lines):
    """{(model, window): (haps, reads [(seq, q)], rows [(seq, values)])}."""
    wins, cur = {}, None
    for ln in lines:
        f = ln.split()
        if f[0] == "window":
            cur = (int(f[1]), int(f[2]))
            wins[cur] = ([], [], [])
        elif cur is None:
            continue
        elif f[0] == "hap":
            wins[cur][0].append(f[1].encode())
        elif f[0] == "read":
            wins[cur][1].append((f[1].encode(), int(f[2])))
        elif f[0] == "row":
            wins[cur][2].append((f[1].encode(), [float(x) for x in f[2:]]))
    return wins


def check_windows_against_oracle(oracle, lines):
    """Each batched window's kept rows against the oracle: raw log10 L per
    pair (the reference rule for the AVX model, the scalar model with
    ORIGINAL_DEFAULT), then the reference's normalise and filter
    (pairhmm.hpp:83-114, restated in the Python mirror, which the
    post-processing tests pin by hand-worked rows)."""
    from biovoltron_amd.pairhmm import PairHMM

    wins = parse_windows(lines)
    assert {m for m, _ in wins} == {0, 1} and len(wins) == 12
    assert any(not reads for _, reads, _ in wins.values())  # a window without reads
    assert len({len(h) for h, _, _ in wins.values()}) == 3   # windows of 1, 2 and 3 haplotypes

    class Rd:
        def __init__(self, seq):
            self.seq = seq

    for (model, w), (haps, reads, rows) in sorted(wins.items()):
        raw = []
        for seq, q in reads:
            n = len(seq)
            rd = dict(seq=seq, q=bytes([q]) * n, i=bytes([45]) * n, d=bytes([45]) * n, c=bytes([10]) * n)
            if model == 0:
                raw.append([oracle.log10_reference(rd, h)[0] for h in haps])
            else:
                raw.append([oracle.scalar_log10(rd, h, PairHMM.ORIGINAL_DEFAULT) for h in haps])
        objs = [Rd(seq) for seq, _ in reads]
        PairHMM.normalize_likelihoods(raw)
        PairHMM.filter_poorly_modeled_reads(objs, raw)
        assert [o.seq for o in objs] == [seq for seq, _ in rows], (model, w)
        for want, (_, got) in zip(raw, rows):
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-3 if model == 0 else 1e-9)


def write_windows(path, wins):
    """Windows [(reads, haps)] in the adapter test's text format."""
    lines = []
    for reads, haps in wins:
        lines.append(f"W {len(haps)} {len(reads)}")
        lines += [f"H {h.decode()}" for h in haps]
        for r in reads:
            assert r["i"] == r["d"] and len(set(r["i"])) == 1 and len(set(r["c"])) == 1
            lines.append(f"R {r['seq'].decode()} {','.join(str(x) for x in r['q'])} {r['i'][0]} {r['c'][0]}")
    path.write_text("\n".join(lines) + "\n")
    return path


@pytest.mark.gpu
def test_shacc_hook_vs_oracle(oracle, tmp_path):
    """shacc_pairhmm::calculate (the weak hook of
    lib/intel_native/pairhmm/shacc_pairhmm.h:10-36, defined by the library)
    called through the reference's declaration: every result is the pair's
    log10 likelihood under the reference's FP32-then-FP64 rule, read-major
    [num_reads][num_haps], as float.  Windows: short-read regions of the
    synthetic generator, and one with N bases, low qualities and reads longer
    than the haplotypes.  An empty batch succeeds; an empty read fails with a
    message."""
    from biovoltron_amd import synth

    b = synth.generate(1, seed=synth.SEED_BASE + 12, count=3)
    wins = []
    for g in range(b.n_regions):
        reads, haps = b.region(g)
        wins.append((reads[:16], haps[:5]))
    rng = np.random.default_rng(5)
    haps = [bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8)) for n in (30, 7, 90)]
    reads = []
    for n in (1, 12, 45, 100):
        s = bytearray(rng.choice(list(b"ACGTN"), n, p=[.23, .23, .23, .23, .08]).astype(np.uint8))
        q = rng.integers(0, 41, n).astype(np.uint8)
        reads.append(dict(seq=bytes(s), q=q.tobytes(), i=bytes([40]) * n, d=bytes([40]) * n, c=bytes([12]) * n))
    wins.append((reads, haps))
    path = write_windows(tmp_path / "windows.txt", wins)
    out = run("shacc", str(path)).strip().splitlines()
    k = 0
    got = {}
    for ln in out:
        f = ln.split()
        if f[0] == "swindow":
            k = int(f[1])
            assert f[2] == "1", ln
            got[k] = []
        elif f[0] == "srow":
            got[k].append([float(x) for x in f[1:]])
    assert len(got) == len(wins)
    worst = 0.0
    for w, (reads, haps) in enumerate(wins):
        assert len(got[w]) == len(reads)
        for rd, row in zip(reads, got[w]):
            assert len(row) == len(haps)
            for hap, v in zip(haps, row):
                want = oracle.log10_reference(rd, hap)[0]
                worst = max(worst, abs(v - want))
                assert abs(v - want) <= 1e-3, (w, len(rd["seq"]), len(hap), v, want)
    print("[shacc] max |dlog10| vs oracle", worst, flush=True)
    assert [ln for ln in out if ln.startswith("sempty")] == ["sempty 1"]
    bad = [ln for ln in out if ln.startswith("sbad")][0]
    assert bad.startswith("sbad 0") and "empty read" in bad


@pytest.mark.gpu
def test_adapter_learned_band_vs_oracle(oracle, tmp_path):
    """GpuPairHMM(LearnedBand{}) with the model shipped beside the library:
    HaplotypeCaller-shaped windows (short-read regions of the synthetic
    generator, up to 24 reads and 6 haplotypes each) in one batched call.
    The kept rows match the oracle's reference rule after the reference's
    normalise and filter, match the full-matrix adapter within 1e-3, and the
    same windows through the C ABI put most pairs on the banded kernel."""
    from biovoltron_amd import synth
    from biovoltron_amd.pairhmm import PairHMM

    b = synth.generate(1, seed=synth.SEED_BASE + 11, count=6)
    wins = []
    for g in range(b.n_regions):
        reads, haps = b.region(g)
        wins.append((reads[:24], haps[:6]))
    path = write_windows(tmp_path / "windows.txt", wins)
    out = run("band", str(path)).strip().splitlines()
    assert out[0] == "learned 1"
    rows = {}
    cur = None
    for ln in out[1:]:
        f = ln.split()
        if f[0] == "bwindow":
            cur = (int(f[1]), int(f[2]))
            rows[cur] = []
        elif f[0] == "brow":
            rows[cur].append((f[1].encode(), np.array([float(x) for x in f[2:]])))
    stats = [ln for ln in out if ln.startswith("bstats")][0].split()
    n_pairs, band_pairs = int(stats[1]), int(stats[2])
    assert band_pairs >= 0.5 * n_pairs, stats
    assert [ln for ln in out if ln.startswith("single")][0].split()[1] == str(len(rows[(0, 0)]))

    class Rd:
        def __init__(self, seq):
            self.seq = seq

    differ = 0
    for w, (reads, haps) in enumerate(wins):
        raw = [[oracle.log10_reference(rd, h)[0] for h in haps] for rd in reads]
        objs = [Rd(rd["seq"]) for rd in reads]
        PairHMM.normalize_likelihoods(raw)
        PairHMM.filter_poorly_modeled_reads(objs, raw)
        for model in (0, 1):
            got = rows[(model, w)]
            assert [s for s, _ in got] == [o.seq for o in objs], (model, w)
            for want, (_, v) in zip(raw, got):
                np.testing.assert_allclose(v, want, rtol=0, atol=1e-3)
        for (_, a), (_, c) in zip(rows[(0, w)], rows[(1, w)]):
            assert np.abs(a - c).max() <= 1e-3
            differ += int((a != c).any())
    # banded FP32 sums round differently from the full matrix: the band ran
    assert differ > 0
