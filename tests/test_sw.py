"""Smith-Waterman haplotype-to-reference alignment (SmithWaterman::align,
smithwaterman.hpp:54-258): the oracle's restatement (oracle/sw_oracle.c)
pinned by the reference test's cases (tests/algo/align/inexact_match/
smithwaterman.cpp), by the reference's own code built from its header
(oracle/_ref/libref_sw.so, where the reference tree exists) and by the
score of the CIGAR it returns; the device kernel (bv_sw_align) equal to it
element for element."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT))

NEW_SW = (200, -150, -260, -11)     # smithwaterman.hpp: NEW_SW_PARAMETERS (align's default)
ORIGINAL = (3, -1, -4, -3)          # ORIGINAL_DEFAULT
STANDARD_NGS = (25, -50, -110, -6)
BEST_HAP = (10, -15, -30, -5)       # ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS
REF_162 = ("AATCGAAGGTCGTAAGGACACGGTTGAGCGTTCAGCGTTCATGTGAGTCCTCACCACTTATGGCTCCATAGCCT"
           "GCTATTTAAGTGGGTTACCGGTCTCCGCCAAGTAGCTGGTGTAAGAACACAGTAACTGAGCCCAGTGTGATCAG"
           "CCCTAACGAGGTAC")


def _olib():
    lib = C.CDLL(str(ROOT / "oracle" / "liboracle_pairhmm.so"))
    f = lib.bvo_sw_align
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def oracle_align(ref: bytes, alt: bytes, params=NEW_SW):
    f = _olib()
    p = np.array(params, np.int32)
    cap = len(ref) + len(alt) + 3
    ops = C.create_string_buffer(cap)
    lens = np.zeros(cap, np.int32)
    off = C.c_int32()
    n = f(len(ref), ref, len(alt), alt, p.ctypes.data, C.byref(off), ops, lens.ctypes.data)
    return off.value, [(ops.raw[k:k + 1].decode(), int(lens[k])) for k in range(n)]


def cigar_str(elems):
    return "".join(f"{n}{op}" for op, n in elems)


def _reference_cases():
    """The reference test's inputs, built as it describes them: the 162-base
    sequence against itself, with 10 substitutions on both sides, with one
    base deleted or inserted in the alternate, and mixed."""
    ref = REF_162
    out = [("same", ref, ref, 0, "162M")]
    r, a = list(ref), list(ref)
    for k in range(70, 80):
        r[k], a[k] = "A", "T"
    out.append(("substitutions", "".join(r), "".join(a), 0, "69M10D1M10I82M"))
    a = list(ref)
    del a[70]
    out.append(("deletion", ref, "".join(a), 0, "70M1D91M"))
    a = list(ref)
    a.insert(70, "T")
    out.append(("insertion", ref, "".join(a), 0, "70M1I92M"))
    r, a = list(ref), list(ref)
    for k in range(len(ref)):
        if 10 < k < 20 or 70 < k < 80 or 120 < k < 130:
            r[k], a[k] = "A", "T"
    del a[60]
    a.insert(90, "T")
    out.append(("mix", "".join(r), "".join(a), 0, "11M9D9I40M1D10M9D9I11M1I28M9D2M9I32M"))
    return out


@pytest.mark.parametrize("name,ref,alt,offset,cigar", _reference_cases())
def test_oracle_reference_test_cases(name, ref, alt, offset, cigar):
    off, el = oracle_align(ref.encode(), alt.encode())
    assert off == offset and cigar_str(el) == cigar, name


def random_pair(rng, n_ref, mut=0.03):
    ref = rng.choice(list(b"ACGT"), n_ref).astype(np.uint8)
    alt = []
    k = 0
    while k < n_ref:
        u = rng.random()
        if u < mut:
            alt.append(int(rng.choice(list(b"ACGT"))))
            k += 1
        elif u < mut * 1.5:
            k += int(rng.integers(1, 6))           # deletion
        elif u < mut * 2:
            alt += list(rng.choice(list(b"ACGT"), int(rng.integers(1, 6))))  # insertion
        else:
            alt.append(int(ref[k]))
            k += 1
    lo = int(rng.integers(0, max(1, len(alt) // 5)))
    hi = len(alt) - int(rng.integers(0, max(1, len(alt) // 5)))
    alt = alt[lo:hi] or alt or [int(ref[0])]  # never empty (align asserts non-empty)
    return bytes(ref), bytes(np.array(alt, np.uint8))


def test_oracle_cigar_consumes_the_alternate_inside_the_reference():
    """Size-independent checks of the restatement on random pairs and every
    parameter set: the CIGAR spans the whole alternate and stays inside the
    reference."""
    rng = np.random.default_rng(5)
    for params in (NEW_SW, ORIGINAL, STANDARD_NGS, BEST_HAP):
        for _ in range(60):
            ref, alt = random_pair(rng, int(rng.integers(20, 200)))
            off, el = oracle_align(ref, alt, params)
            n_alt = sum(n for op, n in el if op in "MIS")
            n_ref = sum(n for op, n in el if op in "MD")
            assert n_alt == len(alt)
            assert 0 <= off and off + n_ref <= len(ref)


def _ref_sw():
    path = ROOT / "oracle" / "_ref" / "libref_sw.so"
    if not path.exists():
        return None
    lib = C.CDLL(str(path))
    f = lib.ref_sw_align
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
    return f


@pytest.mark.skipif(_ref_sw() is None, reason="reference tree not present (oracle/_ref not built)")
def test_oracle_matches_reference_build():
    """The restatement against SmithWaterman::align compiled from the
    reference header, on random pairs of 1-300 bases (ties, clips, runs of
    gaps, equal-length pairs near the well-match shortcut) and every
    parameter set."""
    f = _ref_sw()
    rng = np.random.default_rng(17)
    n = 0
    for params in (NEW_SW, ORIGINAL, STANDARD_NGS, BEST_HAP):
        p = np.array(params, np.int32)
        for k in range(250):
            if k % 5 == 0:
                m = int(rng.integers(1, 40))
                ref = bytes(rng.choice(list(b"ACGT"), m).astype(np.uint8))
                alt = bytes(rng.choice(list(b"ACGT"), int(rng.integers(1, 40))).astype(np.uint8))
            elif k % 5 == 1:
                ref = bytes(rng.choice(list(b"ACGT"), int(rng.integers(10, 120))).astype(np.uint8))
                a = bytearray(ref)
                for _ in range(int(rng.integers(0, 5))):
                    a[int(rng.integers(0, len(a)))] = int(rng.choice(list(b"ACGT")))
                alt = bytes(a)
            else:
                ref, alt = random_pair(rng, int(rng.integers(5, 300)), mut=float(rng.choice([0.01, 0.05, 0.2])))
            off_o, el = oracle_align(ref, alt, params)
            buf = C.create_string_buffer(4 * (len(ref) + len(alt)) + 64)
            off_r = C.c_int32()
            rc = f(len(ref), ref, len(alt), alt, p.ctypes.data, C.byref(off_r), buf, len(buf))
            assert rc >= 0
            assert (off_o, cigar_str(el)) == (off_r.value, buf.value.decode()), (params, ref, alt)
            n += 1
    assert n == 1000


def _device_pairs(rng, n, lo=1, hi=300):
    out = []
    for k in range(n):
        if k % 7 == 0:
            r = bytes(rng.choice(list(b"ACGT"), int(rng.integers(lo, 40))).astype(np.uint8))
            a = bytes(rng.choice(list(b"ACGT"), int(rng.integers(lo, 40))).astype(np.uint8))
        elif k % 7 == 1:
            r = bytes(rng.choice(list(b"ACGT"), int(rng.integers(10, 150))).astype(np.uint8))
            a = bytearray(r)
            for _ in range(int(rng.integers(0, 5))):
                a[int(rng.integers(0, len(a)))] = int(rng.choice(list(b"ACGT")))
            a = bytes(a)
        else:
            r, a = random_pair(rng, int(rng.integers(max(lo, 5), hi)), mut=float(rng.choice([0.01, 0.05, 0.2])))
        out.append((r, a))
    return out


@pytest.mark.gpu
def test_device_sw_reference_cases():
    """The reference test's five cases through the device kernel."""
    from biovoltron_amd import sw

    for name, ref, alt, offset, cigar in _reference_cases():
        print(f"sw case {name}", flush=True)
        assert sw.SmithWaterman.align(ref, alt) == (offset, cigar), name


@pytest.mark.gpu
@pytest.mark.parametrize("params", [NEW_SW, ORIGINAL, STANDARD_NGS, BEST_HAP])
def test_device_sw_equals_oracle(params):
    """bv_sw_align against the oracle, element for element (offset, every
    CIGAR element with zero-length ones), on 400 random pairs of 1-300 bases
    (stripes of 64 rows, equal-length pairs near the well-match shortcut,
    clips, long gaps) in one launch."""
    from biovoltron_amd import sw

    rng = np.random.default_rng(29)
    pairs = _device_pairs(rng, 400)
    print(f"sw: {len(pairs)} pairs, params {params}", flush=True)
    got = sw.align_batch(pairs, params)
    for (r, a), (off, el) in zip(pairs, got):
        assert (off, el) == oracle_align(r, a, params), (r, a)


@pytest.mark.gpu
def test_device_sw_long_and_many():
    """Pairs up to 1500 bases (23 stripes) and a batch larger than the
    persistent grid (scratch reused across pairs)."""
    from biovoltron_amd import sw

    rng = np.random.default_rng(31)
    pairs = [random_pair(rng, int(rng.integers(800, 1500)), mut=0.03) for _ in range(12)]
    pairs += _device_pairs(rng, 3000, hi=200)
    print(f"sw: {len(pairs)} pairs", flush=True)
    got = sw.align_batch(pairs)
    for k in list(range(12)) + list(range(12, len(pairs), 37)):
        assert got[k] == oracle_align(*pairs[k]), k


@pytest.mark.gpu
def test_device_sw_rejects_bad_input():
    from biovoltron_amd import pairhmm, sw

    with pytest.raises(pairhmm.PairHMMError, match="empty"):
        sw.align_batch([(b"ACGT", b"")])
    with pytest.raises(pairhmm.PairHMMError, match="4096"):
        sw.align_batch([(b"A" * 5000, b"ACGT")])
    with pytest.raises(pairhmm.PairHMMError, match="65536"):
        sw.align_batch([(b"ACGT", b"ACGA")], (1 << 20, -1, -4, -3))


@pytest.mark.gpu
def test_device_sw_size_class_edges_and_longest_pairs():
    """Pairs on both sides of every size-class edge (longer side 512/513,
    1024/1025, 2048/2049 bases), pairs of 2,049 to 4,096 bases (the largest
    class, with the largest per-wavefront scratch) and one 4,096 x 4,096
    pair, in one call and again in a second call (the workspace kept from
    the first), element for element against the oracle."""
    from biovoltron_amd import sw

    rng = np.random.default_rng(47)
    pairs = []
    for edge in (512, 1024, 2048):
        for m in (edge, edge + 1):
            r, a = random_pair(rng, m, mut=0.02)
            pairs.append((r[:m], a))              # reference side exactly m bases
            pairs.append((a[: max(1, m // 3)], r[:m]))  # alternate side m bases
    pairs += [random_pair(rng, int(rng.integers(2049, 4000)), mut=0.03) for _ in range(4)]
    r4 = bytes(rng.choice(list(b"ACGT"), 4096).astype(np.uint8))
    a4 = bytearray(r4)
    for k in rng.integers(0, 4096, 60):
        a4[int(k)] = int(rng.choice(list(b"ACGT")))
    pairs.append((r4, bytes(a4)))
    pairs = [(r[:4096], a[:4096]) for r, a in pairs]
    print(f"sw: {len(pairs)} long and edge pairs, longest {max(max(len(r), len(a)) for r, a in pairs)}", flush=True)
    for call in range(2):
        got = sw.align_batch(pairs)
        for k, (r, a) in enumerate(pairs):
            assert got[k] == oracle_align(r, a), (call, k, len(r), len(a))
    # a small call after the large ones reuses the workspace; after a
    # release the next call allocates it again
    small = _device_pairs(rng, 50, hi=120)
    for release in (False, True):
        if release:
            sw.release()
        got = sw.align_batch(small)
        for k in range(len(small)):
            assert got[k] == oracle_align(*small[k]), (release, k)
