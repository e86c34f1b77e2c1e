"""PairHMM throughput on one MI355X.

The default workload is the combined one: 100,000 synthetic regions, 90 % of
them short-read regions (reads of 50-250 bases against 2-32 haplotypes of
100-600 bases) and 10 % long-read regions (reads of 2-10 kb against 4-8
haplotypes of 3-12 kb), about 4.04e8 read x haplotype pairs.  A step is one
full evaluation of it with inputs resident in device memory, as one captured
graph: learned bands for the short reads (full-matrix FP32 fallback), the
moving FP64 band for the long reads (full-matrix FP64 fallback), the FP64
re-run of pairs whose FP32 sum falls below 1e-28, and the final log10.  The
full-matrix evaluation of the same workload is timed beside it.  Prints one
JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 4] [--regions N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GCUPS (read×hap cells/s) learned-band vs full PairHMM at |Δlog10L|≤1e-3"
FLOPS_PER_CELL = 12  # avx-pairhmm-template.h:188-197: 8 mul + 4 add
# Banded kernels are credited with the same 12 flops per band cell.  The
# edge-zone touch mass that judges a band re-runs the recurrence; that second
# pass is overhead of the method, reported beside the roofline, not counted.
BAND_FLOPS_PER_CELL = FLOPS_PER_CELL
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (packed) datasheet peak
PEAK_FP32_NOPACK_TFLOPS = 78.6
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector datasheet peak
LONG_ROWS = 1024  # reads this long run in FP64 (the moving band when banded)
WORKLOADS = {
    0: "one region of 64 reads x 8 haplotypes",
    1: "short-read regions (reads 50-250 bp, 2-32 haplotypes of 100-600 bp)",
    2: "long-read regions (reads 2-10 kb, 4-8 haplotypes of 3-12 kb, 5-10 % errors)",
    4: "combined: 90 % short-read and 10 % long-read regions",
}


def profiled_traffic(profile: str, leg: str, kernels):
    """HBM bytes per launch of the given kernel roles (summed) in one leg of
    the bench, from a committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE
    passes, each scaled by a copy of known size), or None when the summary or
    a role is absent."""
    path = ROOT / "profiles" / profile
    if not path.exists():
        return None
    found = {}
    in_table = False
    for line in path.read_text().splitlines():
        if line.startswith("| leg | kernel (role) | FETCH_SIZE"):
            in_table = True
            continue
        if in_table and line.startswith("|") and not line.startswith("|---"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            name = cells[1].strip("`")
            if cells[0] == leg and name in kernels:
                found[name] = (float(cells[4]) + float(cells[5])) * 1e9
        elif in_table and not line.startswith("|"):
            in_table = False
    if len(found) != len(kernels):
        return None
    return int(sum(found.values())), path.name


def cpu_baseline(batch, out_gpu, budget_s: float):
    """The reference kernels (oracle/_ref, built from the reference sources)
    run as the reference driver runs them -- OpenMP schedule(dynamic,1) over
    reads, min(64, cores) threads, FP32 then FP64 -- region by region over a
    bounded sample of the same workload.  Falls back to the oracle's C
    restatement ("port") where oracle/_ref is absent.  Also reports the GPU's
    max |dlog10 L| against these CPU results on the sampled pairs."""
    ref_so = ROOT / "oracle" / "_ref" / "libref_pairhmm.so"
    port_so = ROOT / "oracle" / "liboracle_pairhmm.so"
    if ref_so.exists():
        lib, fn, kind = C.CDLL(str(ref_so)), "ref_region_log10", "reference"
    else:
        lib, fn, kind = C.CDLL(str(port_so)), "bvo_region_log10", "port"
    f = getattr(lib, fn)
    f.restype = C.c_long
    f.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    cores = len(os.sched_getaffinity(0))
    threads = min(64, cores)  # the reference's cap, pairhmm_avx.hpp:83,88-90
    b = batch
    cells = 0
    elapsed = 0.0
    regions = 0
    worst = 0.0
    rng = np.random.default_rng(1)
    order = rng.permutation(b.n_regions)
    for g in order:
        r0, r1 = int(b.region_read_off[g]), int(b.region_read_off[g + 1])
        h0, h1 = int(b.region_hap_off[g]), int(b.region_hap_off[g + 1])
        nr, nh = r1 - r0, h1 - h0
        ro = np.ascontiguousarray(b.read_off[r0:r1 + 1])
        ho = np.ascontiguousarray(b.hap_off[h0:h1 + 1])
        out = np.zeros(nr * nh)
        t0 = time.perf_counter()
        f(nr, ro.ctypes.data, b.read_bases.ctypes.data, b.read_qual.ctypes.data,
          b.read_ins.ctypes.data, b.read_del.ctypes.data, b.read_gcp.ctypes.data,
          nh, ho.ctypes.data, b.hap_bases.ctypes.data, threads, out.ctypes.data)
        elapsed += time.perf_counter() - t0
        rl = np.diff(ro).sum()
        cells += int(rl) * int(np.diff(ho).sum())
        p0 = int(b.region_pair_off[g])
        fin = np.isfinite(out)
        if fin.any():
            worst = max(worst, float(np.abs(out_gpu[p0:p0 + nr * nh][fin] - out[fin]).max()))
        regions += 1
        if elapsed >= budget_s:
            break
    gcups = cells / elapsed / 1e9
    model, physical = host_cpu()
    return dict(value=round(gcups, 4), unit="GCUPS", cores=threads, kind=kind,
                cpu_model=model, physical_cores=physical,
                threads_note=("min(64, cores) as shipped; equal to every physical core here"
                              if physical is not None and threads >= physical else
                              "min(64, cores) as shipped"),
                sample=f"{regions} of {b.n_regions} regions ({cells:.3e} cells, "
                       f"{100.0 * cells / b.cells:.2f}% of the workload), {elapsed:.1f} s, "
                       f"{os.cpu_count()} host CPUs visible, {cores} in affinity"), worst


def host_cpu():
    """(model name, physical core count) of the host from /proc/cpuinfo."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and core is not None:
                cores.add((phys, core))
                phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, (len(cores) or None)


def sw_pairs(batch, n_regions: int, max_len: int = 1024):
    """Smith-Waterman pairs of the haplotype caller's shape from the first
    regions of the workload: every haplotype of a region aligned to the
    region's first haplotype (standing in for the reference window), as
    haplotype_graph.hpp:219 aligns each haplotype to the reference.  Regions
    with haplotypes longer than max_len (the 3-12 kb long-read regions) are
    left out: a caller's active-region haplotypes are a few hundred bases."""
    b = batch
    pairs = []
    for g in range(min(n_regions, b.n_regions)):
        h0, h1 = int(b.region_hap_off[g]), int(b.region_hap_off[g + 1])
        if h1 - h0 < 2:
            continue
        ref = b.hap_bases[int(b.hap_off[h0]):int(b.hap_off[h0 + 1])].tobytes()
        if len(ref) > max_len:
            continue
        for h in range(h0 + 1, h1):
            alt = b.hap_bases[int(b.hap_off[h]):int(b.hap_off[h + 1])].tobytes()
            if len(alt) <= max_len:
                pairs.append((ref, alt))
    return pairs


def sw_cpu_baseline(pairs, device_results, budget_s: float):
    """SmithWaterman::align built from the reference header (oracle/_ref,
    one thread) on a bounded prefix of the same pairs; also checks the
    device results against it pair by pair."""
    path = ROOT / "oracle" / "_ref" / "libref_sw.so"
    if not path.exists():
        return None
    from biovoltron_amd.sw import SmithWaterman, cigar_string
    f = C.CDLL(str(path)).ref_sw_align
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
    p = np.array(SmithWaterman.NEW_SW_PARAMETERS, np.int32)
    cells = done = mismatches = 0
    t = 0.0
    for (r, a), (off, el) in zip(pairs, device_results):
        buf = C.create_string_buffer(4 * (len(r) + len(a)) + 64)
        o = C.c_int32()
        t0 = time.perf_counter()
        f(len(r), r, len(a), a, p.ctypes.data, C.byref(o), buf, len(buf))
        t += time.perf_counter() - t0
        mismatches += (o.value, buf.value.decode()) != (off, cigar_string(el))
        cells += len(r) * len(a)
        done += 1
        if t >= budget_s:
            break
    return dict(value=round(cells / t / 1e9, 4), unit="GCUPS", cores=1, kind="reference",
                sample=f"{done} of {len(pairs)} pairs, {cells:.3e} cells, {t:.1f} s"), done, mismatches


FORM_NAMES = {0: "fp32", 1: "fp64", 2: "band_fp32", 3: "band_rejected_full_fp32", 4: "band_fp64",
              5: "band_fp32_long"}


def oracle_form_errors(batch, legs, rr, hh, short_per_form: int = 600, long_per_form: int = 6):
    """Max |dlog10 L| of each result form against the FP64 full-matrix oracle
    (oracle/liboracle_pairhmm.so, bvo_log10_exact: per-row renormalised FP64
    forward, never underflows) on a seeded sample of pairs per form.  `legs`
    maps a leg name to its (log10 L, form) arrays; rr, hh give each pair's
    read and haplotype; forms are split by read length (reads of 1024 bases
    or more run FP64 first)."""
    from concurrent.futures import ThreadPoolExecutor

    lib = C.CDLL(str(ROOT / "oracle" / "liboracle_pairhmm.so"))
    lib.bvo_init()
    ex = lib.bvo_log10_exact
    ex.restype = C.c_double
    ex.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_int]
    b = batch
    rlen = np.diff(b.read_off)[rr]
    rng = np.random.default_rng(7)
    base = lambda a, o: a.ctypes.data + int(o)  # noqa: E731

    def exact(p):
        r, h = int(rr[p]), int(hh[p])
        ro, ho = int(b.read_off[r]), int(b.hap_off[h])
        return ex(int(b.read_off[r + 1]) - ro, int(b.hap_off[h + 1]) - ho,
                  base(b.read_bases, ro), base(b.read_qual, ro), base(b.read_ins, ro),
                  base(b.read_del, ro), base(b.read_gcp, ro), base(b.hap_bases, ho), 0)

    jobs = []
    for leg, (out, form) in legs.items():
        for code, name in FORM_NAMES.items():
            for long_reads in (False, True):
                sel = np.flatnonzero((form == code) & ((rlen >= LONG_ROWS) == long_reads))
                if sel.size == 0:
                    continue
                k = min(sel.size, long_per_form if long_reads else short_per_form)
                pick = np.sort(rng.choice(sel, size=k, replace=False))
                jobs.append((f"{leg}/{name}/{'long' if long_reads else 'short'}", out, pick, sel.size))
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    report = {}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as pool:
        for key, out, pick, total in jobs:
            ref = np.array(list(pool.map(exact, pick.tolist())))
            d = np.abs(out[pick] - ref)
            report[key] = dict(pairs_in_form=int(total), sampled=int(pick.size),
                               max_abs_dlog10=float(d.max()),
                               mean_abs_dlog10=float(d.mean()))
    return report, round(time.perf_counter() - t0, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps of the learned-band leg")
    ap.add_argument("--full-steps", type=int, default=2, help="timed steps of the full-matrix leg")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--regions", type=int, default=100000)
    ap.add_argument("--workload", type=int, default=4)
    ap.add_argument("--seed", type=int, default=None,
                    help="generator seed (default: the workload's own, 20261203 + workload)")
    ap.add_argument("--tune-regions", type=int, default=2000,
                    help="the autotuner times runs of a plan of this many regions of the same "
                         "workload and the full plan takes its grids")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU legs (reference timing and the FP64-oracle error per form)")
    ap.add_argument("--sw-regions", type=int, default=10000,
                    help="regions whose haplotypes the Smith-Waterman line aligns")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the FP64-oracle error per form")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--band-model", default=str(ROOT / "biovoltron_amd" / "models" / "band_mlp.npz"),
                    help="trained band model; the learned-band leg is skipped when absent")
    ap.add_argument("--no-band", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the occupancy-derived grids instead of tuning them on the plan")
    ap.add_argument("--band-margin", type=int, default=10)
    ap.add_argument("--band-max-width", type=int, default=68)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="also run a plain device-to-device copy of this size, so byte "
                         "counters can be scaled against a kernel of known traffic")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the PairHMM path has no cross-GPU step; "
              "regions are independent and would be dealt out to GPUs by cell count. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from biovoltron_amd import pairhmm, synth

    t0 = time.perf_counter()
    b = synth.generate(args.workload, seed=args.seed, count=args.regions)
    t_gen = time.perf_counter() - t0
    print(f"[bench] generated workload {args.workload}: {b.n_regions} regions, {b.n_pairs} pairs, "
          f"{b.cells:.4e} cells in {t_gen:.1f} s", file=sys.stderr, flush=True)
    t_plan0 = time.perf_counter()
    plan = pairhmm.Plan(pairhmm.PackedBatch.from_synth(b))
    t_plan = time.perf_counter() - t_plan0
    st0 = plan.stats()
    setup = dict(host_ms=round(st0.setup_ms_host, 1), upload_ms=round(st0.setup_ms_upload, 1),
                 upload_bytes=int(st0.setup_bytes),
                 upload_gbps=round(st0.setup_bytes / (st0.setup_ms_upload / 1e3) / 1e9, 2)
                 if st0.setup_ms_upload > 0 else None)
    print(f"[bench] plan ready ({t_plan:.1f} s: host {setup['host_ms']} ms, upload {setup['upload_ms']} ms, "
          f"{setup['upload_bytes'] / 1e9:.2f} GB)", file=sys.stderr, flush=True)
    use_band = not args.no_band and Path(args.band_model).exists()
    t_model = 0.0
    tuned_full = tuned = None
    if args.autotune:
        from biovoltron_amd import band
        bt = synth.generate(args.workload, seed=args.seed, count=min(args.tune_regions, args.regions))
        with pairhmm.Plan(pairhmm.PackedBatch.from_synth(bt)) as tp:
            tuned_full = tp.autotune(1)
            if use_band:
                band.set_plan_model(tp, band.BandModel.load(args.band_model), margin=args.band_margin,
                                    max_width=args.band_max_width)
                tuned = tp.autotune(1)
        plan.set_grids(tuned_full)
        print(f"[bench] autotuned grids on {bt.n_regions} regions: full matrix {tuned_full}, "
              f"learned band {tuned}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        plan.run()
        plan.results()  # waits for the stream
    # timed region: K full evaluations, synchronised on both sides
    plan.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.full_steps):
        plan.run()
    plan.synchronize()
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1e3 / args.full_steps
    out, form = plan.results()
    legs = {"full": (out, form)}
    # split of one evaluation between the passes, from HIP events on the stream
    ev_ms = plan.time(1)
    st = plan.stats()
    info = pairhmm.device_info()
    print(f"[bench] {ms_step:.2f} ms/step wall, events: total {ev_ms:.2f} ms, fp32 {st.last_ms_f32:.2f} ms, "
          f"fp64 {st.last_ms_f64:.2f} ms ({st.fp64_pairs} pairs, {st.fp64_cells:.3e} cells)",
          file=sys.stderr, flush=True)
    gcups = b.cells / (ms_step / 1e3) / 1e9
    rr, hh = b.pair_table()
    R = np.diff(b.read_off)[rr]
    Hl = np.diff(b.hap_off)[hh]
    short = R < LONG_ROWS
    cells_fp32 = int((R[short] * Hl[short]).sum())  # the FP32 full-matrix pass (short reads)
    achieved = FLOPS_PER_CELL * cells_fp32 / (st.last_ms_f32 / 1e3) / 1e12
    band_info = None
    if use_band:
        from biovoltron_amd import band
        model = band.BandModel.load(args.band_model)
        # the first model set on a plan builds its haplotype k-mer tables
        # (hap_table_kernel); part of the end-to-end time below
        t_m0 = time.perf_counter()
        band.set_plan_model(plan, model, margin=args.band_margin, max_width=args.band_max_width)
        t_model = time.perf_counter() - t_m0
        if tuned:
            plan.set_grids(tuned)
        for _ in range(args.warmup):
            plan.run()
            plan.results()
        plan.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            plan.run()
        plan.synchronize()
        t2 = time.perf_counter()
        ms_band = (t2 - t1) * 1e3 / args.steps
        t_copy0 = time.perf_counter()
        out_b, form_b = plan.results()
        t_copy = time.perf_counter() - t_copy0
        legs["learned_band"] = (out_b, form_b)
        plan.time(1)
        sb = plan.stats()
        fin = np.isfinite(out)
        err = float(np.abs(out_b[fin] - out[fin]).max())
        # band cells the banded kernels computed: R x width for every pair
        # they ran (accepted, form 2, or rejected, form 3)
        used = band.plan_bands(plan)
        ran = (form_b == 2) | (form_b == 3)
        band_cells = int((R[ran] * (used[ran, 1] - used[ran, 0] + 1)).sum())
        # long reads with a start band ran on the moving band (64 diagonals)
        w = used[:, 1] - used[:, 0] + 1
        ran_long = (R >= LONG_ROWS) & (w >= 1) & (w <= 64)
        long_cells = int(R[ran_long].sum()) * 64
        band_info = dict(
            ms_per_step=round(ms_band, 3), gcups=round(b.cells / (ms_band / 1e3) / 1e9, 3),
            speedup_vs_full=round(ms_step / ms_band, 3),
            max_abs_dlog10_vs_full=err,
            band_pairs=int(sb.band_pairs), fallback_pairs=int(sb.fallback_pairs),
            fp64_pairs=int(sb.fp64_pairs),
            cells_swept=int(sb.cells_swept), cells_swept_frac=round(sb.cells_swept / b.cells, 4),
            ms_pass1=round(sb.last_ms_f32, 3), ms_fallback=round(sb.last_ms_fallback, 3),
            ms_fp64=round(sb.last_ms_f64, 3), model=Path(args.band_model).name,
            margin=args.band_margin, max_width=args.band_max_width,
            band_kernel_ms=round(sb.last_ms_band, 3), band_cells=band_cells,
            long_pairs=int(ran_long.sum()), long_band_pairs=int(sb.long_band_pairs),
            long_band_kernel_ms=round(sb.last_ms_long, 3), long_band_cells=long_cells,
            long_band_bits=plan.long_precision(),
            grids=tuned or dict(band=None, fp32=sb.grid_f32, fp64=sb.grid_f64))
        # End to end for one evaluation from host buffers: plan set-up (host
        # work lists, input upload), the band model's set-up (haplotype k-mer
        # tables), one run, copy of the results and form bytes back.
        e2e_s = t_plan + t_model + ms_band / 1e3 + t_copy
        band_info["end_to_end"] = dict(
            plan_create_s=round(t_plan, 3), plan_create_parts=setup, band_model_setup_s=round(t_model, 3),
            run_s=round(ms_band / 1e3, 3), result_copy_s=round(t_copy, 3),
            gcups=round(b.cells / e2e_s / 1e9, 1))
        print(f"[bench] learned band: {ms_band:.2f} ms/step ({band_info['speedup_vs_full']}x), "
              f"max |dlog10| vs full {err:.2e}, fallback {sb.fallback_pairs}, "
              f"cells swept {band_info['cells_swept_frac']}, band kernels {sb.last_ms_band:.2f} ms, "
              f"long reads {int(ran_long.sum())} ({sb.long_band_pairs} accepted, "
              f"{sb.last_ms_long:.2f} ms), fp64 {sb.last_ms_f64:.2f} ms", file=sys.stderr, flush=True)
        out = out_b
    # Device post-processing of the last run (normalise, filter, marginalise
    # over a two-allele map: haplotype h of a region -> allele h % 2), timed
    # with HIP events around its kernel; bytes a caller copies back with it
    # against copying the whole matrix.
    nh = np.diff(b.region_hap_off)
    n_all = np.minimum(nh, 2).astype(np.int32)
    ha = np.concatenate([np.arange(k) % 2 for k in nh]).astype(np.int32) if b.n_regions else np.zeros(0, np.int32)
    keep, _, A = plan.postprocess(ha, n_all)
    post_ms = plan.stats().last_ms_post
    post_bytes = b.n_pairs * 8 + b.n_reads * 9 + 2 * A.size * 8  # read L, write keep + A
    post_info = dict(
        kernel_ms=round(post_ms, 3), reads=int(b.n_reads), reads_kept=int(keep.sum()),
        allele_entries=int(A.size), hbm_bytes=int(post_bytes),
        hbm_gbps=round(post_bytes / (post_ms / 1e3) / 1e9, 1) if post_ms > 0 else None,
        d2h_bytes=int(b.n_reads + A.size * 8), d2h_bytes_full_matrix=int(b.n_pairs * 9))
    # Smith-Waterman (haplotype to reference window) on pairs of the same
    # workload's first regions
    from biovoltron_amd import sw
    swp = sw_pairs(b, args.sw_regions)
    sw_info = None
    if swp:
        # two calls: the first allocates the library's Smith-Waterman
        # workspace, the second finds it in place (a caller's steady state)
        calls = []
        for _ in range(2):
            timing = {}
            t_sw = time.perf_counter()
            sw_res = sw.align_batch(swp, timing=timing)
            calls.append((time.perf_counter() - t_sw, timing["kernel_ms"], timing["call_ms"]))
        sw_cells = sum(len(r) * len(a) for r, a in swp)
        (t_first, _, c_first), (t_sw, k_ms, c_ms) = calls
        sw_info = dict(pairs=len(swp), cells=sw_cells, kernel_ms=round(k_ms, 3),
                       gcups=round(sw_cells / (k_ms / 1e3) / 1e9, 2),
                       library_call_ms=round(c_ms, 1), library_call_gcups=round(sw_cells / (c_ms / 1e3) / 1e9, 2),
                       first_library_call_ms=round(c_first, 1),
                       python_call_s=round(t_sw, 3), first_python_call_s=round(t_first, 3),
                       params="NEW_SW_PARAMETERS (200, -150, -260, -11)")
        print(f"[bench] smith-waterman: {len(swp)} pairs, {k_ms:.2f} ms kernel, {sw_info['gcups']} GCUPS; "
              f"library call {c_ms:.1f} ms ({sw_info['library_call_gcups']} GCUPS; first {c_first:.1f} ms); "
              f"Python wrapper {t_sw:.3f} s", file=sys.stderr, flush=True)
    print(f"[bench] post-processing: {post_ms:.3f} ms, {post_info['hbm_gbps']} GB/s, "
          f"copy-back {post_info['d2h_bytes'] / 1e9:.2f} GB vs {post_info['d2h_bytes_full_matrix'] / 1e9:.2f} GB",
          file=sys.stderr, flush=True)
    line = {
        "metric": METRIC,
        "value": round(band_info["gcups"] if band_info else gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(band_info["ms_per_step"] if band_info else ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded generator, biovoltron_amd/csrc/synth.cpp)",
        "config": {
            "workload": (("learned-band PairHMM (short reads: learned band, full-matrix FP32 "
                          "fallback; long reads: moving FP64 band, full-matrix FP64 fallback; "
                          "FP64 rescue)") if band_info else "batched full-matrix PairHMM")
                        + f", {WORKLOADS.get(args.workload, 'workload %d' % args.workload)}, "
                          f"{b.n_regions} regions",
            "full_matrix": {"ms_per_step": round(ms_step, 3), "gcups": round(gcups, 3),
                            "steps": args.full_steps,
                            "grids": tuned_full or dict(fp32=st.grid_f32, fp64=st.grid_f64)},
            "learned_band": band_info,
            "postprocess": post_info,
            "smith_waterman": sw_info,
            "regions": b.n_regions, "pairs": b.n_pairs, "cells": b.cells,
            "seed": args.seed if args.seed is not None else synth.workload_seed(args.workload),
            "fp64_pairs": st.fp64_pairs, "fp64_cells": st.fp64_cells,
            "device": info.arch.decode(), "compute_units": info.compute_units,
            "grid_fp32": st.grid_f32, "grid_fp64": st.grid_f64,
        },
    }
    full_roof = {
        "bound": "valu", "kernel": "pairhmm_stack_kernel (FP32 full matrix, stacked reads)",
        "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "frac_of_unpacked_peak": round(achieved / PEAK_FP32_NOPACK_TFLOPS, 4),
        "flops_per_cell": FLOPS_PER_CELL, "kernel_ms": round(st.last_ms_f32, 3), "traffic": None,
    }
    tr = profiled_traffic(f"pairhmm_workload{args.workload}.md", "full",
                          ["full-matrix FP32 pass (short reads, stacked)"])
    if tr is not None:
        full_roof["traffic"] = tr[0]
        full_roof["traffic_source"] = f"profiles/{tr[1]} (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)"
    if band_info:
        # dominant kernels of the measured leg: the banded forward (both widths)
        ach = BAND_FLOPS_PER_CELL * band_info["band_cells"] / (band_info["band_kernel_ms"] / 1e3) / 1e12
        roof = {
            "bound": "valu",
            "kernel": "band_quad_kernel<4> + band_forward_kernel<64> + band_forward_kernel<32> (all banded short-read pairs)",
            "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(ach / PEAK_FP32_TFLOPS, 4),
            "frac_of_unpacked_peak": round(ach / PEAK_FP32_NOPACK_TFLOPS, 4),
            "flops_per_cell": BAND_FLOPS_PER_CELL, "kernel_ms": band_info["band_kernel_ms"],
            "traffic": None,
            "note": "12 flops per band cell; the touch-mass pass that judges each band is not counted",
        }
        tb = profiled_traffic(f"pairhmm_workload{args.workload}.md", "learned",
                              ["band_quad_kernel<4>", "band_forward_kernel<64>", "band_forward_kernel<32>"])
        if tb is not None:
            roof["traffic"] = tb[0]
            roof["traffic_source"] = f"profiles/{tb[1]} (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)"
        line["roofline"] = roof
        line["config"]["full_matrix"]["roofline"] = full_roof
        if band_info["long_band_kernel_ms"] > 0:
            achl = (BAND_FLOPS_PER_CELL * band_info["long_band_cells"]
                    / (band_info["long_band_kernel_ms"] / 1e3) / 1e12)
            fp32_long = band_info["long_band_bits"] == 32
            peak = PEAK_FP32_TFLOPS if fp32_long else PEAK_FP64_TFLOPS
            band_info["long_roofline"] = {
                "bound": "valu",
                "kernel": ("band_long_quad_kernel + band_long_kernel<float> (four pairs of a read per "
                           "wavefront, then the rest one per wavefront)" if fp32_long
                           else "band_long_kernel<double>"),
                "achieved": round(achl, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(achl / peak, 4),
                "flops_per_cell": BAND_FLOPS_PER_CELL, "kernel_ms": band_info["long_band_kernel_ms"],
                "note": "shares the GPU with the short-read passes (second stream)"}
    else:
        line["roofline"] = full_roof
    if not args.no_cpu:
        cb, worst = cpu_baseline(b, out, args.cpu_seconds)
        line["cpu_baseline"] = cb
        line["config"]["max_abs_dlog10_vs_cpu_sample"] = worst
        if sw_info is not None:
            r = sw_cpu_baseline(swp, sw_res, 3.0)
            if r is not None:
                sw_info["cpu_baseline"], n_chk, n_bad = r
                sw_info["checked_vs_reference"] = dict(pairs=n_chk, different=n_bad)
        if not args.no_oracle:
            rep, secs = oracle_form_errors(b, legs, rr, hh)
            line["config"]["vs_fp64_oracle"] = dict(
                sample="seeded per form: up to 600 short-read and 6 long-read pairs per form and leg",
                seconds=secs, forms=rep)
            print(f"[bench] vs FP64 oracle ({secs} s): " + ", ".join(
                f"{k} {v['max_abs_dlog10']:.2e}" for k, v in rep.items()), file=sys.stderr, flush=True)
    plan.close()
    if args.copy_calibration > 0:
        # A device-to-device copy of known size through the HIP runtime's own
        # blit kernel; its FETCH_SIZE / WRITE_SIZE scale the byte counters.
        hip = C.CDLL("libamdhip64.so")
        n = args.copy_calibration * (1 << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
        assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
        assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        line["config"]["copy_calibration_bytes"] = n
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
