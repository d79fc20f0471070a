#!/usr/bin/env python
"""Cone-beam projector-pair benchmark on one MI355X.

Workload: 512^3 FP32 volume (0.5 mm voxels) from a seeded 30-ellipsoid
phantom, 360 views over 2*pi, 512 x 512 detector (0.8 mm), DSO 1000 mm,
DSD 1536 mm.  One step = Ax('Siddon') + Atb('FDK') over all 360 views on
device-resident data (the operator pair OS-SART and ASD-POCS use).

Metric: voxel-view updates per second of the pair, 2 * nVoxel * nViews per
step (Gupdates/s).  The roofline entry prices the dominant kernel against
HBM with the per-view sweep byte model: Ax moves n (V + P) * 4 bytes, Atb
n (2V + P) * 4 bytes.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline's OpenMP threads stay on neighbouring cores; set before any
# OpenMP runtime (torch's or the oracle's) reads its environment.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# The CPUs this process may use, read before any OpenMP runtime starts: with a
# bound OpenMP, the first runtime to start pins the main thread to one core, and
# the affinity read afterwards is that one core.
HOST_CPUS = len(os.sched_getaffinity(0))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes
# (scaled by a copy-kernel calibration); parsed from this summary when present.
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_512cube_360views.txt")
# VALU / LDS instruction counts per launch from a rocprofv3 SQ_* counter pass; parsed when present.
VALU_SUMMARY = os.path.join(ROOT, "profiles", "valu_lds_512cube_360views.txt")
PROBE_BLOCKS, PROBE_ITERS = 8192, 4096
# VALU instructions per wavefront of the probe (its ISA, and SQ_INSTS_VALU exactly): mode 0, eight
# FP32 FMA chains, 8 per loop round plus 26 outside the loop; mode 3, four FMA chains interleaved with
# four integer adds, 8 per round plus 25 outside
PROBE_VALU_PER_WAVE = {0: 8 * PROBE_ITERS + 26, 3: 8 * PROBE_ITERS + 25}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 2 and cells[0] == kernel:
                    return float(cells[-1])
    except (OSError, ValueError):
        return None
    return None


def measured_valu(kernel):
    """(VALU, LDS, LDS bank-conflict cycles) per launch of `kernel` from the committed counter
    summary (rows '| key | kernel name | SQ_INSTS_VALU | SQ_INSTS_SALU | SQ_INSTS_LDS |
    SQ_LDS_BANK_CONFLICT | ...'), or None."""
    try:
        with open(VALU_SUMMARY) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 6 and cells[0] == kernel:
                    return float(cells[2]), float(cells[4]), float(cells[5])
    except (OSError, ValueError):
        return None
    return None


def valu_issue_rate(lib, dev):
    """Measured VALU issue rates (wavefront instructions per second) of the
    probe kernel: PROBE_BLOCKS x 4 wavefronts, PROBE_VALU_PER_WAVE each, timed
    by the library's HIP events on the launch stream (median of 5), for a pure
    FP32-FMA stream (mode 0) and for FMAs interleaved with integer adds (mode
    3, the higher of the two: a SIMD issues the mix faster than FMAs alone)."""
    import ctypes

    import numpy as np
    import torch

    from tigre_amd import _native
    from tigre_amd.operators import _stream_handle

    out = torch.empty(PROBE_BLOCKS * 256, device=dev)
    total, count = ctypes.c_double(0.0), ctypes.c_int(0)
    waves = PROBE_BLOCKS * 4
    res = {"waves": waves, "valu_per_wave": {"fma": PROBE_VALU_PER_WAVE[0], "mixed": PROBE_VALU_PER_WAVE[3]}}
    for mode, key in ((0, "fma"), (3, "mixed")):
        _native.check(lib.tigre_valu_probe(ctypes.c_void_p(out.data_ptr()), PROBE_BLOCKS, PROBE_ITERS, mode,
                                           _stream_handle()), "probe")
        torch.cuda.synchronize()
        lib.tigre_kernel_timing(1)
        ms = []
        for _ in range(5):
            _native.check(lib.tigre_valu_probe(ctypes.c_void_p(out.data_ptr()), PROBE_BLOCKS, PROBE_ITERS, mode,
                                               _stream_handle()), "probe")
            _native.check(lib.tigre_kernel_time(_native.KERNEL_PROBE, ctypes.byref(total), ctypes.byref(count)),
                          "kernel_time")
            ms.append(total.value / count.value)
        lib.tigre_kernel_timing(0)
        t = float(np.median(ms))
        res[key + "_ms"] = round(t, 4)
        res[key + "_wave_instr_per_s"] = waves * PROBE_VALU_PER_WAVE[mode] / (t * 1e-3)
    res["wave_instr_per_s"] = max(res["fma_wave_instr_per_s"], res["mixed_wave_instr_per_s"])
    return res


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on, capped by
    OMP_NUM_THREADS when set (the share of a shared host allotted to the job)."""
    n = HOST_CPUS
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, cap) if cap > 0 else n


def host_description():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": HOST_CPUS,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="volume edge / detector edge")
    ap.add_argument("--views", type=int, default=360)
    ap.add_argument("--cpu-views", type=int, default=24,
                    help="views in the timed CPU-baseline sample, indices floor(k * views / cpu_views) (0 = skip)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU-baseline runs (the median is reported)")
    ap.add_argument("--parity-views", type=int, default=32,
                    help="views of the full-size parity check against the oracle (0 = skip)")
    ap.add_argument("--no-extra", action="store_true", help="skip the interpolated / matched timings")
    ap.add_argument("--ossart-iters", type=int, default=20,
                    help="pair: OS-SART iterations (20-view subsets, lambda 1) after the timed region (0 = skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the timed region, copy a buffer of this size once (counter calibration)")
    ap.add_argument("--calib-valu", action="store_true",
                    help="after the timed region, run the SART residual kernel once with a known VALU count "
                         "(8192 workgroups x 4 wavefronts, 16 loop iterations per lane: 216 VALU per wavefront)")
    ap.add_argument("--workload", choices=["pair", "small", "batched", "learned", "precision", "tv"], default="pair",
                    help="pair: Ax+Atb at 512^3 (default); batched: 64 uneven OS-SART cases in one graph; "
                         "learned: ASD-POCS with the learned proximal step vs plain TV at 256^3; "
                         "precision: BF16 / FP16 / FP32 storage at 1024^3 with CGLS; "
                         "small: every operator at 128^3 beside the full FP64 oracle")
    ap.add_argument("--train-steps", type=int, default=12000, help="learned: Adam steps")
    ap.add_argument("--learned-cases", type=int, default=8, help="learned: training phantoms")
    ap.add_argument("--learned-first", type=int, default=3,
                    help="learned: pre-TV images of the first iterations paired with the final image")
    ap.add_argument("--learned-crops", type=int, default=6000, help="learned: 32^3 training crops")
    ap.add_argument("--learned-attempts", type=int, default=3, help="learned: calls that try the network")
    ap.add_argument("--cgls-iters", type=int, default=30, help="precision: CGLS iterations per storage type")
    ap.add_argument("--oracle-views", type=int, default=8, help="precision: views checked against the FP64 oracle")
    ap.add_argument("--no-workloads", action="store_true",
                    help="pair: skip the compact 1024^3 precision, 64-case batched and 256^3 learned summaries")
    ap.add_argument("--cases", type=int, default=64, help="batched: number of cases")
    ap.add_argument("--iters", type=int, default=10, help="batched: OS-SART iterations per case")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload in ("batched", "learned", "precision"):
        fn = {"batched": bench_batched, "learned": bench_learned, "precision": bench_precision}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    if args.workload == "small":
        bench_small(args)
        return
    if args.workload == "tv":
        bench_tv(args)
        return

    import numpy as np
    import torch

    import ctypes

    import tigre_amd
    from tigre_amd import _native
    from tigre_amd.geometry import cone_geometry, prepared
    from tigre_amd.operators import _stream_handle
    from tigre_amd.phantom import ellipsoid_phantom_torch

    dev = torch.device("cuda:0")
    n, nviews = args.n, args.views
    geo = cone_geometry((n, n, n), (256.0, 256.0, 256.0), (n, n), (0.8 * 512 / n, 0.8 * 512 / n), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    log("phantom %d^3 ..." % n)
    img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 30, seed=20261202 + 2, device=dev)
    torch.cuda.synchronize()
    V = n ** 3
    P = n * n

    def ax():
        return tigre_amd.Ax(img, geo, angles, "Siddon")

    proj = ax()

    def atb():
        return tigre_amd.Atb(proj, geo, angles, "FDK")

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ax()
        atb()
    torch.cuda.synchronize()

    lib = _native.lib()

    def kernel_ms(kernel_id):
        """Average duration of the launches of one kernel recorded since the last call."""
        total, count = ctypes.c_double(0.0), ctypes.c_int(0)
        _native.check(lib.tigre_kernel_time(kernel_id, ctypes.byref(total), ctypes.byref(count)), "kernel_time")
        return total.value / count.value if count.value else float("nan")

    for k in range(4):
        kernel_ms(k)
    lib.tigre_kernel_timing(1)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        ax()
        ev[k][1].record(stream)
        atb()
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ax_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    atb_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    k_siddon = kernel_ms(0)
    k_fdk = kernel_ms(2)
    ms_per_step = wall * 1e3 / args.steps
    updates = 2.0 * V * nviews
    value = updates * args.steps / wall / 1e9
    log("Ax Siddon %.2f ms (kernel %.2f)  Atb FDK %.2f ms (kernel %.2f)  step %.2f ms" % (
        ax_ms, k_siddon, atb_ms, k_fdk, ms_per_step))

    # the other two operators of the pair, timed the same way
    def timed(fn, reps=3):
        fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    if args.no_extra:
        interp_ms = matched_ms = k_interp = k_matched = float("nan")
    else:
        interp_ms = timed(lambda: tigre_amd.Ax(img, geo, angles, "interpolated"))
        matched_ms = timed(lambda: tigre_amd.Atb(proj, geo, angles, "matched"))
        k_interp = kernel_ms(1)
        k_matched = kernel_ms(3)
    lib.tigre_kernel_timing(0)
    kernel_ms(4)
    log("Ax interpolated %.2f ms (kernel %.2f)  Atb matched %.2f ms (kernel %.2f)" % (
        interp_ms, k_interp, matched_ms, k_matched))

    ossart = None
    if args.ossart_iters > 0:
        # the reconstruction loop over the same data: OS-SART, 20-view subsets, lambda 1, from zero;
        # the wall-clock includes its set-up (subset plans, W and V weights)
        from tigre_amd import algorithms as algs

        torch.cuda.synchronize()
        t = time.perf_counter()
        rec = algs.ossart(proj, geo, angles, args.ossart_iters, blocksize=20, lmbda=1.0, return_tensor=True)
        torch.cuda.synchronize()
        t_os = time.perf_counter() - t
        rmse = float(torch.sqrt(torch.mean((rec.double() - img.double()) ** 2)))
        ossart = {"iterations": args.ossart_iters, "blocksize": 20, "lambda": 1.0, "seconds": round(t_os, 3),
                  "ms_per_iteration": round(t_os * 1e3 / args.ossart_iters, 2),
                  "gupdates_per_s": round(2.0 * V * nviews * args.ossart_iters / t_os / 1e9, 1),
                  "rmse_vs_phantom": rmse}
        log("OS-SART %d iterations: %.3f s, image RMSE %.4g" % (args.ossart_iters, t_os, rmse))
        del rec

    if args.calib_copy_mib > 0:
        # one streaming elementwise kernel of known size (reads and writes calib_copy_mib MiB each),
        # for scaling the FETCH_SIZE / WRITE_SIZE counters
        n_el = args.calib_copy_mib * 262144
        a = torch.ones(n_el, device=dev)
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB" % args.calib_copy_mib)
        del a, b

    if args.calib_valu:
        # r = W (p - r) over 8192 x 256 x 4 x 16 elements: every lane runs 16 iterations of the
        # 13-instruction loop of sart_residual_kernel, plus 8 instructions outside it
        n_el = 8192 * 256 * 4 * 16
        a = torch.zeros(n_el, device=dev)
        b = torch.ones(n_el, device=dev)
        c = torch.ones(n_el, device=dev)
        _native.check(lib.tigre_sart_residual(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                              ctypes.c_void_p(c.data_ptr()), n_el, _stream_handle()), "calibration")
        torch.cuda.synchronize()
        log("VALU calibration: 32768 wavefronts x 216 VALU instructions = 7077888 expected")
        del a, b, c

    probe = valu_issue_rate(lib, dev)
    log("VALU probe: %.4g (FMA) / %.4g (FMA + integer) wavefront instructions/s" % (
        probe["fma_wave_instr_per_s"], probe["mixed_wave_instr_per_s"]))

    b_ax = nviews * (V + P) * 4.0
    b_atb = nviews * (2 * V + P) * 4.0
    if k_fdk >= k_siddon:
        dom, dom_ms, dom_b = "atb_fdk", k_fdk, b_atb
    else:
        dom, dom_ms, dom_b = "ax_siddon", k_siddon, b_ax
    achieved = dom_b / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom)

    result = {
        "metric": "Ax+Atb Gupdates/s (Siddon + FDK, cone beam)",
        "value": round(value, 3),
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 30-ellipsoid phantom)",
        "config": {"workload": "cone-beam Ax Siddon + Atb FDK, %d^3 volume, %d views, %dx%d detector" % (
            n, nviews, n, n), "nVoxel": [n, n, n], "nViews": nviews, "nDetector": [n, n],
            "DSO_mm": 1000.0, "DSD_mm": 1536.0},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "model_bytes_per_launch": dom_b, "kernel_ms": round(dom_ms, 3)},
        "ops_ms": {"ax_siddon": round(ax_ms, 3), "atb_fdk": round(atb_ms, 3), "ax_interpolated": round(interp_ms, 3),
                   "atb_matched": round(matched_ms, 3)},
        "kernel_ms": {"ax_siddon": round(k_siddon, 3), "atb_fdk": round(k_fdk, 3),
                      "ax_interpolated": round(k_interp, 3), "atb_matched": round(k_matched, 3)},
    }
    # The same four kernels against the VALU issue rate the probe measured: instruction counts per
    # launch from the committed counter summary, durations from this run.
    valu = {}
    for name, ms in (("ax_siddon", k_siddon), ("ax_interpolated", k_interp), ("atb_fdk", k_fdk),
                     ("atb_matched", k_matched)):
        counts = measured_valu(name) if n == 512 and nviews == 360 else None
        if counts is None or not ms == ms:
            continue
        rate = counts[0] / (ms * 1e-3)
        valu[name] = {"valu_instr_per_launch": counts[0], "kernel_ms": round(ms, 3),
                      "achieved": round(rate, 4), "frac": round(rate / probe["wave_instr_per_s"], 4),
                      "lds_instr_per_launch": counts[1], "lds_bank_conflict_cycles": counts[2]}
    result["valu_roofline"] = {"bound": "valu", "unit": "wavefront instructions/s",
                               "peak": round(probe["wave_instr_per_s"], 1), "probe": probe, "kernels": valu}
    if ossart is not None:
        result["ossart"] = ossart

    if args.cpu_views > 0 or args.parity_views > 0:
        import oracle

        threads = cpu_threads()
        oracle.set_threads(threads)
        img_h = img.cpu().numpy()
    if args.parity_views > 0:
        log("parity: FP64 oracle on %d views ..." % args.parity_views)
        # full-size parity of the 360-view calls on a fixed view subset
        idx = np.array([k * nviews // args.parity_views for k in range(args.parity_views)])
        gp = prepared(geo, angles[idx])
        ref_p = oracle.ax(img_h, gp, "Siddon")
        proj_sub = proj[torch.as_tensor(idx, device=dev)].contiguous()
        ref_v = oracle.atb(proj_sub.cpu().numpy(), gp, "FDK")
        got_v = tigre_amd.Atb(proj_sub, geo, angles[idx], "FDK").cpu().numpy()
        result["parity_full_size"] = {
            "ax_siddon_rel_l2": oracle.rel_l2(proj_sub.cpu().numpy(), ref_p),
            "atb_fdk_rel_l2": oracle.rel_l2(got_v, ref_v), "views": [int(i) for i in idx]}
        log("parity on %d views: Ax %.2e, Atb %.2e" % (len(idx), result["parity_full_size"]["ax_siddon_rel_l2"],
                                                       result["parity_full_size"]["atb_fdk_rel_l2"]))
        del ref_p, ref_v, got_v
    if args.cpu_views > 0:
        # the FP64 OpenMP oracle on the host: Ax Siddon + Atb FDK over a view sample, median of runs
        idx = np.array([k * nviews // args.cpu_views for k in range(args.cpu_views)])
        gp = prepared(geo, angles[idx])
        proj_h = proj[torch.as_tensor(idx, device=dev)].contiguous().cpu().numpy()
        runs = []
        for r in range(max(1, args.cpu_runs)):
            t = time.perf_counter()
            oracle.ax(img_h, gp, "Siddon")
            oracle.atb(proj_h, gp, "FDK")
            runs.append(time.perf_counter() - t)
            log("cpu baseline run %d: %.1f s on %d threads" % (r + 1, runs[-1], threads))
        t_med = float(np.median(runs))
        cpu_rate = 2.0 * V * len(idx) / t_med / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_rate, 4), "unit": "Gupdates/s", "cores": oracle.get_threads(), "kind": "port",
            "sample": "FP64 OpenMP oracle: Ax Siddon + Atb FDK over %d of the %d views (views floor(k*%d/%d)), "
                      "median of %d runs (%s s)" % (len(idx), nviews, nviews, len(idx), len(runs),
                                                   ", ".join("%.1f" % r for r in runs)),
            "host": host_description()}
        log("cpu baseline %.4f Gupdates/s on %d threads" % (cpu_rate, oracle.get_threads()))
    if not args.no_workloads:
        # the other benchmark workloads, compact: run after the timed region, each summarised in the line
        del img, proj
        torch.cuda.empty_cache()
        result["workloads"] = run_workloads(args)
    print(json.dumps(result), flush=True)


def _compact(d, drop):
    return {k: v for k, v in d.items() if k not in drop}


def run_workloads(args):
    """Compact runs of the precision (1024^3, BF16/FP16/FP32 storage, CGLS with
    10 iterations, 2 oracle views), batched (64 cases, 2 graph replays) and
    learned (256^3 ASD-POCS) workloads; each entry is that workload's JSON
    without its per-iteration curves, or the error it raised."""
    import copy

    import torch

    out = {}
    for name, fn, over, drop in (
            ("precision_1024", bench_precision, {"cgls_iters": 10, "oracle_views": 2}, ()),
            ("batched_64", bench_batched, {"steps": 2}, ("tuned",)),
            ("learned_256", bench_learned, {}, ("runs",))):
        a = copy.copy(args)
        for k, v in over.items():
            setattr(a, k, v)
        t = time.perf_counter()
        try:
            r = fn(a)
            if "config" in r and isinstance(r["config"], dict):
                r["config"] = _compact(r["config"], ("cases", "oracle_views"))
            out[name] = _compact(r, drop + ("n_gpus", "scaling", "vs_baseline"))
        except Exception as e:  # a failed summary must not cost the main measurement
            out[name] = {"error": repr(e)}
        out[name]["wall_s"] = round(time.perf_counter() - t, 1)
        log("%s: %.1f s" % (name, out[name]["wall_s"]))
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return out


def bench_batched(args):
    """64 independent OS-SART reconstructions of uneven sizes: edge in {64, 96,
    128, 192, 256} (1 mm voxels), 60-360 views over 2*pi, edge x edge detector
    at 1.6 mm, DSO U[800, 1200] mm, DSD 1.536 DSO, 10-ellipsoid phantoms, 10
    iterations with 20-view subsets, lambda 1.  Times the same work three
    ways: 64 calls of algorithms.ossart (per-call setup and launches), the
    prepared plans launched case by case from the host, and one replay of the
    graph holding every case's iterations.  Setup (weights, plans, autotuning,
    capture) is outside the timed regions; projections come from Ax of the
    phantom."""
    import numpy as np
    import torch

    import tigre_amd
    from tigre_amd import algorithms as algs
    from tigre_amd.algorithms.batched import BatchedOSSART
    from tigre_amd.geometry import cone_geometry
    from tigre_amd.phantom import ellipsoid_phantom_torch
    from tigre_amd.plan import Autotuner

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(20261202 + 4)
    cases, desc, updates = [], [], 0.0
    for k in range(args.cases):
        edge = int(rng.choice([64, 96, 128, 192, 256]))
        nviews = int(rng.integers(60, 361))
        dso = float(rng.uniform(800.0, 1200.0))
        geo = cone_geometry((edge,) * 3, (float(edge),) * 3, (edge, edge), (1.6, 1.6), dso, 1.536 * dso)
        angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
        img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 10, seed=20261202 + 4 + 1000 * k, device=dev)
        proj = tigre_amd.Ax(img, geo, angles, "Siddon")
        cases.append((proj, geo, angles))
        desc.append((edge, nviews))
        updates += 2.0 * edge ** 3 * nviews * args.iters
    torch.cuda.synchronize()
    log("batched: %d cases, %.1f Gupdates per run" % (len(cases), updates / 1e9))

    t = time.perf_counter()
    tuner = Autotuner(reps=3)
    bo = BatchedOSSART(cases, niter=args.iters, blocksize=20, tuner=tuner)
    bo.run()  # builds the wave tables, loads the kernels
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    log("setup + autotuning %.2f s; tuned %s" % (t_setup, tuner.table))

    def clock(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # 1) the public API, one call per case (default plan parameters)
    t_api = clock(lambda: [algs.ossart(p, g, a, args.iters, blocksize=20, return_tensor=True) for p, g, a in cases])
    log("64 x ossart(): %.3f s" % t_api)
    # 2) the same plans as the batches, host-driven, case after case
    bo.reset()
    t_eager = clock(bo.run_sequential)
    ref = [r.clone() for r in bo.results()]
    log("plans, case after case: %.3f s" % t_eager)
    # 3) wave batches: one launch per stage over every case of a wave
    t_waves = float(np.median([clock(bo.run) for _ in range(max(1, args.steps))]))
    same_waves = all(torch.equal(a, b) for a, b in zip(bo.results(), ref))
    log("wave batches: median %.3f s (%d waves), bit-identical: %s" % (t_waves, len(bo.waves), same_waves))
    # 4) the waves captured in one graph
    t = time.perf_counter()
    bo.capture()
    t_capture = time.perf_counter() - t
    reps = [clock(bo.replay) for _ in range(max(1, args.steps))]
    t_graph = float(np.median(reps))
    same = all(torch.equal(a, b) for a, b in zip(bo.results(), ref))
    log("graph replay: median %.3f s over %d (capture %.2f s), bit-identical: %s" % (t_graph, len(reps), t_capture,
                                                                                 same))
    del ref
    result = {
        "metric": "batched OS-SART wall-clock, %d cases x %d iterations" % (len(cases), args.iters),
        "value": round(t_graph, 4), "unit": "s", "n_gpus": 1, "steps": len(reps), "warmup": 1,
        "ms_per_step": round(t_graph * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded 10-ellipsoid phantoms, projections from Ax)",
        "config": {"workload": "64 uneven OS-SART cases (edge 64-256, 60-360 views, DSO 800-1200 mm), "
                               "10 iterations, 20-view subsets, wave batches in one graph replay",
                   "cases": [list(d) for d in desc], "waves": len(bo.waves)},
        "gupdates_per_s": round(updates / t_graph / 1e9, 2),
        "wave_batches_s": round(t_waves, 4), "sequential_api_s": round(t_api, 4),
        "sequential_plans_s": round(t_eager, 4), "speedup_vs_api": round(t_api / t_graph, 2),
        "speedup_vs_sequential_plans": round(t_eager / t_graph, 2),
        "bit_identical_to_sequential_plans": bool(same and same_waves), "setup_s": round(t_setup, 2),
        "capture_s": round(t_capture, 2), "tuned": {k: list(v) for k, v in tuner.table.items()},
    }
    bo.close()
    return result


def bench_learned(args):
    """ASD-POCS with the learned proximal step against plain TV.

    256^3 volume (1 mm voxels), 180 views over 2*pi, 256 x 256 detector at
    1.6 mm, DSO 1000 / DSD 1536 mm; seeded 40-ellipsoid phantoms scaled to
    attenuation 0.02/mm per unit density; Poisson noise at I0 = 1e5 photons;
    ASD-POCS with tviter 20, alpha 0.002.  Training data: plain ASD-POCS runs
    of up to 40 iterations on ``learned_cases`` phantoms; pairs (pre-TV image
    of iteration k, the run's final post-TV image) for the first
    ``learned_first`` iterations k; random 32^3 crops (mostly inside the
    object), Adam on the library's MFMA convolution kernels.  Evaluation on a
    held-out phantom: image RMSE after every iteration of plain ASD-POCS, of
    ASD-POCS with the learned step (plain TV judging each learned step by the
    data constraint and TV objective, at most ``learned_attempts`` attempts,
    ended by the first fallback), and of the loop with no regularization
    step; time to reach the plain run's final RMSE."""
    import numpy as np
    import torch

    import tigre_amd
    from tigre_amd import algorithms as algs
    from tigre_amd import learned as L
    from tigre_amd.geometry import cone_geometry
    from tigre_amd.phantom import ellipsoid_phantom_torch

    n, nviews, niter = 256, 180, 20
    geo = cone_geometry((n, n, n), (float(n),) * 3, (n, n), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    seed0 = 20261202 + 5

    def case(k):
        img = 0.02 * ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 40, seed=seed0 + 100 * k, device="cuda")
        proj = tigre_amd.Ax(img, geo, angles, "Siddon")
        return img, L.poisson_sinogram(proj, 1e5, seed=seed0 + 100 * k + 1)

    t = time.perf_counter()
    pairs = []
    for k in range(1, args.learned_cases + 1):
        pre, post = L.collect_trajectory(case(k)[1], geo, angles, 40)
        pairs += [(x, post[-1], 1.0) for x in pre[:args.learned_first]]
        del pre, post
    x, target = L.make_patches(pairs, args.learned_crops, 32, seed=seed0, object_level=2e-3)
    del pairs
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t
    log("pairs: %d crops in %.1f s" % (x.shape[0], t_pairs))
    net = L.ResidualCNN(seed=seed0, data_scale=50.0)
    t = time.perf_counter()
    losses = L.train(net, x, target, args.train_steps, batch=8, lr=3e-4, seed=seed0, lr_final=3e-5,
                     log=lambda *a: None)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    log("training: %d steps in %.1f s, loss %.3g -> %.3g" % (args.train_steps, t_train, float(np.mean(losses[:200])),
                                                            float(np.mean(losses[-200:]))))
    del x, target

    img, b = case(0)

    def rmse(a):
        return float(torch.sqrt(torch.mean((a - img).double() ** 2)))

    resid = L.data_residual(b, geo, angles)
    eps = L.data_epsilon(b, geo, angles, resid)
    runs = {}
    for name in ("plain", "learned", "no_step"):
        if name == "plain":
            reg = None
        elif name == "learned":
            reg = L.LearnedProx(net, 20, residual=resid, judge="tv", epsilon=eps, attempts=args.learned_attempts,
                                latch=True)
        else:
            reg = lambda z, s: z  # noqa: E731
        curve = []
        res, stats = algs.asd_pocs(b, geo, angles, niter, regularizer=reg, return_tensor=True, return_stats=True,
                                   callback=lambda k, r: curve.append(rmse(r)))
        times = [float(v) for v in np.cumsum([s_[5] for s_ in stats])]
        runs[name] = {"rmse": [round(c, 7) for c in curve], "seconds": [round(v, 4) for v in times]}
        if name == "learned":
            runs[name]["taken"], runs[name]["fallbacks"] = reg.taken, reg.fallbacks
        log("%s: final RMSE %.6f after %d iterations, %.2f s" % (name, curve[-1], len(curve), times[-1]))
    goal = runs["plain"]["rmse"][-1]
    t_plain = runs["plain"]["seconds"][-1]
    reach = [i for i, c in enumerate(runs["learned"]["rmse"]) if c <= goal]
    t_learned = runs["learned"]["seconds"][reach[0]] if reach else None

    # one learned step vs one TV step at this size, and the convolution kernel rate
    x0 = img.contiguous()

    def clock(fn, reps=5):
        fn()
        s_, e_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s_.record()
        for _ in range(reps):
            fn()
        e_.record()
        e_.synchronize()
        return s_.elapsed_time(e_) / reps

    ms_cnn = clock(lambda: net.forward(x0, scale=1e-3))
    ms_tv = clock(lambda: tigre_amd.minTV(x0, 1e-3, 20))
    V = float(n) ** 3
    flops_model = 2.0 * 27 * (32 + 4 * 32 * 32 + 32) * V  # useful multiply-adds of the 6 layers
    # MFMA work issued: four 32 -> 32 layers in full, the first layer with its 27 taps packed into
    # k = 32, the last (32 -> 1) with the taps as the 32 GEMM columns
    flops_exec = (4 * 2.0 * 27 * 32 * 32 + 2.0 * 32 * 32 + 2.0 * 32 * 32) * V
    result = {
        "metric": "ASD-POCS wall-clock to the plain-TV RMSE, learned proximal step vs TV",
        "value": round(t_plain / t_learned, 3) if t_learned else 0.0, "unit": "x (speedup; 0 = RMSE not reached)",
        "n_gpus": 1, "steps": niter, "warmup": 0, "ms_per_step": round(1e3 * t_plain / len(runs["plain"]["rmse"]), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (CNN), f32 (projectors)",
        "data": "synthetic (seeded 40-ellipsoid phantoms, Poisson noise I0 1e5)",
        "config": {"workload": "ASD-POCS 256^3, 180 views, 256x256 detector, tviter 20, alpha 0.002; 6-layer "
                               "32-channel residual CNN", "train_cases": args.learned_cases,
                   "train_iterates": args.learned_first, "train_crops": args.learned_crops, "crop": 32,
                   "train_steps": args.train_steps, "attempts": args.learned_attempts},
        "plain_final_rmse": goal, "plain_seconds": round(t_plain, 3),
        "learned_seconds_to_plain_rmse": round(t_learned, 3) if t_learned else None,
        "learned_iterations_to_plain_rmse": reach[0] + 1 if reach else None,
        "runs": runs, "train": {"pairs_s": round(t_pairs, 2), "train_s": round(t_train, 2),
                                "loss_first": float(np.mean(losses[:200])), "loss_last": float(np.mean(losses[-200:]))},
        "step_ms": {"cnn_256cube": round(ms_cnn, 3), "minTV_20iter_256cube": round(ms_tv, 3)},
        "roofline": {"bound": "mfma", "kernel": "conv_zm_kernel + conv_taps_kernel + conv_out_kernel (6 layers, "
                                                "useful FLOPs)",
                     "achieved": round(flops_model / ms_cnn / 1e9, 1),
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": round(flops_model / ms_cnn / 1e9 / 2500.0, 4),
                     "traffic": None, "model_tflop": round(flops_model / 1e12, 3),
                     "issued_tflop": round(flops_exec / 1e12, 3)},
    }
    return result


def bench_precision(args):
    """Storage precision at scale: 1024^3 volume (sVoxel 256 mm, 0.25 mm voxels),
    720 views over 2*pi, 1024 x 768 detector at 0.4 mm, DSO 1000 / DSD 1536 mm,
    per-view detector offsets drawn from U[-2, 2] pixels on both axes (seeded);
    seeded 30-ellipsoid phantom, projections from Ax.  For FP32, BF16 and FP16
    storage (FP32 arithmetic): Ax Siddon and Atb matched timed and compared with
    the FP32 result and, on a view subset, with the FP64 oracle; then CGLS
    (FP64 inner products; 30 iterations by default) with narrow working vectors
    and a float32 iterate, and with the iterate stored narrow as well, the final
    iterate compared with the FP32 run's and with the phantom."""
    import numpy as np
    import torch

    import oracle
    import tigre_amd
    from tigre_amd import algorithms as algs
    from tigre_amd.geometry import cone_geometry, prepared
    from tigre_amd.phantom import ellipsoid_phantom_torch

    nx, nviews, niter = 1024, 720, args.cgls_iters
    geo = cone_geometry((nx, nx, nx), (256.0,) * 3, (768, 1024), (0.4, 0.4), 1000.0, 1536.0)
    rng = np.random.default_rng(20261202 + 3)
    geo.offDetector = rng.uniform(-2.0, 2.0, (nviews, 2)) * np.asarray(geo.dDetector)
    angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 30, seed=20261202 + 3, device="cuda")
    proj = tigre_amd.Ax(img, geo, angles, "Siddon")
    torch.cuda.synchronize()
    log("precision: phantom and data ready")

    def rel(a, b):
        a, b = a.double(), b.double()
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    def clock(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    sub = np.array([k * nviews // args.oracle_views for k in range(args.oracle_views)])
    oracle.set_threads(cpu_threads())
    img_h = img.cpu().numpy()
    gp = prepared(copy_geo(geo, sub), angles[sub])
    t = time.perf_counter()
    ref_ax = oracle.ax(img_h, gp, "Siddon")
    ref_atb = oracle.atb(proj[torch.as_tensor(sub, device="cuda")].cpu().numpy(), gp, "matched")
    t_oracle = time.perf_counter() - t
    log("oracle on %d views: %.1f s" % (len(sub), t_oracle))
    res = {}
    x32 = None
    for name, dt in (("f32", torch.float32), ("bf16", torch.bfloat16), ("f16", torch.float16)):
        imgd, projd = img.to(dt), proj.to(dt)
        ax, t_ax = clock(lambda: tigre_amd.Ax(imgd, geo, angles, "Siddon"))
        atb, t_atb = clock(lambda: tigre_amd.Atb(projd, geo, angles, "matched"))
        idx = torch.as_tensor(sub, device="cuda")
        subgeo = copy_geo(geo, sub)
        atb_sub = tigre_amd.Atb(projd[idx].contiguous(), subgeo, angles[sub], "matched")
        entry = {"ax_ms": round(t_ax * 1e3, 2), "atb_matched_ms": round(t_atb * 1e3, 2),
                 "ax_rel_l2_vs_f32": rel(ax.float(), proj), "ax_rel_l2_vs_fp64":
                     oracle.rel_l2(ax[idx].float().cpu().numpy(), ref_ax),
                 "atb_rel_l2_vs_fp64": oracle.rel_l2(atb_sub.float().cpu().numpy(), ref_atb)}
        del ax, atb, atb_sub
        torch.cuda.synchronize()
        t = time.perf_counter()
        x = algs.cgls(proj, geo, angles, niter, storage=dt, return_tensor=True)
        torch.cuda.synchronize()
        entry["cgls_s"] = round(time.perf_counter() - t, 2)
        xf = x.float()
        if name == "f32":
            x32 = xf.clone()
        entry["cgls_rel_l2_vs_f32_run"] = rel(xf, x32)
        entry["cgls_rel_l2_vs_phantom"] = rel(xf, img)
        if name != "f32":
            # the iterate stored narrow as well (every update rounded)
            del x, xf
            t = time.perf_counter()
            x = algs.cgls(proj, geo, angles, niter, storage=dt, narrow_iterate=True, return_tensor=True)
            torch.cuda.synchronize()
            xf = x.float()
            entry["cgls_narrow_iterate"] = {"s": round(time.perf_counter() - t, 2),
                                            "rel_l2_vs_f32_run": rel(xf, x32), "rel_l2_vs_phantom": rel(xf, img)}
        res[name] = entry
        log("%s: %s" % (name, entry))
        del x, xf, imgd, projd
        torch.cuda.empty_cache()
    V, n = float(nx) ** 3, nviews
    result = {
        "metric": "storage precision at 1024^3 / 720 views: rel-L2 of Ax, Atb and CGLS(%d) per storage type" % niter,
        "value": res["bf16"]["cgls_rel_l2_vs_f32_run"],
        "unit": "rel-L2 (CGLS iterate with BF16 working vectors vs the FP32 run)",
        "n_gpus": 1, "steps": niter, "warmup": 1, "ms_per_step": round(res["f32"]["cgls_s"] * 1e3 / niter, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16/f16/f32 storage, f32 math",
        "data": "synthetic (seeded 30-ellipsoid phantom, projections from Ax)",
        "config": {"workload": "1024^3 (0.25 mm), 720 views, 1024x768 detector at 0.4 mm, jittered offsets, "
                               "CGLS %d iterations" % niter, "oracle_views": [int(i) for i in sub],
                   "oracle_seconds": round(t_oracle, 1)},
        "by_storage": res,
        "gupdates_per_s_ax_f32": round(V * n / (res["f32"]["ax_ms"] * 1e-3) / 1e9, 1),
    }
    return result


def bench_tv(args):
    """The TV steepest-descent step (minimizeTV, 20 iterations) on seeded
    40-ellipsoid phantoms at 256^3 and 512^3: kernel time per iteration and the
    rate against the 12 bytes per voxel and iteration of the two-pass model."""
    import ctypes

    import numpy as np
    import torch

    import tigre_amd
    from tigre_amd import _native
    from tigre_amd.phantom import ellipsoid_phantom_torch

    lib = _native.lib()
    out = {}
    for n in (256, 512):
        img = 0.02 * ellipsoid_phantom_torch((n, n, n), (float(n),) * 3, 40, seed=20261202 + 5, device="cuda")
        tigre_amd.minTV(img, 1e-3, 20)
        torch.cuda.synchronize()
        total, count = ctypes.c_double(0.0), ctypes.c_int(0)
        lib.tigre_kernel_time(_native.KERNEL_POCS_TV, ctypes.byref(total), ctypes.byref(count))
        lib.tigre_kernel_timing(1)
        s_, e_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s_.record()
        for _ in range(max(1, args.steps)):
            tigre_amd.minTV(img, 1e-3, 20)
        e_.record()
        e_.synchronize()
        lib.tigre_kernel_timing(0)
        _native.check(lib.tigre_kernel_time(_native.KERNEL_POCS_TV, ctypes.byref(total), ctypes.byref(count)),
                      "kernel_time")
        it_ms = total.value / count.value
        call_ms = s_.elapsed_time(e_) / max(1, args.steps)
        V = float(n) ** 3
        out["%d^3" % n] = {"iteration_us": round(it_ms * 1e3, 1), "call_20_iterations_ms": round(call_ms, 3),
                           "model_gb_s": round(12.0 * V / (it_ms * 1e-3) / 1e9, 1),
                           "frac_of_hbm_peak": round(12.0 * V / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log("TV %d^3: %.1f us per iteration, 20 iterations %.3f ms" % (n, it_ms * 1e3, call_ms))
        del img
        torch.cuda.empty_cache()
    result = {"metric": "minimizeTV iteration time (two passes + norm)", "value": out["512^3"]["iteration_us"],
              "unit": "us per iteration at 512^3", "n_gpus": 1, "steps": args.steps, "warmup": 1,
              "higher_is_better": False, "dtype": "f32", "data": "synthetic (seeded 40-ellipsoid phantoms)",
              "config": {"workload": "minTV 20 iterations at 256^3 and 512^3"}, "by_size": out}
    print(json.dumps(result), flush=True)


def bench_small(args):
    """Every operator at 128^3 (1 mm voxels), 128 views over 2*pi, 128 x 128
    detector at 1.6 mm, DSO 1000 / DSD 1536 mm, seeded 10-ellipsoid phantom:
    Ax 'Siddon' and 'interpolated', Atb 'FDK' and 'matched', and FDK with the
    ram-lak filter, each timed on the GPU (device-resident, median of 10 after
    3 warm-ups) and in full by the FP64 OpenMP oracle on the host cores, with
    the relative L2 error between the two."""
    import numpy as np
    import torch

    import oracle
    import oracle.algorithms as oalg
    import tigre_amd
    from tigre_amd import algorithms as algs
    from tigre_amd.geometry import cone_geometry, prepared
    from tigre_amd.phantom import ellipsoid_phantom

    geo = cone_geometry((128, 128, 128), (128.0,) * 3, (128, 128), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 128, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 10, seed=20261202 + 1)
    gp = prepared(geo, angles)
    oracle.set_threads(cpu_threads())
    dimg = torch.as_tensor(img, device="cuda")
    proj_ref = oracle.ax(img, gp, "Siddon").astype(np.float32)
    dproj = torch.as_tensor(proj_ref, device="cuda")

    def gpu_ms(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(10):
            s_, e_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s_.record()
            fn()
            e_.record()
            e_.synchronize()
            ts.append(s_.elapsed_time(e_))
        return float(np.median(ts))

    def cpu_s(fn):
        t = time.perf_counter()
        out = fn()
        return out, time.perf_counter() - t

    ops = {
        "ax_siddon": (lambda: tigre_amd.Ax(dimg, geo, angles, "Siddon"), lambda: oracle.ax(img, gp, "Siddon")),
        "ax_interpolated": (lambda: tigre_amd.Ax(dimg, geo, angles, "interpolated"),
                            lambda: oracle.ax(img, gp, "interpolated")),
        "atb_fdk": (lambda: tigre_amd.Atb(dproj, geo, angles, "FDK"), lambda: oracle.atb(proj_ref, gp, "FDK")),
        "atb_matched": (lambda: tigre_amd.Atb(dproj, geo, angles, "matched"),
                        lambda: oracle.atb(proj_ref, gp, "matched")),
        "fdk_ram_lak": (lambda: algs.fdk(dproj, geo, angles, filter="ram_lak", return_tensor=True),
                        lambda: oalg.fdk(proj_ref, geo, angles, "ram_lak")[0]),
    }
    res = {}
    for name, (g_fn, c_fn) in ops.items():
        ms = gpu_ms(g_fn)
        got = g_fn().float().cpu().numpy()
        ref, t_cpu = cpu_s(c_fn)
        res[name] = {"gpu_ms": round(ms, 3), "cpu_s": round(t_cpu, 3), "rel_l2": oracle.rel_l2(got, ref),
                     "speedup": round(t_cpu * 1e3 / ms, 1)}
        log("%s: %s" % (name, res[name]))
    V, n = 128.0 ** 3, 128
    pair_ms = res["ax_siddon"]["gpu_ms"] + res["atb_fdk"]["gpu_ms"]
    result = {
        "metric": "Ax+Atb Gupdates/s (Siddon + FDK, cone beam), 128^3 / 128 views",
        "value": round(2 * V * n / (pair_ms * 1e-3) / 1e9, 2), "unit": "Gupdates/s", "n_gpus": 1, "steps": 10,
        "warmup": 3, "ms_per_step": round(pair_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded 10-ellipsoid phantom)",
        "config": {"workload": "cone beam 128^3 at 1 mm, 128 views, 128x128 detector at 1.6 mm, DSO 1000 / DSD 1536"},
        "operators": res,
        "cpu_baseline": {"value": round(2 * V * n / (res["ax_siddon"]["cpu_s"] + res["atb_fdk"]["cpu_s"]) / 1e9, 4),
                         "unit": "Gupdates/s", "cores": oracle.get_threads(), "kind": "port",
                         "sample": "FP64 OpenMP oracle, every operator in full"},
    }
    print(json.dumps(result), flush=True)


def copy_geo(geo, idx):
    """The geometry restricted to views idx (per-view arrays sliced)."""
    import copy

    import numpy as np

    g = copy.deepcopy(geo)
    off = np.asarray(g.offDetector)
    if off.ndim == 2:
        g.offDetector = off[idx]
    return g


if __name__ == "__main__":
    main()
