"""Benchmark of the per-grain level-set timestep on one MI355X.

Workload (the configuration the headline metric is quoted on): 8000 grains from
a seeded Voronoi tessellation (seed 2), 20 points per grain, domain border 3
(323^3 interior, 329^3 grid), Gaussian convolution with TimeStepSize 0.2,
isotropic, no coarsening.  A step is one pass of the hot path over every grain
box: convolution, neighbour lists, comparison, level set (existence, volume,
area, resize) and redistancing.  The timed steps of the FP32 batched path run
as replays of a captured HIP graph; the all-FP64 path is timed the same way and
reported beside it.

value = voxel updates per second = sum over steps of sum_boxes n_b^3 / wall
time of the timed steps, with every input resident in HBM.

Output: one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gragles_3d_amd as G  # noqa: E402

METRIC = "grain-box voxel updates/sec per timestep, FP32-batched vs FP64-plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CLASS_NMAX = (32, 48, 64, 96)  # box-size classes of the work plan, one comparison launch each


def size_class(n):
    for c, m in enumerate(CLASS_NMAX):
        if n <= m:
            return c
    return len(CLASS_NMAX) - 1


def kbar_and_voxels(ctx):
    """Exact mean number of comparison-list boxes covering a voxel, K = sum_i sum_{j in list(i)}
    |B_i cap B_j| / sum_i |B_i|, over the grains alive now, from their box bounds; overall and
    per box-size class (numerator and voxel count per class)."""
    st, _, _ = ctx.stats()
    alive = st == G.ALIVE
    B = ctx.all_bounds(1).astype(np.int64)
    offs, ids = ctx.all_neighbour_lists(0)
    n = B[:, 3] - B[:, 0]
    cls = np.searchsorted(np.array(CLASS_NMAX), n)
    cls = np.minimum(cls, len(CLASS_NMAX) - 1)
    own = np.repeat(np.arange(len(offs) - 1), np.diff(offs))
    keep = alive[own] & alive[ids]
    own, nb = own[keep], ids[keep]
    w = np.ones(len(own), dtype=np.int64)
    for a in range(3):
        w *= np.maximum(0, np.minimum(B[own, a + 3], B[nb, a + 3]) - np.maximum(B[own, a], B[nb, a]))
    num = np.bincount(cls[own], weights=w, minlength=len(CLASS_NMAX)).astype(np.float64)
    g = np.nonzero(alive)[0]
    den = np.bincount(cls[g], weights=n[g].astype(np.float64) ** 3, minlength=len(CLASS_NMAX))
    return num.sum() / max(den.sum(), 1), num, den


def run_path(precision, inputs, p, steps, warmup, use_graph=True):
    ctx = G.LevelSetContext(p, precision=precision)
    ctx.set_boxes(inputs["bounds"], inputs["exists"])
    ctx.init_distance(inputs["idfield"])
    ctx.set_neighbour_lists(inputs["offsets"], inputs["ids"])
    ctx.synchronize()
    ctx.timestep(warmup, use_graph=use_graph)
    ctx.synchronize()
    vox_before = ctx.box_voxels()
    t0 = time.perf_counter()
    ctx.timestep(steps, use_graph=use_graph)
    ctx.synchronize()
    t1 = time.perf_counter()
    vox_after = ctx.box_voxels()
    return ctx, t1 - t0, vox_before, vox_after


def kernel_pass(ctx, steps):
    """Eager steps with every kernel segment bracketed by HIP events on the library's stream."""
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    ctx.timestep(steps, use_graph=False)
    ctx.synchronize()
    seg = ctx.stage_times(reset=True)
    ctx.enable_stage_timing(False)
    return seg


def container_leg(ctx):
    """One voxel container (saveNetworkAsVoxelContainer's id grid) of the
    context's state: the kernel clears size^3 ids and scatters every alive
    grain's phi >= 0 voxels; algorithmic bytes = 4 B per box voxel read + 4 B
    per container voxel cleared and written once."""
    st, _, _ = ctx.stats()
    b = ctx.all_bounds(1)[st == G.ALIVE]
    box_vox = int(((b[:, 3] - b[:, 0]).astype(np.int64) ** 3).sum())
    ctx.voxel_container()  # first call: buffer allocation and the kernels' first load
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    t = time.perf_counter()
    cont = ctx.voxel_container()
    wall = time.perf_counter() - t
    ms, _ = ctx.stage_times(reset=True)["container"]
    ctx.enable_stage_timing(False)
    nbytes = 4.0 * box_vox + 8.0 * cont.size
    return {"size": int(cont.shape[0]), "call_ms": wall * 1e3, "kernel_ms": ms, "kernel_bytes": nbytes,
            "kernel_GBs": nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
            "empty_voxel_fraction": float((cont == 0).mean())}


def coarsening_leg(ctx, p, frac=0.9):
    """One grid coarsening (gridCoarsement's coarsening branch) of the context's
    state onto a grid frac times the size, then one timestep on it.  The
    resampling kernel reads every alive grain's old box and writes its new box:
    4 B per voxel each way (FP32) is its algorithmic traffic."""
    st, _, _ = ctx.stats()
    alive = st == G.ALIVE
    old_b = ctx.all_bounds(1)[alive]
    src_vox = int(((old_b[:, 3] - old_b[:, 0]).astype(np.int64) ** 3).sum())
    rds = p.ngridpoints - 2 * p.grid_blowup
    new = int(rds * frac)
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    ctx.synchronize()
    t = time.perf_counter()
    ctx.coarsen_grid(new, 0.2 / float(new * new))  # Gaussian mode, TimeStepSize 0.2 (derive_params' default)
    ctx.synchronize()
    wall = time.perf_counter() - t
    ms, nl = ctx.stage_times(reset=True)["resample"]
    ctx.enable_stage_timing(False)
    new_b = ctx.all_bounds(1)[alive]
    dst_vox = int(((new_b[:, 3] - new_b[:, 0]).astype(np.int64) ** 3).sum())
    nbytes = 4.0 * (src_vox + dst_vox)
    ctx.timestep(1, use_graph=False)
    ctx.synchronize()
    st2, _, _ = ctx.stats()
    return {"grid_from": rds, "grid_to": new, "call_ms": wall * 1e3, "resample_ms": ms,
            "resample_bytes": nbytes, "resample_GBs": nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
            "grains_alive_after_next_step": int((st2 == G.ALIVE).sum())}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_fp32_8000grains.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed profiler summary (FETCH_SIZE and
    WRITE_SIZE passes of rocprofv3 on this workload, scaled by a copy of known size), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch %s:" % kernel):
                    return float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size (the runtime's blit kernel), so that HBM
    byte counters read under a profiler can be scaled: a copy of B bytes must read and
    write B bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d device-to-device copies of %d bytes" % (reps, nbytes))


def cpu_baseline(inputs, p, nthreads):
    from oracle import oracle as O
    O.set_threads(nthreads)
    orc = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, p.n_grains)
    orc.set_boxes(inputs["bounds"], inputs["exists"])
    orc.init_distance(inputs["idfield"])
    orc.set_lists(inputs["offsets"], inputs["ids"])
    vox = 0
    for g in range(1, p.n_grains + 1):
        if orc.state(g) == O.ALIVE:
            b, _ = orc.bounds(g, 1)
            vox += int(b[3] - b[0]) ** 3
    t0 = time.perf_counter()
    orc.timestep(1)
    dt = time.perf_counter() - t0
    orc.close()
    return vox / dt, dt, vox


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--grains", type=int, default=8000)
    ap.add_argument("--ppg", type=int, default=20)
    ap.add_argument("--border", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--kernel-steps", type=int, default=4, help="eager event-timed steps for the roofline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run device-to-device copies of 1 GiB (scale for profiler byte counters)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the voxel-container and grid-coarsening legs (they follow the timed steps; "
                         "profiles of the timestep leave them out so per-kernel means cover one grid)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the per-grain timestep runs every grain box on a single MI355X "
              "(no sharding); --gpus %d is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        copy_calibration()
    p = G.derive_params(args.grains, args.ppg, args.border)
    t = time.perf_counter()
    idf, bounds, exists = G.generate_voronoi(p, args.seed, True, 0)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    log("setup: %d grains, grid %d^3, generator %.2fs" % (args.grains, p.ngridpoints, time.perf_counter() - t))
    inputs = dict(idfield=idf, bounds=bounds, exists=exists, offsets=offs, ids=ids)

    # ---- FP32 batched, graph replays ------------------------------------------
    ctx32, wall32, vb32, va32 = run_path(G.FP32_BATCHED, inputs, p, args.steps, args.warmup)
    vox_step32 = 0.5 * (vb32 + va32)
    value = vox_step32 * args.steps / wall32
    log("fp32: %d steps %.3fs -> %.3f ms/step, %.3e voxel-updates/s (boxes %.3e voxels)" %
        (args.steps, wall32, 1e3 * wall32 / args.steps, value, vox_step32))
    kbar, knum, kden = kbar_and_voxels(ctx32)

    # ---- FP64 plain, same steps; then the FP32 error against it ------------------
    fp64 = None
    err = None
    if not args.no_fp64:
        ctx64, wall64, vb64, va64 = run_path(G.FP64_PLAIN, inputs, p, args.steps, args.warmup)
        vox_step64 = 0.5 * (vb64 + va64)
        v64 = vox_step64 * args.steps / wall64
        log("fp64: %d steps %.3fs -> %.3f ms/step, %.3e voxel-updates/s" % (args.steps, wall64,
                                                                          1e3 * wall64 / args.steps, v64))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from fp32_error import phi_volume_errors, trajectory_error
        err = phi_volume_errors(ctx32, ctx64, p.delta, grains=np.arange(1, p.n_grains + 1, 16))
        err.update(trajectory_error(ctx32.count_history(), ctx64.count_history()))
        err["grains_sampled"] = "every 16th grain id"
        seg64 = kernel_pass(ctx64, max(1, args.kernel_steps // 2))
        ctx64.close()
        fp64 = {"value": v64, "ms_per_step": 1e3 * wall64 / args.steps,
                "segment_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in seg64.items() if v[1]}}

    seg32 = kernel_pass(ctx32, args.kernel_steps)
    counts32 = ctx32.count_history()
    container = coarsen = None
    if not args.no_extras:
        container = container_leg(ctx32)
        log("voxel container %d^3: kernel %.3f ms (%.0f GB/s), call with copy-out %.1f ms" % (
            container["size"], container["kernel_ms"], container["kernel_GBs"], container["call_ms"]))
        coarsen = coarsening_leg(ctx32, p)
        log("coarsening %d -> %d: call %.2f ms, resampling %.3f ms (%.0f GB/s)" % (
            coarsen["grid_from"], coarsen["grid_to"], coarsen["call_ms"], coarsen["resample_ms"],
            coarsen["resample_GBs"]))
    ctx32.close()

    # Roofline of the comparison kernel (neighbour max + wall + set comparison), one launch per
    # box-size class: per voxel it reads K neighbour values and its own value (4 B each) and its
    # previous IDLocal (4 B), and writes the new value and IDLocal (8 B) -> (4K + 12) B per voxel.
    # The dominant kernel is the class launch with the largest event-timed duration.
    cls_rows = []
    for c, m in enumerate(CLASS_NMAX):
        ms, nl = seg32["comparison_n%d" % m]
        avg_s = ms / max(nl, 1) * 1e-3
        kb = knum[c] / kden[c] if kden[c] else 0.0
        bts = kden[c] * (4.0 * kb + 12.0)
        cls_rows.append({"kernel": "k_comparison<float, %d>" % m, "box_voxels": int(kden[c]), "kbar": kb,
                         "bytes_per_launch": bts, "avg_launch_ms": avg_s * 1e3,
                         "achieved_GBs": bts / avg_s / 1e9 if avg_s > 0 and bts > 0 else 0.0})
    dom = max(cls_rows, key=lambda r: r["avg_launch_ms"])
    achieved = dom["achieved_GBs"]
    cmp_total_ms = sum(r["avg_launch_ms"] for r in cls_rows)
    cmp_total_bytes = sum(r["bytes_per_launch"] for r in cls_rows)
    seg_ms_per_step = {k: round(v[0] / max(v[1], 1), 4) for k, v in seg32.items() if v[1]}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "%d grains, seeded Voronoi (seed %d), %d points/grain, border %d, grid %d^3, "
                        "FP32 batched path, hipGraph-replayed steps" % (args.grains, args.seed, args.ppg,
                                                                        args.border, p.ngridpoints),
            "grains": args.grains, "points_per_grain": args.ppg, "border": args.border,
            "ngridpoints": p.ngridpoints, "box_voxels_per_step": vox_step32, "kbar": kbar,
        },
        "roofline": {
            "kernel": dom["kernel"],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom["kernel"]),
            "traffic_source": "profiles/hbm_traffic_fp32_8000grains.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                              "copy-calibrated)",
            "bytes_per_launch": dom["bytes_per_launch"],
            "avg_launch_ms": dom["avg_launch_ms"],
            "kbar": dom["kbar"],
        },
        "comparison_by_class": cls_rows,
        "comparison_all_classes": {"ms_per_step": cmp_total_ms, "bytes_per_step": cmp_total_bytes,
                                   "GBs": cmp_total_bytes / (cmp_total_ms * 1e-3) / 1e9 if cmp_total_ms else 0.0},
        "segment_ms_per_step_fp32": seg_ms_per_step,
        "grain_count_trajectory_fp32": counts32[-3:].tolist(),
        "voxel_container_fp32": container,
        "grid_coarsening_fp32": coarsen,
    }

    if fp64 is not None:
        out["fp64_plain"] = fp64
        out["fp32_speedup_vs_fp64"] = value / fp64["value"]
        out["fp32_error_vs_fp64"] = err

    if not args.no_cpu:
        try:
            v_cpu, dt_cpu, vox_cpu = cpu_baseline(inputs, p, args.cpu_threads)
            out["cpu_baseline"] = {"value": v_cpu, "unit": "voxel-updates/s", "cores": args.cpu_threads,
                                   "kind": "port",
                                   "sample": "FP64 CPU oracle (g++ -O3 -fopenmp), one full timestep of the same "
                                             "%d-grain workload (%.3e box voxels) in %.1fs" % (args.grains, vox_cpu,
                                                                                              dt_cpu)}
            log("cpu oracle: %.1fs for one step -> %.3e voxel-updates/s" % (dt_cpu, v_cpu))
        except Exception as e:  # the checker may be absent on a product install
            log("cpu baseline skipped: %s" % e)
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
