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

Order of the run: set-up; one autotuning timestep (the comparison's tile
shape per box-size class, gg_autotune); an event-timed pass of --kernel-steps
eager steps (every kernel segment bracketed by HIP events on the
library's stream, all launches on that one stream, so each kernel runs alone:
the per-kernel roofline rows come from it, and a profiler run of
`--steps 0 --warmup 0 --no-fp64 --no-cpu --no-extras --no-accuracy` sees
exactly these dispatches); warm-up and timed graph steps; the FP64 path; the
200-step accuracy leg; the CPU baseline.

Output: one JSON line on stdout.  `--summarize-rocprof DIR` instead prints the
markdown tables of a rocprofv3 output directory (kernel stats, and FETCH_SIZE /
WRITE_SIZE passes scaled by the copy calibration).
"""
import argparse
import csv
import glob
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gragles_3d_amd as G  # noqa: E402

METRIC = "grain-box voxel updates/sec per timestep, FP32-batched vs FP64-plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0
CLASS_NMAX = (32, 48, 64, 96)  # box-size classes of the work plan, one launch each per stage
CMP_SHAPES = G.levelset.CMP_SHAPES  # comparison tile shapes of gg_autotune (layers, gathers in flight, run-major)
CMP_SPACE_SHAPES = [(8, 8), (16, 8), (16, 16), (32, 8)]  # shapes of the spatially ordered launch
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_fp32_8000grains.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kbar_and_voxels(state, B, offs, ids):
    """Exact mean number of comparison-list boxes covering a voxel,
    K = sum_i sum_{j in list(i)} |B_i cap B_j| / sum_i |B_i|, over the alive grains, from box
    bounds B and lists (CSR offs, ids); overall and per box-size class (numerator and voxel count
    per class)."""
    alive = state == G.ALIVE
    B = B.astype(np.int64)
    n = B[:, 3] - B[:, 0]
    cls = np.minimum(np.searchsorted(np.array(CLASS_NMAX), n), len(CLASS_NMAX) - 1)
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


def new_context(precision, inputs, p):
    ctx = G.LevelSetContext(p, precision=precision)
    ctx.set_boxes(inputs["bounds"], inputs["exists"])
    ctx.init_distance(inputs["idfield"])
    ctx.set_neighbour_lists(inputs["offsets"], inputs["ids"])
    ctx.synchronize()
    return ctx


def tune_comparison(precision, inputs, p, reps):
    """gg_autotune in both comparison orders, each on a context of its own taking the same first
    timestep: the tile shape per size class (grain order) and of the one launch (spatial order),
    and the faster order.  Returns (shapes per class, order, spatial shape, report)."""
    tctx = new_context(precision, inputs, p)
    choice, ms = tctx.autotune(reps=reps)
    tctx.close()
    choice = choice.tolist()
    grain_ms = float(sum(ms[c][choice[c]] for c in range(len(CLASS_NMAX))))
    tctx = new_context(precision, inputs, p)
    tctx.set_comparison_order(G.CMP_ORDER_SPACE)
    schoice, sms = tctx.autotune(reps=reps)
    tctx.close()
    shape = int(schoice[0])
    space_ms = float(sms[0][shape])
    order = G.CMP_ORDER_SPACE if space_ms < grain_ms else G.CMP_ORDER_GRAIN
    report = {"choice_per_class": choice, "ms_per_class_and_shape": np.round(ms, 4).tolist(),
              "grain_order_ms": grain_ms, "spatial_order_ms_per_shape": np.round(sms[0][:4], 4).tolist(),
              "spatial_order_shape": shape, "spatial_order_ms": space_ms,
              "order": "spatial" if order == G.CMP_ORDER_SPACE else "grain",
              "source": "autotuner (one timestep on a separate context per order)"}
    return choice, order, shape, report


def timed_steps(ctx, steps, warmup, use_graph=True):
    ctx.timestep(warmup, use_graph=use_graph)
    ctx.synchronize()
    vox_before = ctx.box_voxels()
    t0 = time.perf_counter()
    ctx.timestep(steps, use_graph=use_graph)
    ctx.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, 0.5 * (vox_before + ctx.box_voxels())


def kernel_pass(ctx, steps, fp32):
    """`steps` eager steps, every kernel segment bracketed by HIP events on the library's stream
    (all launches then go to that stream: each kernel runs alone).  Per step the state before it
    gives the algorithmic bytes of each stage.  Returns the per-segment times, the per-stage bytes
    and the per-class comparison bytes, summed over the steps."""
    esz = 4 if fp32 else 8
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    by = {"convolution": 0.0, "comparison": 0.0, "hull": 0.0, "redistancing": 0.0}
    cls_bytes = np.zeros(len(CLASS_NMAX))
    cls_vox = np.zeros(len(CLASS_NMAX))
    cls_num = np.zeros(len(CLASS_NMAX))
    for _ in range(steps):
        # The convolution, comparison and hull of a step run on the boxes it starts with, the
        # redistancing on the new ones (nearly the same volume); the comparison reads the boxes of
        # its comparison list, which is the neighbour list after the step.
        state = ctx.stats()[0]
        B = ctx.all_bounds(1)
        ctx.timestep(1, use_graph=False)
        offs, ids = ctx.all_neighbour_lists(0)
        _, num, den = kbar_and_voxels(state, B, offs, ids)
        vox = den.sum()
        by["convolution"] += 2 * esz * vox
        by["hull"] += esz * vox
        by["redistancing"] += 2 * esz * vox
        # K neighbour values and the own value read, the value and the IDLocal id written
        cb = esz * num + (2 * esz + 4) * den
        by["comparison"] += cb.sum()
        cls_bytes += cb
        cls_vox += den
        cls_num += num
    ctx.synchronize()
    seg = ctx.stage_times(reset=True)
    ctx.enable_stage_timing(False)
    return seg, by, cls_bytes, cls_vox, cls_num


STAGE_SEGMENTS = {
    "convolution": ("conv_xy", "conv_z", "conv_xy_inv"),
    "comparison": ("comparison_n32", "comparison_n48", "comparison_n64", "comparison_n96"),
    "hull": ("hull",),
    "redistancing": ("redist_xy", "redist_z"),
}


def roofline_rows(seg, by, steps):
    rows = {}
    for st, segs in STAGE_SEGMENTS.items():
        ms = sum(seg[s][0] for s in segs) / max(steps, 1)
        b = by[st] / max(steps, 1)
        rows[st] = {"ms_per_step": ms, "algorithmic_bytes_per_step": b,
                    "GBs": b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
                    "frac_of_hbm_peak": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS if ms > 0 else 0.0}
    return rows


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed profiler summary (FETCH_SIZE and
    WRITE_SIZE passes of rocprofv3 on the kernel pass of this workload, scaled by a copy of known
    size), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                # a row of the summary table: | kernel | dispatches | ... | traffic GB/dispatch |
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 7 and cells[0] == kernel:
                    return float(cells[6]) * 1e9
    except (OSError, ValueError):
        pass
    return None


def copy_calibration(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size (the runtime's blit kernel), so that HBM byte
    counters read under a profiler can be scaled: a copy of B bytes must read and write B bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b),
                                                                                 ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d device-to-device copies of %d bytes" % (reps, nbytes))


def cpu_description():
    """CPU model and physical core count from lscpu, and the CPUs this process may use."""
    model, sockets, cores_per_socket = platform.processor() or "unknown", 1, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Socket(s)":
                sockets = int(v)
            elif k.strip() == "Core(s) per socket":
                cores_per_socket = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    physical = sockets * cores_per_socket if cores_per_socket else None
    usable = len(os.sched_getaffinity(0))
    return {"model": model, "physical_cores": physical, "usable_cpus": usable}


def cpu_baseline(inputs, p, threads, full_steps, sample_stride):
    """The FP64 CPU oracle (g++ -O3 -fopenmp) on the same workload: `full_steps` whole timesteps
    with `threads` OpenMP threads, and one timestep with 1 thread over the grains whose id is a
    multiple of sample_stride (a bounded sample).  OMP_PROC_BIND / OMP_PLACES are those of the
    environment (set before the process starts)."""
    from oracle import oracle as O

    def sim():
        orc = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, p.n_grains)
        orc.set_boxes(inputs["bounds"], inputs["exists"])
        orc.init_distance(inputs["idfield"])
        orc.set_lists(inputs["offsets"], inputs["ids"])
        return orc

    def box_voxels(orc, ids):
        v = 0
        for g in ids:
            if orc.state(g) == O.ALIVE:
                b, _ = orc.bounds(g, 1)
                v += int(b[3] - b[0]) ** 3
        return v

    O.set_sample_stride(1)
    O.set_threads(threads)
    orc = sim()
    vox = 0
    t0 = time.perf_counter()
    for _ in range(full_steps):
        vox += box_voxels(orc, range(1, p.n_grains + 1))
        orc.timestep(1)
    t_full = time.perf_counter() - t0
    orc.close()
    O.set_threads(1)
    orc = sim()
    O.set_sample_stride(sample_stride)
    sample_ids = range(sample_stride, p.n_grains + 1, sample_stride)
    vox1 = box_voxels(orc, sample_ids)
    t0 = time.perf_counter()
    orc.timestep(1)
    t_one = time.perf_counter() - t0
    O.set_sample_stride(1)
    orc.close()
    return vox / t_full, t_full, vox, vox1 / t_one, t_one, vox1, len(sample_ids)


# ---- rocprofv3 output summaries ------------------------------------------------
def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def short_name(name):
    """k_comparison<float, 48> from the demangled 'void gg::(anonymous namespace)::k_comparison<float, 48>(...)'."""
    n = name.replace("void ", "", 1) if name.startswith("void ") else name
    n = n.replace("gg::(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def summarize_rocprof(d):
    """Markdown tables of a rocprofv3 output directory: kernel stats (from *kernel_stats.csv) and,
    if present, FETCH_SIZE / WRITE_SIZE counter passes (*counter_collection.csv) per kernel,
    scaled by the device-to-device copy kernel of the --calibrate-copy run (1 GiB per copy)."""
    stats = sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True))
    for s in stats:
        print("kernel stats (%s)\n" % os.path.relpath(s, d))
        print("| kernel | calls | avg (us) | min (us) | max (us) | total (ms) |")
        print("|---|---|---|---|---|---|")
        for r in sorted(_rows(s), key=lambda r: -float(r["TotalDurationNs"])):
            print("| %s | %s | %.1f | %.1f | %.1f | %.3f |" % (
                short_name(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                float(r["MaxNs"]) / 1e3, float(r["TotalDurationNs"]) / 1e6))
        print()
    counters = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    per = {}
    for c in counters:
        for r in _rows(c):
            name = short_name(r.get("Kernel_Name") or r.get("Kernel-Name") or r.get("Name"))
            cn = r.get("Counter_Name") or r.get("Counter-Name")
            val = float(r.get("Counter_Value") or r.get("Counter-Value") or 0.0)
            per.setdefault((name, cn), []).append(val)
    if not per:
        return
    sq = sorted({cn for _, cn in per if cn.startswith("SQ_")})
    if sq:
        # SQ counters: mean per dispatch; LDS bank-conflict cycles as a share of the cycles LDS
        # instructions were active (both counted by the SQ in the same units)
        print("| kernel | dispatches | " + " | ".join(c[3:] for c in sq) +
              (" | LDS bank-conflict share |" if {"SQ_LDS_BANK_CONFLICT", "SQ_ACTIVE_INST_LDS"} <= set(sq) else " |"))
        print("|---|---|" + "---|" * len(sq) + ("---|" if "SQ_LDS_BANK_CONFLICT" in sq else ""))
        names = sorted({n for n, cn in per if cn in sq}, key=lambda n: -np.mean(per.get((n, sq[0]), [0])))
        for n in names:
            vals = [np.mean(per.get((n, c), [0.0])) for c in sq]
            row = "| %s | %d | " % (n, len(per.get((n, sq[0]), []))) + " | ".join("%.3g" % v for v in vals)
            if "SQ_LDS_BANK_CONFLICT" in sq and "SQ_ACTIVE_INST_LDS" in sq:
                a = np.mean(per.get((n, "SQ_ACTIVE_INST_LDS"), [0.0]))
                b = np.mean(per.get((n, "SQ_LDS_BANK_CONFLICT"), [0.0]))
                row += " | %s" % ("%.3f" % (b / a) if a > 0 else "-")
            print(row + " |")
        print()
        if not any(cn in ("FETCH_SIZE", "WRITE_SIZE") for _, cn in per):
            return
    cal = {}
    for (name, cn), vals in per.items():
        if "copyBuffer" in name:
            big = [v for v in vals if v > 1e5]
            if big:
                cal[cn] = float(np.median(big)) / (1 << 20)  # KiB counted per KiB copied
    print("calibration (KiB counted per KiB of a 1 GiB device-to-device copy): %s\n" % cal)
    print("| kernel | dispatches | FETCH_SIZE KiB/dispatch (raw) | read GB (scaled) | "
          "WRITE_SIZE KiB/dispatch (raw) | write GB (scaled) | traffic GB/dispatch |")
    print("|---|---|---|---|---|---|---|")
    names = sorted({n for n, _ in per}, key=lambda n: -np.mean(per.get((n, "FETCH_SIZE"), [0])))
    for n in names:
        f = per.get((n, "FETCH_SIZE"), [])
        w = per.get((n, "WRITE_SIZE"), [])
        fr = np.mean(f) if f else 0.0
        wr = np.mean(w) if w else 0.0
        rb = fr * 1024 / cal.get("FETCH_SIZE", 1.0) / 1e9
        wb = wr * 1024 / cal.get("WRITE_SIZE", 1.0) / 1e9
        print("| %s | %d | %.0f | %.3f | %.0f | %.3f | %.3f |" % (n, max(len(f), len(w)), fr, rb, wr, wb, rb + wb))
        if fr and wr and n.startswith("k_"):
            print("traffic_bytes_per_launch %s: %.6e" % (n, (rb + wb) * 1e9), file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--grains", type=int, default=8000)
    ap.add_argument("--ppg", type=int, default=20)
    ap.add_argument("--border", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--kernel-steps", type=int, default=4, help="eager event-timed steps for the roofline rows")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 200-step FP32-vs-FP64 accuracy leg")
    ap.add_argument("--accuracy-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run device-to-device copies of 1 GiB (scale for profiler byte counters)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the CPUs this process may use")
    ap.add_argument("--cpu-steps", type=int, default=2, help="whole timesteps of the multi-thread CPU baseline")
    ap.add_argument("--cpu-sample-stride", type=int, default=16,
                    help="the 1-thread CPU baseline times one step over every n-th grain")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default comparison tile shapes instead of autotuning them (one timestep)")
    ap.add_argument("--cmp-order", choices=("auto", "grain", "spatial"), default="auto",
                    help="comparison order of the FP32 leg (auto: the autotuner's faster one)")
    ap.add_argument("--space-shape", type=int, choices=range(4), default=None,
                    help="tile shape of the spatially ordered comparison, instead of autotuning it")
    ap.add_argument("--tuning", metavar="S0,S1,S2,S3",
                    help="comparison tile shape per size class (0..3), instead of autotuning them")
    ap.add_argument("--no-physics", action="store_true",
                    help="skip the configurations with the convolution correction (anisotropic, stored energy)")
    ap.add_argument("--physics-steps", type=int, default=10)
    ap.add_argument("--physics-accuracy-steps", type=int, default=50)
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven batch (50,000 bimodal grains, volume-sorted queue)")
    ap.add_argument("--uneven-steps", type=int, default=10)
    ap.add_argument("--uneven-accuracy-steps", type=int, default=100)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the voxel-container and grid-coarsening legs")
    ap.add_argument("--summarize-rocprof", metavar="DIR", help="print markdown tables of a rocprofv3 directory")
    args = ap.parse_args()
    if args.summarize_rocprof:
        summarize_rocprof(args.summarize_rocprof)
        return
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

    # ---- FP32 batched: autotuned comparison shapes (one timestep), event-timed kernel pass, ----
    # ---- then graph replays                                                                  ----
    # The autotuner runs a timestep; it runs on a context of its own, so that the measured
    # context starts from the initial state (as a profiled run with --tuning does).
    tune32 = None
    choice32 = None
    order32, shape32 = G.CMP_ORDER_GRAIN, 0
    if args.tuning:
        choice32 = [int(v) for v in args.tuning.split(",")]
        tune32 = {"choice_per_class": choice32, "source": "--tuning"}
    elif not args.no_tune:
        choice32, order32, shape32, tune32 = tune_comparison(G.FP32_BATCHED, inputs, p, 3)
        log("fp32 comparison: grain order %.3f ms (shapes per class %s), spatial order %.3f ms (shape %d) -> %s"
            % (tune32["grain_order_ms"], choice32, tune32["spatial_order_ms"], shape32, tune32["order"]))
    if args.space_shape is not None:
        shape32 = args.space_shape
    if args.cmp_order == "grain":
        order32 = G.CMP_ORDER_GRAIN
    elif args.cmp_order == "spatial":
        order32 = G.CMP_ORDER_SPACE
    ctx32 = new_context(G.FP32_BATCHED, inputs, p)
    if choice32 is not None:
        ctx32.set_tuning(choice32)
        log("fp32 comparison shapes per class (n<=32,48,64,96): %s" % choice32)
    ctx32.set_comparison_order(order32, shape32)
    seg32, by32, cls_bytes, cls_vox, cls_num = kernel_pass(ctx32, args.kernel_steps, True)
    ks = max(args.kernel_steps, 1)
    rows32 = roofline_rows(seg32, by32, ks)
    wall32, vox_step32 = timed_steps(ctx32, args.steps, args.warmup)
    value = vox_step32 * args.steps / wall32 if args.steps else 0.0
    log("fp32: %d steps %.3fs -> %.3f ms/step, %.3e voxel-updates/s (boxes %.3e voxels)" %
        (args.steps, wall32, 1e3 * wall32 / max(args.steps, 1), value, vox_step32))

    cls_rows = []
    tuning = ctx32.tuning()
    if order32 == G.CMP_ORDER_SPACE:
        # one launch for every class; its segment also holds the neighbour table and the
        # work-list build (about 2 % of it), so `achieved` is a lower bound for the kernel
        ms, nl = seg32["comparison_n32"]
        avg_s = ms / max(nl, 1) * 1e-3
        bts = cls_bytes.sum() / ks
        zl, qb = CMP_SPACE_SHAPES[shape32]
        b16 = "true" if args.grains < 0xFFFF else "false"  # winner ids packed as 16-bit halves below 0xFFFF grains
        cls_rows.append({"kernel": "k_cmp_space<float, %d, %d, %s>" % (zl, qb, b16), "box_voxels": int(cls_vox.sum() / ks),
                         "kbar": cls_num.sum() / cls_vox.sum() if cls_vox.sum() else 0.0, "bytes_per_launch": bts,
                         "avg_launch_ms": avg_s * 1e3,
                         "achieved_GBs": bts / avg_s / 1e9 if avg_s > 0 and bts > 0 else 0.0,
                         "segment": "neighbour table + work list + comparison"})
    else:
        for c, m in enumerate(CLASS_NMAX):
            ms, nl = seg32["comparison_n%d" % m]
            avg_s = ms / max(nl, 1) * 1e-3
            bts = cls_bytes[c] / ks
            zl, qb, zmaj = CMP_SHAPES[tuning[c]]
            cls_rows.append({"kernel": "k_cmp_cols<float, %d, %d, %d, %s>" % (m, zl, qb, "true" if zmaj else "false"),
                             "box_voxels": int(cls_vox[c] / ks),
                             "kbar": cls_num[c] / cls_vox[c] if cls_vox[c] else 0.0, "bytes_per_launch": bts,
                             "avg_launch_ms": avg_s * 1e3,
                             "achieved_GBs": bts / avg_s / 1e9 if avg_s > 0 and bts > 0 else 0.0})
    dom = max(cls_rows, key=lambda r: r["avg_launch_ms"])
    cmp_red_ms = rows32["comparison"]["ms_per_step"] + rows32["redistancing"]["ms_per_step"]
    cmp_red_b = rows32["comparison"]["algorithmic_bytes_per_step"] + rows32["redistancing"][
        "algorithmic_bytes_per_step"]
    seg_ms_per_step = {k: round(v[0] / ks, 4) for k, v in seg32.items() if v[1]}

    # ---- FP64 plain, same schedule ------------------------------------------------------------
    fp64 = None
    ctx64 = None
    if not args.no_fp64:
        order64, shape64 = order32, shape32
        if not args.no_tune and not args.tuning:
            choice64, order64, shape64, _ = tune_comparison(G.FP64_PLAIN, inputs, p, 2)
        else:
            choice64 = choice32
        ctx64 = new_context(G.FP64_PLAIN, inputs, p)
        if choice64 is not None:
            ctx64.set_tuning(choice64)
        ctx64.set_comparison_order(order64, shape64)
        seg64, by64, _, _, _ = kernel_pass(ctx64, max(1, args.kernel_steps // 2), False)
        wall64, vox_step64 = timed_steps(ctx64, args.steps, args.warmup)
        v64 = vox_step64 * args.steps / wall64 if args.steps else 0.0
        log("fp64: %d steps %.3fs -> %.3f ms/step, %.3e voxel-updates/s" % (
            args.steps, wall64, 1e3 * wall64 / max(args.steps, 1), v64))
        k64 = max(1, args.kernel_steps // 2)
        fp64 = {"value": v64, "ms_per_step": 1e3 * wall64 / max(args.steps, 1),
                "comparison_order": "spatial" if order64 == G.CMP_ORDER_SPACE else "grain",
                "segment_ms_per_step": {k: round(v[0] / k64, 4) for k, v in seg64.items() if v[1]},
                "stages": roofline_rows(seg64, by64, k64)}
        ctx64.close()

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

    # ---- configurations with the convolution correction ------------------------------------------
    physics = None
    if not args.no_physics:
        physics = {}
        for kind, grains in (("anisotropic", args.grains), ("stored_energy", 20000)):
            leg = physics_leg(kind, grains, args.ppg, args.border, args.seed, args.physics_steps, 2, choice32,
                              args.physics_accuracy_steps)
            physics["%s_%d_grains" % (kind, grains)] = leg
            log("%s, %d grains: %.3f ms/step, %.3e voxel-updates/s, correction %.3f ms + elements %.3f ms per step"
                % (kind, grains, leg["ms_per_step"], leg["value"], leg["segment_ms_per_step"].get("correction", 0),
                   leg["segment_ms_per_step"].get("elements", 0)))

    # ---- the uneven batch: 50,000 bimodal grains -------------------------------------------------
    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(50000, 16, args.border, 4, args.uneven_steps, 2, args.uneven_accuracy_steps)
        log("uneven batch, %d grains: %.3f ms/step (%s order), %.3e voxel-updates/s; the other order %.3f ms/step"
            % (uneven["grains"], uneven["ms_per_step"], uneven["work_order"], uneven["value"],
               uneven["ms_per_step_other_order"]))

    # ---- accuracy: FP32 beside FP64 (bit-identical to the oracle) over 200 steps ---------------
    accuracy = None
    if not args.no_accuracy and not args.no_fp64:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from fp32_error import phi_volume_errors, trajectories, trajectory_error
        a32 = new_context(G.FP32_BATCHED, inputs, p)
        a64 = new_context(G.FP64_PLAIN, inputs, p)
        tr = trajectories(a32, a64, args.accuracy_steps, every=10)
        c32, c64 = a32.count_history(), a64.count_history()
        e = phi_volume_errors(a32, a64, p.delta, grains=np.arange(1, p.n_grains + 1, 16))
        e.update(trajectory_error(c32, c64))
        e["grains_sampled_for_phi"] = "every 16th grain id"
        rel = np.abs(np.array(tr["mean_volume_fp32"]) / np.array(tr["mean_volume_fp64"]) - 1)
        accuracy = {"steps": args.accuracy_steps, "reference": "FP64 GPU path (bit-identical to the FP64 oracle)",
                    "after_last_step": e, "count_fp32_every_10": c32[9::10].tolist(),
                    "count_fp64_every_10": c64[9::10].tolist(),
                    "mean_radius_fp32": tr["mean_radius_fp32"], "mean_radius_fp64": tr["mean_radius_fp64"],
                    "mean_volume_rel_err_max": float(rel.max()),
                    "volume_rel_err_max_every_10": tr["volume_rel_err_max"],
                    "volume_rel_err_median_every_10": tr["volume_rel_err_median"],
                    "grains_above_1e-3_every_10": tr["grains_above_1e-3"],
                    "largest_voxels_above_1e-3_every_10": tr["largest_voxels_above_1e-3"],
                    "volume_rel_err_max_from_100_voxels_every_10": tr["volume_rel_err_max_from_100_voxels"]}
        log("accuracy %d steps: counts %d/%d, max count error %.3g, mean volume error %.2e" % (
            args.accuracy_steps, c32[-1], c64[-1], e["count_rel_err_max"], rel.max()))
        a32.close()
        a64.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / max(args.steps, 1),
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
            "ngridpoints": p.ngridpoints, "box_voxels_per_step": vox_step32,
        },
        "roofline": {
            "kernel": dom["kernel"],
            "bound": "hbm",
            "achieved": dom["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom["kernel"]),
            "bytes_per_launch": dom["bytes_per_launch"],
            "avg_launch_ms": dom["avg_launch_ms"],
            "kbar": dom["kbar"],
            "timing": "HIP events around each launch, %d eager steps from the initial state, one stream" % ks,
            "traffic_source": "profiles/hbm_traffic_fp32_8000grains.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "copy-calibrated)",
        },
        "stages_fp32": rows32,
        "comparison_plus_redistancing": {"ms_per_step": cmp_red_ms, "algorithmic_bytes_per_step": cmp_red_b,
                                         "GBs": cmp_red_b / (cmp_red_ms * 1e-3) / 1e9 if cmp_red_ms else 0.0,
                                         "frac_of_hbm_peak": cmp_red_b / (cmp_red_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                         if cmp_red_ms else 0.0, "target_frac": 0.6},
        "comparison_by_class": cls_rows,
        "comparison_autotune_fp32": tune32,
        "segment_ms_per_step_fp32": seg_ms_per_step,
        "voxel_container_fp32": container,
        "grid_coarsening_fp32": coarsen,
    }
    if fp64 is not None:
        out["fp64_plain"] = fp64
        out["fp32_speedup_vs_fp64"] = value / fp64["value"] if fp64["value"] else None
    if accuracy is not None:
        out["fp32_accuracy_vs_fp64"] = accuracy
    if physics is not None:
        out["configurations_with_correction_fp32"] = physics
    if uneven is not None:
        out["uneven_batch_fp32"] = uneven

    if not args.no_cpu:
        try:
            cpu = cpu_description()
            threads = args.cpu_threads or min(cpu["usable_cpus"], int(os.environ.get("OMP_NUM_THREADS", "64")))
            v_cpu, t_full, vox_full, v1, t1, vox1, nsample = cpu_baseline(inputs, p, threads, args.cpu_steps,
                                                                         args.cpu_sample_stride)
            out["cpu_baseline"] = {
                "value": v_cpu, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
                "sample": "FP64 CPU oracle (g++ -O3 -fopenmp -ffp-contract=off; its own mixed-radix FFT in place "
                          "of the reference's MKL/FFTW, hull topology every step as the reference), %d whole timesteps "
                          "of the same %d-grain workload from its initial state (%.3e box voxels) in %.1f s with "
                          "%d threads" % (args.cpu_steps, args.grains, vox_full, t_full, threads),
                "cpu_model": cpu["model"], "physical_cores": cpu["physical_cores"],
                "usable_cpus": cpu["usable_cpus"],
                "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
                "one_thread": {"value": v1, "unit": "voxel-updates/s", "seconds": t1, "box_voxels": vox1,
                               "sample": "one timestep over the %d grains whose id is a multiple of %d" % (
                                   nsample, args.cpu_sample_stride)},
            }
            log("cpu oracle: %.1fs for %d steps (%d threads) -> %.3e; 1 thread sample %.1fs -> %.3e voxel-updates/s"
                % (t_full, args.cpu_steps, threads, v_cpu, t1, v1))
        except Exception as e:  # the checker may be absent on a product install
            log("cpu baseline skipped: %s" % e)
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def physics_leg(kind, grains, ppg, border, seed, steps, warmup, tuning, accuracy_steps):
    """A configuration with the convolution correction on (src/box.cpp:428-536), FP32 batched:
    'anisotropic' -- Read-Shockley energies and sigmoidal mobilities from uniform random
    orientations; 'stored_energy' -- unit energies and mobilities, lognormal stored energies
    (median 0.1, sigma 1.0) as a driving-force shift.  TripleLineDrag 1e6.  Graph-replayed steps
    for the rate, one eager event-timed step for the stage split, and FP32 against the FP64 path
    (equal to the FP64 oracle bit for bit) over accuracy_steps."""
    p = G.derive_params(grains, ppg, border)
    t = time.perf_counter()
    idf, bounds, exists = G.generate_voronoi(p, seed, True, 0)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    inputs = dict(idfield=idf, bounds=bounds, exists=exists, offsets=offs, ids=ids)
    info = {"grains": grains, "points_per_grain": ppg, "border": border, "seed": seed,
            "ngridpoints": p.ngridpoints}
    if kind == "anisotropic":
        q = G.textured_orientations(grains, seed)
        phys = dict(correction=True, mobility_model=True, unique_energies=False, stored_energy=False,
                    triple_line_drag=1e6, quat=q)
        th = np.array([G.misorientation(q[a], q[b]) for a in range(1, grains + 1)
                       for b in ids[offs[a]:offs[a + 1]] if b > a])
        info["orientations"] = ("uniform random unit quaternions (Shoemake), 45%% of them replaced by a texture "
                                "component scattered by up to 7.5 degrees (seed %d)" % seed)
        info["low_angle_fraction_initial_neighbour_pairs"] = float((th < 15 * np.pi / 180).mean())
    else:
        see = G.lognormal_energies(grains, seed, 0.1, 1.0)
        phys = dict(correction=True, mobility_model=False, unique_energies=True, stored_energy=True,
                    triple_line_drag=1e6, see=see)
        info["stored_energy"] = "lognormal, median 0.1, sigma 1.0 (normalised units), seed %d" % seed
    info["setup_s"] = time.perf_counter() - t

    def context(precision, form=G.CONV_DEFAULT):
        c = new_context(precision, inputs, p)
        if tuning is not None:
            c.set_tuning(tuning)
        c.set_physics(**phys)
        if form != G.CONV_DEFAULT:
            c.set_convolution_form(form)
        return c

    ctx = context(G.FP32_BATCHED)
    ctx.timestep(1, use_graph=False)  # the first convolution has no hull to correct with
    ctx.synchronize()
    ctx.enable_stage_timing(True)
    ctx.stage_times(reset=True)
    ctx.timestep(2, use_graph=False)
    ctx.synchronize()
    seg = ctx.stage_times(reset=True)
    ctx.enable_stage_timing(False)
    wall, vox = timed_steps(ctx, steps, warmup)
    info.update({"value": vox * steps / wall if steps else 0.0, "unit": "voxel-updates/s",
                 "ms_per_step": 1e3 * wall / max(steps, 1), "box_voxels_per_step": vox, "steps": steps,
                 "segment_ms_per_step": {k: round(v[0] / 2, 4) for k, v in seg.items() if v[1]}})
    ctx.close()
    if kind == "stored_energy":
        # the split-BF16 (3-term) convolution on the matrix cores, timed the same way
        cb = context(G.FP32_BATCHED, G.CONV_SPLIT_BF16)
        cb.timestep(1, use_graph=False)
        cb.synchronize()
        cb.enable_stage_timing(True)
        cb.stage_times(reset=True)
        cb.timestep(2, use_graph=False)
        cb.synchronize()
        segb = cb.stage_times(reset=True)
        cb.enable_stage_timing(False)
        wall_b, vox_b = timed_steps(cb, steps, warmup)
        info["split_bf16_convolution"] = {
            "value": vox_b * steps / wall_b if steps else 0.0, "ms_per_step": 1e3 * wall_b / max(steps, 1),
            "segment_ms_per_step": {k: round(v[0] / 2, 4) for k, v in segb.items() if v[1]}}
        cb.close()
    if accuracy_steps:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from fp32_error import phi_volume_errors, trajectory_error
        a32, a64 = context(G.FP32_BATCHED), context(G.FP64_PLAIN)
        a32.timestep(accuracy_steps)
        a64.timestep(accuracy_steps)
        sample = np.arange(1, grains + 1, 16)
        e = phi_volume_errors(a32, a64, p.delta, grains=sample)
        e.update(trajectory_error(a32.count_history(), a64.count_history()))
        e["steps"] = accuracy_steps
        e["reference"] = "FP64 GPU path with the correction (bit-identical to the FP64 oracle)"
        e["grains_sampled_for_phi"] = "every 16th grain id"
        info["fp32_accuracy_vs_fp64"] = e
        if kind == "stored_energy":
            ab = context(G.FP32_BATCHED, G.CONV_SPLIT_BF16)
            ab.timestep(accuracy_steps)
            eb = phi_volume_errors(ab, a64, p.delta, grains=sample)
            eb.update(trajectory_error(ab.count_history(), a64.count_history()))
            eb["steps"] = accuracy_steps
            eb["reference"] = e["reference"]
            info["split_bf16_accuracy_vs_fp64"] = eb
            ef = phi_volume_errors(ab, a32, p.delta, grains=sample)
            ef.update(trajectory_error(ab.count_history(), a32.count_history()))
            ef["reference"] = "FP32 batched path with the FP32 circulant convolution"
            info["split_bf16_accuracy_vs_fp32_circulant"] = ef
            ab.close()
        a32.close()
        a64.close()
    return info


def uneven_leg(grains, ppg, border, seed, steps, warmup, accuracy_steps):
    """The uneven batch: 10 % of the grains four times larger, lognormal sizes (sigma 0.2)
    about each mode, power-diagram microstructure (gg_bimodal_seeds, gg_generate_power), FP32
    batched.  The comparison order and tile shapes come from the autotuner (one timestep on a
    separate context per order); the work order (device-side volume-sorted queue or spatial order) is chosen by
    timing graph-replayed steps in each; then FP32 against the FP64 path over accuracy_steps."""
    p = G.derive_params(grains, ppg, border)
    t = time.perf_counter()
    xyz, rho = G.bimodal_seeds(grains, seed)
    idf, bounds, exists = G.generate_power(p, xyz, rho)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    inputs = dict(idfield=idf, bounds=bounds, exists=exists, offsets=offs, ids=ids)
    cb = bounds[1:].copy()
    for r in cb:
        G.resize_to_cube(r, p.ngridpoints)
    side = (cb[:, 3] - cb[:, 0])[exists[1:] == 1]
    big = rho > np.median(rho) * 2.5
    info = {"grains": grains, "points_per_mean_diameter": ppg, "border": border, "seed": seed,
            "ngridpoints": p.ngridpoints, "setup_s": round(time.perf_counter() - t, 2),
            "microstructure": "bimodal: 10%% of the grains 4x larger, lognormal sigma 0.2, random sequential "
                              "addition + power diagram (seed %d)" % seed,
            "large_grain_fraction": float(big.mean()),
            "box_side_min_median_max": [int(side.min()), float(np.median(side)), int(side.max())],
            "box_side_median_small_large": [float(np.median(side[~big[exists[1:] == 1]])),
                                            float(np.median(side[big[exists[1:] == 1]]))],
            "grains_per_class": np.bincount(np.searchsorted(CLASS_NMAX, side), minlength=4).tolist()}
    choice, corder, cshape, crep = tune_comparison(G.FP32_BATCHED, inputs, p, 2)
    info["comparison_shapes_per_class"] = choice
    info["comparison_order"] = crep["order"]
    info["comparison_ms_grain_spatial"] = [crep["grain_order_ms"], crep["spatial_order_ms"]]
    timing = {}
    for order, name in ((G.ORDER_VOLUME, "volume-sorted"), (G.ORDER_SPATIAL, "spatial")):
        c = new_context(G.FP32_BATCHED, inputs, p)
        c.set_tuning(choice)
        c.set_comparison_order(corder, cshape)
        c.set_work_order(order)
        wall, vox = timed_steps(c, steps, warmup)
        timing[name] = (wall, vox)
        c.close()
    best = min(timing, key=lambda k: timing[k][0])
    other = [k for k in timing if k != best][0]
    wall, vox = timing[best]
    info.update({"work_order": best, "value": vox * steps / wall if steps else 0.0, "unit": "voxel-updates/s",
                 "ms_per_step": 1e3 * wall / max(steps, 1), "box_voxels_per_step": vox, "steps": steps,
                 "ms_per_step_other_order": 1e3 * timing[other][0] / max(steps, 1)})
    if accuracy_steps:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from fp32_error import phi_volume_errors, trajectories, trajectory_error
        a32 = new_context(G.FP32_BATCHED, inputs, p)
        a64 = new_context(G.FP64_PLAIN, inputs, p)
        for c in (a32, a64):
            c.set_tuning(choice)
            c.set_comparison_order(corder, cshape)
            c.set_work_order(G.ORDER_VOLUME if best == "volume-sorted" else G.ORDER_SPATIAL)
        tr = trajectories(a32, a64, accuracy_steps, every=10)
        e = phi_volume_errors(a32, a64, p.delta, grains=np.arange(1, grains + 1, 50))
        e.update(trajectory_error(a32.count_history(), a64.count_history()))
        e["steps"] = accuracy_steps
        e["reference"] = "FP64 GPU path (bit-identical to the FP64 oracle)"
        e["grains_sampled_for_phi"] = "every 50th grain id"
        e.update(tr)
        info["fp32_accuracy_vs_fp64"] = e
        a32.close()
        a64.close()
    return info


def container_leg(ctx):
    """One voxel container (saveNetworkAsVoxelContainer's id grid) of the context's state: the
    kernel clears size^3 ids and scatters every alive grain's phi >= 0 voxels; algorithmic bytes =
    4 B per box voxel read + 4 B per container voxel cleared and written once."""
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
    """One grid coarsening (gridCoarsement's coarsening branch) of the context's state onto a grid
    frac times the size, then one timestep on it.  The resampling kernel reads every alive grain's
    old box and writes its new box: 4 B per voxel each way (FP32) is its algorithmic traffic."""
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


if __name__ == "__main__":
    main()
