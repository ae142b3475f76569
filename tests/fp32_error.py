"""Error of the FP32 batched path against the FP64 path (which is bit-identical to
the FP64 oracle), shared by the GPU tests and bench.py.

Reported, as the north-star asks: per-voxel |phi32 - phi64| max and RMS inside
the band |phi64| < delta, per-grain volume relative error, grain-count
trajectories.
"""
import numpy as np

import gragles_3d_amd as G


def phi_volume_errors(ctx32, ctx64, delta, grains=None):
    st32, v32, _ = ctx32.stats()
    st64, v64, _ = ctx64.stats()
    both = np.nonzero((st32 == G.ALIVE) & (st64 == G.ALIVE))[0]
    if grains is not None:
        both = np.intersect1d(both, grains)
    sq = 0.0
    cnt = 0
    mx = 0.0
    same_bounds = 0
    for g in both:
        b32 = ctx32.bounds(int(g), 1)
        b64 = ctx64.bounds(int(g), 1)
        if not np.array_equal(b32, b64):
            continue
        same_bounds += 1
        _, p32 = ctx32.phi(int(g), 1)
        _, p64 = ctx64.phi(int(g), 1)
        band = np.abs(p64) < delta
        d = np.abs(p32[band] - p64[band])
        if d.size:
            mx = max(mx, float(d.max()))
            sq += float((d * d).sum())
            cnt += int(d.size)
    vrel = np.abs(v32[both] - v64[both]) / np.maximum(np.abs(v64[both]), 1e-300)
    return {
        "phi_max_abs_err_in_band": mx,
        "phi_rms_err_in_band": (sq / max(cnt, 1)) ** 0.5,
        "phi_max_err_over_delta": mx / delta,
        "band_voxels_compared": cnt,
        "grains_compared": int(len(both)),
        "grains_with_equal_boxes": same_bounds,
        "volume_rel_err_max": float(vrel.max()) if vrel.size else 0.0,
        "volume_rel_err_median": float(np.median(vrel)) if vrel.size else 0.0,
    }


def trajectory_error(c32, c64):
    c32 = np.asarray(c32, dtype=np.float64)
    c64 = np.asarray(c64, dtype=np.float64)
    n = min(len(c32), len(c64))
    rel = np.abs(c32[:n] - c64[:n]) / np.maximum(c64[:n], 1)
    return {"count_rel_err_max": float(rel.max()) if n else 0.0, "final_count_fp32": int(c32[n - 1]) if n else 0,
            "final_count_fp64": int(c64[n - 1]) if n else 0, "steps": int(n)}
