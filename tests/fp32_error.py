"""Error of the FP32 batched path against an FP64 reference: the FP64 oracle
itself (OracleView) or the FP64 GPU path (bit-identical to the oracle), shared
by the GPU tests and bench.py.

Reported, as the north-star asks: per-voxel |phi32 - phi64| max and RMS inside
the band |phi64| < delta, per-grain volume relative error, grain-count and
mean-grain-size trajectories.
"""
import numpy as np

import gragles_3d_amd as G


class OracleView:
    """The FP64 oracle with the inspection surface of a LevelSetContext
    (stats, bounds, phi) that phi_volume_errors reads."""

    def __init__(self, orc):
        self.orc = orc

    def stats(self):
        v, a, _ = self.orc.stats()
        return self.orc.states(), v, a

    def bounds(self, gid, which):
        return self.orc.bounds(gid, which)[0]

    def phi(self, gid, which):
        return self.orc.phi(gid, which)


def phi_volume_errors(ctx32, ref, delta, grains=None):
    """ctx32: the FP32 context; ref: the FP64 reference (a LevelSetContext or an
    OracleView).  Distance buffers are compared where both sides have the same
    output box (an FP32 rounding can move a box bound by one plane)."""
    st32, v32, _ = ctx32.stats()
    st64, v64, _ = ref.stats()
    both = np.nonzero((st32 == G.ALIVE) & (st64 == G.ALIVE))[0]
    if grains is not None:
        both = np.intersect1d(both, grains)
    sq = 0.0
    cnt = 0
    mx = 0.0
    same_bounds = 0
    for g in both:
        b32 = ctx32.bounds(int(g), 1)
        b64 = ref.bounds(int(g), 1)
        if not np.array_equal(b32, b64):
            continue
        same_bounds += 1
        _, p32 = ctx32.phi(int(g), 1)
        _, p64 = ref.phi(int(g), 1)
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


def mean_grain_size(ctx):
    """Mean volume of the alive grains and the mean radius of equal-volume
    spheres (grid units of the unit cube)."""
    st, v, _ = ctx.stats()
    alive = st == G.ALIVE
    vol = v[alive]
    if vol.size == 0:
        return 0.0, 0.0
    return float(vol.mean()), float(np.mean(np.cbrt(3.0 * vol / (4.0 * np.pi))))


def trajectories(ctx32, ctx64, steps, every=1):
    """Runs both contexts `steps` eager timesteps side by side and records, every
    `every` steps, the mean grain size of each and the per-grain volume error of
    the grains alive in both.  Returns a dict of lists."""
    out = {"step": [], "mean_volume_fp32": [], "mean_volume_fp64": [], "mean_radius_fp32": [],
           "mean_radius_fp64": [], "volume_rel_err_max": [], "volume_rel_err_median": [],
           "grains_above_1e-3": [], "largest_voxels_above_1e-3": [], "volume_rel_err_max_from_100_voxels": []}
    h3 = float(ctx64.params.h) ** 3
    done = 0
    while done < steps:
        k = min(every, steps - done)
        ctx32.timestep(k, use_graph=False)
        ctx64.timestep(k, use_graph=False)
        done += k
        v32, r32 = mean_grain_size(ctx32)
        v64, r64 = mean_grain_size(ctx64)
        st32, vol32, _ = ctx32.stats()
        st64, vol64, _ = ctx64.stats()
        both = (st32 == G.ALIVE) & (st64 == G.ALIVE)
        rel = np.abs(vol32[both] - vol64[both]) / np.maximum(np.abs(vol64[both]), 1e-300)
        out["step"].append(done)
        out["mean_volume_fp32"].append(v32)
        out["mean_volume_fp64"].append(v64)
        out["mean_radius_fp32"].append(r32)
        out["mean_radius_fp64"].append(r64)
        out["volume_rel_err_max"].append(float(rel.max()) if rel.size else 0.0)
        out["volume_rel_err_median"].append(float(np.median(rel)) if rel.size else 0.0)
        vox = vol64[both] / h3  # volume in voxels
        above = rel > 1e-3
        out["grains_above_1e-3"].append(int(above.sum()))
        out["largest_voxels_above_1e-3"].append(float(vox[above].max()) if above.any() else 0.0)
        big = vox >= 100.0
        out["volume_rel_err_max_from_100_voxels"].append(float(rel[big].max()) if big.any() else 0.0)
    return out
