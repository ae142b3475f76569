"""The 8000-grain workload at full size (seeded Voronoi, seed 2, 20 points per
grain, border 3, grid 329^3; boxes of side 26-55):

* two timesteps of the FP64 GPU path bit for bit against the oracle (every
  grain's state, bounds, distance buffer, IDLocal and neighbour list), and of
  the FP32 path against the oracle within tolerances;
* 200 timesteps of the FP32 path beside the FP64 GPU path (the oracle's
  stand-in: bit-identical to it above and over 50 steps at 512 grains), with
  the grain-count and mean-grain-size trajectories and per-grain volume errors.

Each case prints its name, seed and shape before touching the GPU.
"""
import json
import os

import numpy as np
import pytest

import gragles_3d_amd as G
from fp32_error import OracleView, phi_volume_errors, trajectories, trajectory_error
from parity import compare_all
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG = (8000, 20, 3, 2)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)


def _inputs():
    n, ppg, border, seed = CFG
    p = G.derive_params(n, ppg, border)
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    return p, dict(idfield=idf, bounds=bounds, exists=exists, offsets=offs, ids=ids)


def _ctx(p, inp, precision):
    ctx = G.LevelSetContext(p, precision=precision)
    ctx.set_boxes(inp["bounds"], inp["exists"])
    ctx.init_distance(inp["idfield"])
    ctx.set_neighbour_lists(inp["offsets"], inp["ids"])
    return ctx


def test_config2_two_steps_against_oracle():
    _say("8000 grains ppg=20 border=3 seed=2: FP64 bit-exact and FP32 vs oracle, 2 steps")
    p, inp = _inputs()
    orc = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, p.n_grains)
    orc.set_boxes(inp["bounds"], inp["exists"])
    orc.init_distance(inp["idfield"])
    orc.set_lists(inp["offsets"], inp["ids"])
    ctx64 = _ctx(p, inp, G.FP64_PLAIN)
    ctx32 = _ctx(p, inp, G.FP32_BATCHED)
    ctx64.timestep(2, use_graph=True)
    ctx32.timestep(2, use_graph=True)
    counts = orc.timestep(2)
    np.testing.assert_array_equal(ctx64.count_history(), counts)
    n = compare_all(ctx64, orc, "FP64 after 2 steps")
    assert n > 7900
    e = phi_volume_errors(ctx32, OracleView(orc), p.delta, grains=np.arange(1, p.n_grains + 1, 4))
    _say("FP32 vs oracle after 2 steps: %s" % e)
    assert e["grains_with_equal_boxes"] >= 0.98 * e["grains_compared"]
    assert e["phi_max_abs_err_in_band"] < 1e-4 * p.delta
    assert e["phi_rms_err_in_band"] < 1e-5 * p.delta
    assert e["volume_rel_err_max"] < 1e-4
    np.testing.assert_array_equal(ctx32.count_history(), counts)
    ctx64.close()
    ctx32.close()
    orc.close()


def test_config2_fp32_200_steps_trajectories():
    _say("8000 grains ppg=20 border=3 seed=2: FP32 vs FP64, 200 steps")
    p, inp = _inputs()
    ctx64 = _ctx(p, inp, G.FP64_PLAIN)
    ctx32 = _ctx(p, inp, G.FP32_BATCHED)
    tr = trajectories(ctx32, ctx64, 200, every=10)
    c32, c64 = ctx32.count_history(), ctx64.count_history()
    t = trajectory_error(c32, c64)
    e = phi_volume_errors(ctx32, ctx64, p.delta, grains=np.arange(1, p.n_grains + 1, 8))
    rel_size = np.abs(np.array(tr["mean_volume_fp32"]) / np.array(tr["mean_volume_fp64"]) - 1)
    summary = {"count_fp32": c32.tolist(), "count_fp64": c64.tolist(), "trajectory_error": t,
               "errors_after_200_steps": e, "mean_size": tr, "mean_volume_rel_err_max": float(rel_size.max())}
    _say(json.dumps(summary))
    out = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "config2_fp32_200_steps.json"), "w") as f:
            json.dump(summary, f)
    assert len(c32) == len(c64) == 200
    assert c64[-1] < 8000, "no grain vanished in 200 steps; the trajectory is vacuous"
    assert t["count_rel_err_max"] <= 0.005
    assert rel_size.max() < 1e-3
    assert np.median(tr["volume_rel_err_median"]) < 1e-4
    ctx64.close()
    ctx32.close()
