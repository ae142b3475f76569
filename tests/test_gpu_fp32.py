"""FP32 batched path: error against the FP64 path (bit-identical to the FP64 oracle).

Tolerances are the north-star's where it states one (per-grain volume <= 1e-3
relative, grain count within 0.5%), and for phi a bound derived from FP32
rounding of values of size delta over the steps run.
"""
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from fp32_error import phi_volume_errors, trajectory_error

pytestmark = pytest.mark.gpu


def _pair(n, ppg, border, seed):
    ctx64, p, inp = G.setup_simulation(n, ppg, border, seed, precision=G.FP64_PLAIN)
    ctx32 = G.LevelSetContext(p, precision=G.FP32_BATCHED)
    ctx32.set_boxes(inp["bounds"], inp["exists"])
    ctx32.init_distance(inp["idfield"])
    ctx32.set_neighbour_lists(inp["offsets"], inp["ids"])
    return ctx32, ctx64, p


def test_fp32_initial_field_matches_fp64():
    print("fp32 init n=125 ppg=10 border=3 seed=4", flush=True)
    ctx32, ctx64, p = _pair(125, 10, 3, 4)
    e = phi_volume_errors(ctx32, ctx64, p.delta)
    assert e["grains_with_equal_boxes"] == e["grains_compared"] == 125
    # the initial field is +-h swept by +-h steps: exact in FP32 up to the rounding of h
    assert e["phi_max_abs_err_in_band"] <= 8 * np.finfo(np.float32).eps * p.delta


def test_fp32_errors_over_steps():
    print("fp32 vs fp64 n=216 ppg=10 border=3 seed=8, 20 steps", flush=True)
    ctx32, ctx64, p = _pair(216, 10, 3, 8)
    ctx32.timestep(20, use_graph=True)
    ctx64.timestep(20, use_graph=True)
    e = phi_volume_errors(ctx32, ctx64, p.delta)
    t = trajectory_error(ctx32.count_history(), ctx64.count_history())
    print(e, t, flush=True)
    assert e["grains_with_equal_boxes"] >= 0.95 * e["grains_compared"]
    assert e["phi_rms_err_in_band"] < 1e-4 * p.delta
    assert e["volume_rel_err_median"] < 1e-4
    assert e["volume_rel_err_max"] < 1e-3
    assert t["count_rel_err_max"] <= 0.005


@pytest.mark.parametrize("n,ppg,border,seed", [(512, 16, 7, 1)])
def test_fp32_trajectory_512_grains(n, ppg, border, seed):
    print("fp32 trajectory n=%d ppg=%d border=%d seed=%d, 50 steps" % (n, ppg, border, seed), flush=True)
    ctx32, ctx64, p = _pair(n, ppg, border, seed)
    ctx32.timestep(50, use_graph=True)
    ctx64.timestep(50, use_graph=True)
    t = trajectory_error(ctx32.count_history(), ctx64.count_history())
    e = phi_volume_errors(ctx32, ctx64, p.delta, grains=np.arange(1, n + 1, 7))
    print(e, t, flush=True)
    assert t["count_rel_err_max"] <= 0.005
    assert e["volume_rel_err_median"] < 1e-3
