"""The split-BF16 (3-term) convolution of the FP32 batched path: the three
circulant passes as matrix products on the matrix cores with BF16 hi/lo
operand parts, accumulated in FP32.

Checked against the FP32 circulant on the same input, stage by stage (its
error is about 2^-16 of the summed |c| |phi| per product, so 1e-4 delta is a
loose bound that a wrong operand slot, row or column misses by orders of
magnitude), and against the FP64 oracle over whole steps: equal grain counts,
per-voxel band error and per-grain volume error within the north-star bound.
Sizes cover the four size classes.  Each case prints its name before touching
the GPU.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
from fp32_error import OracleView, phi_volume_errors
from parity import make_pair

pytestmark = pytest.mark.gpu

CASES = [(64, 10, 3, 5), (27, 40, 3, 11)]  # box sides 14-30 and 39-82


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n_grains,ppg,border,seed", CASES)
def test_split_bf16_convolution_near_fp32_circulant(n_grains, ppg, border, seed):
    _say("split-BF16 convolution vs FP32 circulant n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    a, orc, p = make_pair(n_grains, ppg, border, seed, precision=G.FP32_BATCHED)
    b, _, _ = make_pair(n_grains, ppg, border, seed, precision=G.FP32_BATCHED)
    b.set_convolution_form(G.CONV_SPLIT_BF16)
    for ctx in (a, b):
        ctx.switch()
        ctx.convolution()
    st, _, _ = a.stats()
    worst = 0.0
    for g in np.nonzero(st == G.ALIVE)[0]:
        ba, pa = a.phi(g, 1)
        bb, pb = b.phi(g, 1)
        assert np.array_equal(ba, bb)
        worst = max(worst, float(np.abs(pa - pb).max()))
    _say("max |split-BF16 - FP32| after one convolution: %.3e (delta %.3e)" % (worst, p.delta))
    assert 0 < worst < 1e-4 * p.delta
    a.close()
    b.close()


@pytest.mark.parametrize("n_grains,ppg,border,seed", CASES)
def test_split_bf16_steps_against_oracle(n_grains, ppg, border, seed):
    _say("split-BF16 vs oracle, 4 graph steps, n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    ctx, orc, p = make_pair(n_grains, ppg, border, seed, precision=G.FP32_BATCHED)
    ctx.set_convolution_form(G.CONV_SPLIT_BF16)
    ctx.timestep(4, use_graph=True)
    counts = orc.timestep(4)
    np.testing.assert_array_equal(ctx.count_history(), counts)
    e = phi_volume_errors(ctx, OracleView(orc), p.delta)
    _say(str(e))
    assert e["grains_with_equal_boxes"] >= 0.9 * e["grains_compared"]
    assert e["phi_max_abs_err_in_band"] < 1e-3 * p.delta
    assert e["volume_rel_err_max"] < 1e-3
    ctx.close()


def test_split_bf16_is_refused_on_the_fp64_path():
    _say("split-BF16 on an FP64 context")
    ctx, orc, p = make_pair(8, 8, 3, 1)
    with pytest.raises(G.GGError, match="FP32 batched"):
        ctx.set_convolution_form(G.CONV_SPLIT_BF16)
    ctx.close()
