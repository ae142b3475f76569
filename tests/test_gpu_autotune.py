"""The comparison's tile shapes and the autotuner that picks one per box-size
class: every shape gives the same result (FP64 bit for bit against the
oracle, FP32 bit for bit against the default shape), and a timestep taken by
gg_autotune equals an ordinary one.

Each case prints its name, seed and shape before touching the GPU.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
from parity import compare_all, make_pair

pytestmark = pytest.mark.gpu

SHAPES = range(G.levelset.CMP_VARIANTS)
CASE = (27, 40, 3, 11)  # box sides 39-82: classes n <= 48, 64 and 96


@pytest.mark.parametrize("shape", SHAPES)
def test_every_shape_is_bitexact_fp64(shape):
    print("comparison shape %d FP64 n=%d ppg=%d border=%d seed=%d, 2 steps" % ((shape,) + CASE), flush=True)
    ctx, orc, p = make_pair(*CASE)
    ctx.set_tuning([shape] * 4)
    assert list(ctx.tuning()) == [shape] * 4
    ctx.timestep(2, use_graph=True)
    orc.timestep(2)
    compare_all(ctx, orc, "shape %d after 2 steps" % shape)
    ctx.close()


def test_every_shape_gives_the_same_fp32_result():
    print("comparison shapes FP32 n=64 ppg=10 border=3 seed=5 and n=27 ppg=40, 2 steps", flush=True)
    for case in [(64, 10, 3, 5), CASE]:
        ref = None
        for shape in SHAPES:
            ctx, p, _ = G.setup_simulation(*case, precision=G.FP32_BATCHED)
            ctx.set_tuning([shape] * 4)
            ctx.timestep(2, use_graph=True)
            st, vol, _ = ctx.stats()
            phis = [ctx.phi(g, 1)[1] for g in range(1, case[0] + 1) if st[g] == G.ALIVE]
            if ref is None:
                ref = (st, vol, phis)
            else:
                assert np.array_equal(st, ref[0]) and np.array_equal(vol, ref[1])
                for a, b in zip(phis, ref[2]):
                    assert np.array_equal(a, b)
            ctx.close()


def test_autotuned_step_equals_a_timestep():
    print("autotune n=%d ppg=%d border=%d seed=%d" % CASE, flush=True)
    ctx, orc, p = make_pair(*CASE)
    choice, ms = ctx.autotune(reps=2)
    assert list(ctx.tuning()) == list(choice)
    assert all(0 <= v < G.levelset.CMP_VARIANTS for v in choice)
    assert (ms > 0).all()
    orc.timestep(1)
    compare_all(ctx, orc, "after the autotuned step")
    ctx.timestep(2, use_graph=True)  # then ordinary steps with the chosen shapes
    orc.timestep(2)
    compare_all(ctx, orc, "two steps later")
    np.testing.assert_array_equal(ctx.count_history(), [orc.count()] * 3)
    with pytest.raises(G.GGError, match="out of range"):
        ctx.set_tuning([0, 1, 2, G.levelset.CMP_VARIANTS])
    ctx.close()
