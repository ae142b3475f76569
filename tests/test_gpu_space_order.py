"""The comparison in spatial order (gg_set_comparison_order): one launch for
all size classes whose wavefront tiles are sorted by the cell of their centre,
with the work list rebuilt on the device at every comparison.  It computes the
same voxels as the per-class launches, so FP64 stays bit for bit equal to the
oracle (every shape, stage by stage, graph replays, vanishing grains) and FP32
bit for bit equal to the grain order, with winner ids packed (fewer than 0xFFFF
grains) and unpacked.

Each case prints its name, seed and shape before touching the GPU.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
from parity import compare_all, make_pair, step_both

pytestmark = pytest.mark.gpu

CASE = (27, 40, 3, 11)  # box sides 39-82: classes n <= 48, 64 and 96


@pytest.mark.parametrize("shape", range(4))
def test_every_shape_is_bitexact_fp64(shape):
    print("spatial comparison shape %d FP64 n=%d ppg=%d border=%d seed=%d, 2 steps" % ((shape,) + CASE), flush=True)
    ctx, orc, p = make_pair(*CASE)
    ctx.set_comparison_order(G.CMP_ORDER_SPACE, shape)
    assert ctx.comparison_order() == (G.CMP_ORDER_SPACE, shape)
    ctx.timestep(2, use_graph=True)
    orc.timestep(2)
    compare_all(ctx, orc, "spatial shape %d after 2 steps" % shape)
    with pytest.raises(G.GGError, match="shape out of range"):
        ctx.set_comparison_order(G.CMP_ORDER_SPACE, 4)
    ctx.close()


def test_stagewise_bitexact_fp64():
    print("spatial comparison stagewise FP64 n=64 ppg=10 border=7 seed=5", flush=True)
    ctx, orc, p = make_pair(64, 10, 7, 5)
    ctx.set_comparison_order(G.CMP_ORDER_SPACE)
    for s in range(3):
        step_both(ctx, orc, compare_every_stage=True, tag="step %d" % s)
    ctx.close()


def test_many_steps_with_vanishing_fp64():
    print("spatial comparison many steps FP64 n=64 ppg=6 border=3 seed=3", flush=True)
    ctx, orc, p = make_pair(64, 6, 3, 3)
    ctx.set_comparison_order(G.CMP_ORDER_SPACE)
    ctx.timestep(30, use_graph=True)
    counts = orc.timestep(30)
    compare_all(ctx, orc, "after 30 steps")
    np.testing.assert_array_equal(ctx.count_history(), counts)
    assert counts[-1] < 64, "no grain vanished; the case does not exercise deletion"
    ctx.close()


# 70000 grains: ids above 0xFFFF, so the spatial kernel keeps its winner ids
# unpacked (below that it packs them as 16-bit halves)
@pytest.mark.parametrize("case", [(64, 10, 3, 5), CASE, (1000, 12, 3, 2), (8000, 20, 3, 2), (70000, 6, 3, 2)])
def test_fp32_equals_grain_order(case):
    print("spatial vs grain order FP32 n=%d ppg=%d border=%d seed=%d, 3 steps" % case, flush=True)
    res = []
    for order in (G.CMP_ORDER_GRAIN, G.CMP_ORDER_SPACE):
        ctx, p, _ = G.setup_simulation(*case, precision=G.FP32_BATCHED)
        ctx.set_comparison_order(order)
        ctx.timestep(3, use_graph=True)
        st, vol, _ = ctx.stats()
        alive = [g for g in range(1, case[0] + 1) if st[g] == G.ALIVE]
        alive = alive[::16] if case[0] > 2000 else alive  # every 16th grain of the large case (host memory)
        res.append((st, vol, [ctx.phi(g, 1) for g in alive], [ctx.idlocal(g) for g in alive]))
        ctx.close()
    sa, va, pa, ia = res[0]
    for sb, vb, pb, ib in res[1:]:
        assert np.array_equal(sa, sb) and np.array_equal(va, vb)
        for (b1, x1), (b2, x2) in zip(pa, pb):
            assert np.array_equal(b1, b2) and np.array_equal(x1, x2)
        for (b1, x1), (b2, x2) in zip(ia, ib):
            assert np.array_equal(b1, b2) and np.array_equal(x1, x2)


def test_autotune_in_spatial_order():
    print("autotune spatial order n=%d ppg=%d border=%d seed=%d" % CASE, flush=True)
    ctx, orc, p = make_pair(*CASE)
    ctx.set_comparison_order(G.CMP_ORDER_SPACE)
    choice, ms = ctx.autotune(reps=2)
    order, shape = ctx.comparison_order()
    assert order == G.CMP_ORDER_SPACE and list(choice) == [shape] * 4 and 0 <= shape < 4
    assert (ms[0, :4] > 0).all()
    orc.timestep(1)
    compare_all(ctx, orc, "after the autotuned step")
    ctx.timestep(2, use_graph=True)
    orc.timestep(2)
    compare_all(ctx, orc, "two steps later")
    with pytest.raises(G.GGError, match="unknown comparison order"):
        ctx.set_comparison_order(2)
    ctx.close()
