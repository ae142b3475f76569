"""The uneven-batch configuration (10 % of the grains four times larger,
lognormal sizes, power-diagram microstructure) at a reduced grain count: the
FP64 path against the oracle bit for bit with the volume-sorted work queue
(stage by stage) and with the spatial order (whole steps), graph replays of
the volume-sorted step, and the FP32 path against the oracle within
tolerances.  The box sides span the four size classes of the work plan.

Each case prints its name, seed and shape before touching the GPU.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
import physics_cases as PC
from fp32_error import OracleView, phi_volume_errors
from parity import compare_all, step_both
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASE = (300, 28, 3, 4)  # grains, points per grain, border, seed: box sides 14-73, all four classes


def _say(msg):
    print(msg, flush=True)


def _pair(order, precision=G.FP64_PLAIN):
    case = PC.bimodal_case(*CASE)
    ctx = PC.gpu_context(case, precision)
    ctx.set_work_order(order)
    assert ctx.work_order() == order
    return ctx, PC.oracle_sim(O, case), case[0]


def test_uneven_boxes_volume_order_stagewise_bitexact():
    _say("uneven batch, volume-sorted queue, stagewise: grains %d ppg %d border %d seed %d" % CASE)
    ctx, orc, p = _pair(G.ORDER_VOLUME)
    sides = np.array([orc.bounds(g, 1)[0][3] - orc.bounds(g, 1)[0][0] for g in range(1, CASE[0] + 1)
                      if orc.state(g) == O.ALIVE])
    assert sides.min() <= 32 and sides.max() > 64, (sides.min(), sides.max())
    assert ((sides > 32) & (sides <= 48)).any() and ((sides > 48) & (sides <= 64)).any()
    compare_all(ctx, orc, "after init")
    for k in range(3):
        step_both(ctx, orc, compare_every_stage=True, tag="step %d" % k)
    ctx.close()


def test_uneven_boxes_spatial_order_and_graph_replay_bitexact():
    _say("uneven batch, spatial order eager + volume order graph, 4 steps: grains %d ppg %d border %d seed %d"
         % CASE)
    a, orc, p = _pair(G.ORDER_SPATIAL)
    b, _, _ = _pair(G.ORDER_VOLUME)
    a.timestep(4, use_graph=False)
    for _ in range(2):  # replays with eager work and synchronisations between them
        b.timestep(2, use_graph=True)
        b.synchronize()
    counts = orc.timestep(4)
    np.testing.assert_array_equal(a.count_history(), counts)
    np.testing.assert_array_equal(b.count_history(), counts)
    compare_all(a, orc, "spatial order after 4 steps")
    compare_all(b, orc, "volume order, graph, after 4 steps")
    a.close()
    b.close()


def test_uneven_boxes_fp32_against_oracle():
    _say("uneven batch FP32 vs oracle, volume order, 4 steps: grains %d ppg %d border %d seed %d" % CASE)
    ctx32, orc, p = _pair(G.ORDER_VOLUME, G.FP32_BATCHED)
    ctx32.timestep(4, use_graph=True)
    counts = orc.timestep(4)
    np.testing.assert_array_equal(ctx32.count_history(), counts)
    e = phi_volume_errors(ctx32, OracleView(orc), p.delta)
    _say(str(e))
    assert e["grains_with_equal_boxes"] >= 0.9 * e["grains_compared"]
    assert e["phi_max_abs_err_in_band"] < 1e-4 * p.delta
    assert e["volume_rel_err_max"] < 1e-4
    ctx32.close()
