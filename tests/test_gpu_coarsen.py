"""Grid coarsening (gg_coarsen_grid) on the GPU against the oracle: FP64
bit-exact after the coarsening (resampled field, IDLocal, bounds, states,
volumes) and over the timesteps after it, eagerly and through a recaptured
graph; the case with grains that vanished just before the coarsening; FP32
against FP64.

Each case prints its name, seed and shape before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from oracle import oracle as O
from parity import compare_all, make_pair, step_both

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _new_grid(p, frac):
    new = int((p.ngridpoints - 2 * p.grid_blowup) * frac)
    return new, 0.2 / (new * new)


@pytest.mark.parametrize("n_grains,ppg,border,seed,frac", [(64, 10, 3, 5, 0.75), (27, 14, 4, 2, 0.6)])
def test_coarsen_bitexact_then_steps(n_grains, ppg, border, seed, frac):
    _say("coarsen n=%d ppg=%d border=%d seed=%d frac=%.2f" % (n_grains, ppg, border, seed, frac))
    ctx, orc, p = make_pair(n_grains, ppg, border, seed)
    for s in range(3):
        step_both(ctx, orc)
    new, dt = _new_grid(p, frac)
    ctx.coarsen_grid(new, dt)
    orc.coarsen(new, dt)
    n = compare_all(ctx, orc, "after coarsening")
    assert n > 0
    for s in range(3):
        step_both(ctx, orc)
        compare_all(ctx, orc, "step %d after coarsening" % s)
    ctx.close()


def test_coarsen_then_graph_timesteps():
    _say("coarsen then graph n=64 ppg=10 border=3 seed=8")
    ctx, orc, p = make_pair(64, 10, 3, 8)
    ctx.timestep(2, use_graph=True)  # a graph captured on the old grid
    orc.timestep(2)
    new, dt = _new_grid(p, 0.8)
    ctx.coarsen_grid(new, dt)
    orc.coarsen(new, dt)
    for _ in range(2):
        ctx.timestep(2, use_graph=True)
        ctx.synchronize()
    counts = orc.timestep(4)
    compare_all(ctx, orc, "graph after coarsening")
    np.testing.assert_array_equal(ctx.count_history()[-4:], counts)
    ctx.close()


def test_coarsen_with_vanished_grains():
    _say("coarsen with vanished grains n=64 ppg=6 border=3 seed=3")
    ctx, orc, p = make_pair(64, 6, 3, 3)
    for s in range(60):
        step_both(ctx, orc)
        if (orc.states() == O.VANISHED).any():
            break
    assert (orc.states() == O.VANISHED).any(), "no grain vanished; the case does not exercise the rule"
    compare_all(ctx, orc, "before coarsening")
    new, dt = _new_grid(p, 0.85)
    ctx.coarsen_grid(new, dt)
    orc.coarsen(new, dt)
    compare_all(ctx, orc, "after coarsening")
    for s in range(4):
        step_both(ctx, orc)
    compare_all(ctx, orc, "4 steps after coarsening")
    ctx.close()


def test_coarsen_fp32_follows_fp64():
    _say("coarsen fp32 vs fp64 n=64 ppg=10 border=3 seed=5")
    c64, _, p = make_pair(64, 10, 3, 5)
    c32, _, _ = make_pair(64, 10, 3, 5, precision=G.FP32_BATCHED)
    for c in (c64, c32):
        c.timestep(3)
    new, dt = _new_grid(p, 0.75)
    for c in (c64, c32):
        c.coarsen_grid(new, dt)
    s64, v64, _ = c64.stats()
    s32, v32, _ = c32.stats()
    both = np.nonzero((s64 == G.ALIVE) & (s32 == G.ALIVE))[0]
    assert both.size > 50
    worst = 0.0
    for g in both:
        b64, f64 = c64.phi(g, 1)
        b32, f32 = c32.phi(g, 1)
        if not np.array_equal(b64, b32):
            continue
        worst = max(worst, float(np.abs(f64 - f32).max()))
    # FP32 distances (|phi| <= delta) after three steps and one resampling
    assert worst < 1e-5
    np.testing.assert_allclose(v32[both], v64[both], rtol=1e-4)
    c64.close()
    c32.close()


def test_two_coarsenings_in_one_run():
    _say("two coarsenings n=64 ppg=12 border=3 seed=4")
    ctx, orc, p = make_pair(64, 12, 3, 4)
    ctx.timestep(2, use_graph=True)
    orc.timestep(2)
    size = p.ngridpoints - 2 * p.grid_blowup
    for frac in (0.85, 0.8):
        size = int(size * frac)
        ctx.coarsen_grid(size, 0.2 / (size * size))
        orc.coarsen(size, 0.2 / (size * size))
        compare_all(ctx, orc, "after coarsening to %d" % size)
        ctx.timestep(2, use_graph=True)
        orc.timestep(2)
        compare_all(ctx, orc, "2 graph steps on %d" % size)
    assert ctx.params.ngridpoints == size + 2 * p.grid_blowup
    ctx.close()
