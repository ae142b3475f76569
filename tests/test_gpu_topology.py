"""Hull topology (the grains each marching-cubes hull touches) of the FP64 GPU
path vs the oracle, grain by grain and in the reference's order.

Each case prints its name, seed and shape before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from parity import make_pair, step_both

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(ctx, orc, where):
    cnt, wall = ctx.hull_neighbour_counts()
    junc = ctx.hull_junction_counts()
    st, _, _ = ctx.stats()
    checked = 0
    for g in range(1, len(st)):
        want_j = orc.hull_junctions(g)
        if not np.array_equal(junc[g], want_j):
            raise AssertionError("%s: grain %d junction counts %s, oracle %s" % (where, g, list(junc[g]), list(want_j)))
        want = orc.hull_neighbours(g)
        got = ctx.hull_neighbours(g)
        if not np.array_equal(got, want):
            raise AssertionError("%s: grain %d hull neighbours %s, oracle %s" % (where, g, list(got), list(want)))
        assert cnt[g] == len(want) and wall[g] == (0 in want), (where, g)
        if st[g] == G.ALIVE:
            eg, eo = ctx.hull_elements(g), orc.hull_elements(g)
            assert [(e[0], e[1]) for e in eg] == [(e[0], e[1]) for e in eo], (where, g, eg[:3], eo[:3])
            for a, b in zip(eg, eo):
                np.testing.assert_allclose(a[2], b[2], rtol=0, atol=1e-9, err_msg="%s grain %d" % (where, g))
            assert sum(1 for e in eg if len(e[0]) == 2) == junc[g][0]
        if st[g] == G.ALIVE:
            assert g in want, (where, g)  # a hull that exists touches its own grain
            checked += 1
    return checked


@pytest.mark.parametrize("n_grains,ppg,border,seed", [(64, 10, 3, 5), (64, 6, 3, 3)])
def test_hull_topology_matches_oracle(n_grains, ppg, border, seed):
    _say("hull topology n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    ctx, orc, p = make_pair(n_grains, ppg, border, seed)
    ctx.enable_hull_topology(True)
    walls = 0
    junctions = np.zeros(4, dtype=np.int64)
    for s in range(6):
        step_both(ctx, orc)
        assert _compare(ctx, orc, "step %d" % s) > 0
        walls += int(ctx.hull_neighbour_counts()[1].sum())
        junctions += ctx.hull_junction_counts().sum(axis=0)
    assert walls > 0, "no hull touched the domain wall; the case does not exercise id 0"
    assert junctions[0] > 0 and junctions[1] > 0 and junctions[2] > 0, junctions
    ctx.close()


def test_hull_topology_in_graph_steps():
    _say("hull topology graph n=64 ppg=10 border=3 seed=7")
    ctx, orc, p = make_pair(64, 10, 3, 7)
    ctx.enable_hull_topology(True)
    ctx.timestep(4, use_graph=True)
    orc.timestep(4)
    assert _compare(ctx, orc, "after 4 graph steps") > 0
    ctx.close()


def test_hull_topology_of_a_grid_coarsening():
    # extractContour in the coarsening branch computes the topology of the
    # resampled field, against the IDLocal just recalculated
    _say("hull topology coarsening n=64 ppg=10 border=3 seed=5")
    ctx, orc, p = make_pair(64, 10, 3, 5)
    ctx.enable_hull_topology(True)
    for s in range(2):
        step_both(ctx, orc)
    new = int((p.ngridpoints - 2 * p.grid_blowup) * 0.8)
    ctx.coarsen_grid(new, 0.2 / (new * new))
    orc.coarsen(new, 0.2 / (new * new))
    assert _compare(ctx, orc, "after coarsening") > 0
    step_both(ctx, orc)
    assert _compare(ctx, orc, "step after coarsening") > 0
    ctx.close()
