"""Hull topology (the grains each marching-cubes hull touches) of the FP64 GPU
path vs the oracle, grain by grain and in the reference's order, and the
grain-boundary faces (get_Faces: pairs and areas).

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


def _compare_faces(ctx, orc, h, where):
    # get_Faces (src/grainHull.cpp:1330-1338): per alive grain its boundary
    # lists with a smaller neighbour id, area = summed triangle areas * h^2
    st = orc.states()
    want = []
    for g in np.nonzero(st == G.ALIVE)[0]:
        els, areas = orc.hull_elements(g), orc.hull_element_areas(g)
        faces = sorted((e[0][0] if e[0][1] == g else e[0][1], a * h * h) for e, a in zip(els, areas)
                       if len(e[0]) == 2 and g in e[0] and g > min(e[0]))
        want += [(int(g), int(nb), a) for nb, a in faces]
    ga, gb, ar = ctx.faces()
    assert [(int(a), int(b)) for a, b in zip(ga, gb)] == [(w[0], w[1]) for w in want], where
    np.testing.assert_allclose(ar, [w[2] for w in want], rtol=1e-10, atol=0, err_msg=where)
    return len(want)


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
        assert _compare_faces(ctx, orc, p.h, "faces, step %d" % s) > 0
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
    # the element positions are summed in fixed point: eager steps give the
    # same bits as graph steps, whatever order the tiles added in
    ctx2, _, _ = make_pair(64, 10, 3, 7)
    ctx2.enable_hull_topology(True)
    ctx2.timestep(4, use_graph=False)
    st, _, _ = ctx.stats()
    for g in range(1, len(st)):
        if st[g] == G.ALIVE:
            a, b = ctx.hull_elements(g), ctx2.hull_elements(g)
            assert len(a) == len(b)
            for ea, eb in zip(a, b):
                assert ea[0] == eb[0] and ea[1] == eb[1] and np.array_equal(ea[2], eb[2]), (g, ea, eb)
    ctx.close()
    ctx2.close()


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


def test_hull_topology_large_boxes():
    # box sides 39-82: the hull's n <= 64 and n <= 96 instantiations (the FP64
    # n <= 96 class adds element sums straight to the grain's table)
    _say("hull topology large boxes n=27 ppg=40 border=3 seed=11")
    ctx, orc, p = make_pair(27, 40, 3, 11)
    ctx.enable_hull_topology(True)
    for s in range(3):
        step_both(ctx, orc)
        assert _compare(ctx, orc, "step %d" % s) > 0
    ctx.close()


@pytest.mark.parametrize("n_grains,ppg,border,seed", [(64, 10, 3, 5), (27, 40, 3, 11)])
def test_hull_topology_fp32_matches_oracle_on_the_same_field(n_grains, ppg, border, seed):
    # The FP32 path's topology against the oracle's topology routines run on
    # the FP32 field the hull read (its values are exact doubles): the same
    # inside/outside corners, so ids, their order and the junction counts
    # must be equal; element positions differ by the FP32 interpolation factor
    # only (atol 1e-4 grid units).
    from oracle import oracle as O
    _say("hull topology FP32 n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    ctx, _, p = make_pair(n_grains, ppg, border, seed, precision=G.FP32_BATCHED)
    ctx.enable_hull_topology(True)
    checked = 0
    for s in range(3):
        for name in ("switch", "convolution", "switch", "update_lists", "comparison", "switch", "level_set"):
            getattr(ctx, name)()
        st, _, _ = ctx.stats()
        junc = ctx.hull_junction_counts()
        for g in range(1, n_grains + 1):
            if st[g] != G.ALIVE:
                continue
            b, phi = ctx.phi(g, 0)
            n = int(b[3] - b[0])
            ib, idl = ctx.idlocal(g)
            args = (phi.reshape(n, n, n), b[:3], ib, idl, g)
            want, want_j = O.hull_neighbours_box(*args, junctions=True)
            assert np.array_equal(ctx.hull_neighbours(g), want), (s, g)
            assert np.array_equal(junc[g], want_j), (s, g, junc[g], want_j)
            eg, eo = ctx.hull_elements(g), O.hull_elements_box(*args)
            assert [(e[0], e[1]) for e in eg] == [(e[0], e[1]) for e in eo], (s, g)
            for a, c in zip(eg, eo):
                np.testing.assert_allclose(a[2], c[2], rtol=0, atol=1e-4, err_msg="step %d grain %d" % (s, g))
            checked += 1
        ctx.redistancing()
    assert checked > 0
    ctx.close()
