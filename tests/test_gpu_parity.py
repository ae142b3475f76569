"""FP64 plain GPU path vs the FP64 CPU oracle: bit-exact, stage by stage.

Each case prints its name, seed and shape before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from parity import compare_all, make_pair, step_both

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n_grains,ppg,border,seed", [(27, 12, 3, 11), (64, 10, 7, 5)])
def test_stagewise_bitexact(n_grains, ppg, border, seed):
    _say("stagewise n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    ctx, orc, p = make_pair(n_grains, ppg, border, seed)
    compare_all(ctx, orc, "after init", check_phi=(1,), check_idlocal=True, check_lists=True)
    for s in range(3):
        step_both(ctx, orc, compare_every_stage=True, tag="step %d" % s)
    ctx.close()


def test_many_steps_with_vanishing():
    # small grains, many steps: grains vanish and get deleted, lists shrink
    _say("many steps n=64 ppg=6 border=3 seed=3")
    ctx, orc, p = make_pair(64, 6, 3, 3)
    counts_o = []
    for s in range(40):
        step_both(ctx, orc)
        counts_o.append(orc.count())
        if s % 10 == 9:
            compare_all(ctx, orc, "step %d" % s)
    compare_all(ctx, orc, "final")
    st, _, _ = ctx.stats()
    assert (st != G.DELETED).sum() == counts_o[-1]
    assert counts_o[-1] < 64, "no grain vanished; the case does not exercise deletion"
    ctx.close()


def test_timestep_entry_matches_oracle():
    _say("timestep entry n=125 ppg=8 border=3 seed=7")
    ctx, orc, p = make_pair(125, 8, 3, 7)
    ctx.timestep(6, use_graph=False)
    counts_o = orc.timestep(6)
    compare_all(ctx, orc, "after 6 timesteps")
    np.testing.assert_array_equal(ctx.count_history(), counts_o)
    ctx.close()


def test_graph_replay_matches_eager():
    _say("graph replay n=125 ppg=8 border=3 seed=9")
    ctx_a, orc, p = make_pair(125, 8, 3, 9)
    ctx_b, _, _ = make_pair(125, 8, 3, 9)
    ctx_a.timestep(7, use_graph=False)
    for _ in range(3):  # several replays with eager work and syncs between them
        ctx_b.timestep(2, use_graph=True)
        ctx_b.synchronize()
    ctx_b.timestep(1, use_graph=True)
    orc.timestep(7)
    compare_all(ctx_a, orc, "eager after 7 steps")
    compare_all(ctx_b, orc, "graph after 7 steps")
    ctx_a.close()
    ctx_b.close()


def test_bulk_getters_match_per_grain_getters():
    _say("bulk getters n=27 ppg=12 border=3 seed=11")
    ctx, orc, p = make_pair(27, 12, 3, 11)
    step_both(ctx, orc)
    for which in (0, 1):
        allb = ctx.all_bounds(which)
        offs, ids = ctx.all_neighbour_lists(which)
        assert offs[0] == 0 and offs[1] == 0
        for g in range(1, p.n_grains + 1):
            assert np.array_equal(allb[g], ctx.bounds(g, which))
            assert np.array_equal(ids[offs[g]:offs[g + 1]], ctx.neighbour_list(g, which))
    ctx.close()


def test_512_grains_50_steps_bitexact():
    # the CPU-runnable configuration: 512 grains, 16 points per grain, border 7
    # (grid 118^3), 50 timesteps; GPU steps replayed from the captured graph
    _say("512 grains ppg=16 border=7 seed=1, 50 steps (graph) vs oracle")
    ctx, orc, p = make_pair(512, 16, 7, 1)
    ctx.timestep(50, use_graph=True)
    counts_o = orc.timestep(50)
    np.testing.assert_array_equal(ctx.count_history(), counts_o)
    compare_all(ctx, orc, "after 50 steps")
    ctx.close()


def test_list_capacity_overflow_is_reported():
    # Initial lists cut to their first two entries and a capacity of two: the
    # list build adds the neighbours-of-neighbours whose boxes intersect, which
    # exceeds the capacity, and the step must report it (not truncate silently).
    _say("list overflow n=64 ppg=10 border=3 seed=5 max_list=2")
    p = G.derive_params(64, 10, 3)
    idf, bounds, exists = G.generate_voronoi(p, 5)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    lens = np.diff(offs)
    assert lens.max() > 2
    short = [ids[offs[g]:offs[g] + min(2, lens[g])] for g in range(len(lens))]
    soffs = np.concatenate([[0], np.cumsum([len(x) for x in short])]).astype(np.int32)
    sids = np.concatenate(short).astype(np.int32)
    ctx = G.LevelSetContext(p, precision=G.FP64_PLAIN, max_list=2)
    ctx.set_boxes(bounds, exists)
    ctx.init_distance(idf)
    ctx.set_neighbour_lists(soffs, sids)
    ctx.timestep(1, use_graph=False)  # asynchronous: the device error surfaces at the next synchronisation
    with pytest.raises(G.GGError, match="comparison list capacity exceeded"):
        ctx.synchronize()
    ctx.close()


@pytest.mark.parametrize("n_grains,ppg,border,seed", [(64, 10, 3, 5), (512, 16, 7, 1)])
def test_device_initial_lists_match_host(n_grains, ppg, border, seed):
    _say("device initial lists n=%d ppg=%d border=%d seed=%d" % (n_grains, ppg, border, seed))
    p = G.derive_params(n_grains, ppg, border)
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    ctx = G.LevelSetContext(p, precision=G.FP64_PLAIN)
    ctx.set_boxes(bounds, exists)
    ctx.build_neighbour_lists()
    d_offs, d_ids = ctx.all_neighbour_lists(0)
    np.testing.assert_array_equal(d_offs, offs)
    np.testing.assert_array_equal(d_ids, ids)
    ctx.close()


def test_redistancing_order_is_enforced():
    # gg_redistancing writes the output boxes the last level set laid out; a
    # second call (or one after a coarsening) has none and is refused
    _say("redistancing order n=27 ppg=12 border=3 seed=11")
    ctx, orc, p = make_pair(27, 12, 3, 11)
    step_both(ctx, orc)
    with pytest.raises(G.GGError, match="needs the output boxes"):
        ctx.redistancing()
    new = int((p.ngridpoints - 2 * p.grid_blowup) * 0.9)
    ctx.coarsen_grid(new, 0.2 / (new * new))
    with pytest.raises(G.GGError, match="needs the output boxes"):
        ctx.redistancing()
    ctx.timestep(1)  # the normal order still runs
    ctx.synchronize()
    ctx.close()


@pytest.mark.parametrize("precision", [G.FP64_PLAIN, G.FP32_BATCHED])
def test_upload_phi_round_trip(precision):
    _say("upload phi n=27 ppg=12 border=3 seed=11 precision=%d" % precision)
    ctx, orc, p = make_pair(27, 12, 3, 11, precision=precision)
    rng = np.random.default_rng(5)
    for g in (1, 7, 27):
        b, phi = ctx.phi(g, 1)
        new = rng.uniform(-p.delta, p.delta, phi.size)
        ctx.upload_phi(g, 1, new)
        _, back = ctx.phi(g, 1)
        want = new.astype(np.float32).astype(np.float64) if precision == G.FP32_BATCHED else new
        assert np.array_equal(back, want)
        with pytest.raises(G.GGError, match="element count"):
            ctx.upload_phi(g, 1, new[:-1])
    # the uploaded fields are the ones the next step convolves: a context that
    # received the other's fields continues bit-identically
    ctx2, _, _ = make_pair(27, 12, 3, 11, precision=precision)
    for g in range(1, 28):
        if ctx.stats()[0][g] == G.ALIVE:
            ctx2.upload_phi(g, 1, ctx.phi(g, 1)[1])
    ctx.timestep(2)
    ctx2.timestep(2)
    for g in range(1, 28):
        if ctx.stats()[0][g] == G.ALIVE:
            assert np.array_equal(ctx.phi(g, 1)[1], ctx2.phi(g, 1)[1]), g
    ctx.close()
    ctx2.close()


def test_second_set_boxes_stays_within_the_first_arena():
    # the first gg_set_boxes sizes the arenas; larger boxes later are refused
    _say("set_boxes twice n=27 ppg=12 border=3 seed=11")
    p = G.derive_params(27, 12, 3)
    idf, bounds, exists = G.generate_voronoi(p, 11)
    ctx = G.LevelSetContext(p, precision=G.FP64_PLAIN)
    b = np.array(bounds, dtype=np.int32).reshape(-1, 6).copy()
    small = b.copy()
    small[1:, 3:] = small[1:, :3] + 4  # boxes of side 4
    ctx.set_boxes(small.ravel(), exists)
    ctx.set_boxes(small.ravel(), exists)  # the same again fits
    big = b.copy()
    big[1:, :3] = 0
    big[1:, 3:] = min(p.ngridpoints, 96)  # every box the whole grid (sides up to 96)
    with pytest.raises(G.GGError, match="exceed the arena"):
        ctx.set_boxes(big.ravel(), exists)
    ctx.close()
