"""gg_init_distance_noalloc (the entry point that allocates no device memory)
against gg_init_distance: every grain's initial distance buffer must be the same
bit for bit, in both precisions, on two calls in a row, the second on a smaller
grid (the shared scratch then holds the first call's data beyond what the
second one writes).  The first grid needs several layer runs through the
scratch (176^3 > 4 Mi entries)."""
import numpy as np
import pytest

import gragles_3d_amd as G

pytestmark = pytest.mark.gpu

CALLS = [(343, 30, 3, 4), (27, 12, 3, 11)]  # grid 176^3 (two layer runs), then 43^3


@pytest.mark.parametrize("precision", [G.FP64_PLAIN, G.FP32_BATCHED])
def test_init_distance_noalloc_equals_init_distance(precision):
    for n, ppg, border, seed in CALLS:
        print("no-alloc init n=%d ppg=%d border=%d seed=%d precision=%d" % (n, ppg, border, seed, precision),
              flush=True)
        p = G.derive_params(n, ppg, border)
        idf, bounds, exists = G.generate_voronoi(p, seed)
        a = G.LevelSetContext(p, precision=precision)
        a.set_boxes(bounds, exists)
        a.init_distance(idf)
        b = G.LevelSetContext(p, precision=precision)
        b.set_boxes(bounds, exists)
        b.init_distance_noalloc(idf)
        st, _, _ = a.stats()
        alive = np.nonzero(st == G.ALIVE)[0]
        assert len(alive) > 0.9 * n
        for g in alive:
            for which in (0, 1):
                ba, pa = a.phi(int(g), which)
                bb, pb = b.phi(int(g), which)
                assert np.array_equal(ba, bb)
                if which == 1:  # the redistanced field (the input buffer is stale after the redistancing)
                    assert np.array_equal(pa, pb), (n, int(g))
        a.close()
        b.close()
