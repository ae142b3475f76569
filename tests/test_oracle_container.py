"""The oracle's voxel container (saveNetworkAsVoxelContainer with
copyDataToContainer, src/grainhdl.cpp:1074-1122, src/box.cpp:3008-3032) and
the .uds line format (src/grainhdl.cpp:1088-1107), on the CPU.

Pins: right after the distance initialisation every grain's phi >= 0 region is
its Voronoi cell, so the container equals the generator's label field inside
the border, voxel for voxel (a transposed axis, a missing border shift or a
wrong sign fails it); after timesteps every container id is a grain whose
output buffer is >= 0 at that voxel; the .uds line of a hand-worked grain.
"""
import numpy as np

import gragles_3d_amd as G
from oracle import oracle as O


def _sim(n, ppg, border, seed):
    p = G.derive_params(n, ppg, border)
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    s = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, n)
    s.set_boxes(bounds, exists)
    s.init_distance(idf)
    s.set_lists(offs, ids)
    return s, p, idf


def test_container_after_init_is_the_label_field():
    s, p, idf = _sim(64, 10, 3, 5)
    c = s.voxel_container()
    L, gb, n = p.conv_domain_size, p.grid_blowup, p.ngridpoints
    assert c.shape == (L, L, L)
    ref = idf.reshape(n, n, n)[gb:gb + L, gb:gb + L, gb:gb + L]
    np.testing.assert_array_equal(c, ref.astype(np.uint32))


def test_container_ids_have_nonnegative_phi_after_steps():
    s, p, _ = _sim(27, 12, 3, 11)
    s.timestep(3)
    c = s.voxel_container()
    gb = p.grid_blowup
    st = s.states()
    alive = set(np.nonzero(st == O.ALIVE)[0].tolist())
    assert set(np.unique(c).tolist()) - {0} <= alive
    assert (c == 0).mean() < 0.01
    for g in list(alive)[:8]:
        b, phi = s.phi(g, 1)
        phi = phi.reshape(b[5] - b[2], b[4] - b[1], b[3] - b[0])
        zz, yy, xx = np.nonzero(c == g)
        local = phi[zz + gb - b[2], yy + gb - b[1], xx + gb - b[0]]
        assert local.size and (local >= 0).all()


def test_uds_line_by_hand():
    txt = G.uds_text([5], np.array([[3, 4, 5, 13, 16, 15]]), 0.1, [0.0123456789])
    assert txt == "5\t0.8\t1\t1\t0\t0\t0\t3\t13\t4\t16\t5\t15\t0.0123457\t0\n"
    txt = G.uds_text([7], np.array([[0, 0, 0, 3, 3, 3]]), 0.5, [2.0], bunge=[[0.1, 0.25, 3.0]], stored=[1e-13])
    assert txt == "7\t0.5\t0.5\t0.5\t0.1\t0.25\t3\t0\t3\t0\t3\t0\t3\t2\t1e-13\n"
