"""The oracle's comparison step (executeComparison + boundaryCondition +
executeSetComparison, src/box.cpp:814-1088) on hand-derived cases (CPU)."""
import numpy as np

from oracle import oracle as O

NGP, GB = 40, 3
H = 1.0 / 40
DELTA = 3 * H


def field(bounds, f):
    x0, y0, z0, x1, y1, z1 = bounds
    z, y, x = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
    return f(x, y, z).astype(np.float64)


def test_two_half_spaces_keep_the_field_and_name_the_neighbour():
    b = [10, 10, 10, 20, 20, 20]
    phi1 = field(b, lambda x, y, z: np.clip((14.5 - x) * H, -DELTA, DELTA))
    out, idl, nb = O.comparison_box(NGP, GB, H, DELTA, b, phi1, [(5, b, -phi1)], np.zeros(1000))
    # in the band 0.5*(phi1 - (-phi1)) = phi1, outside it phi1 is copied
    np.testing.assert_array_equal(out, phi1)
    assert np.all(idl == 5)
    # bbox of phi >= 0 is x in [10,14], y,z in [10,19]; min - gb, max + gb
    assert list(nb) == [7, 7, 7, 17, 22, 22]


def test_ties_go_to_the_first_neighbour_in_list_order():
    b = [10, 10, 10, 16, 16, 16]
    phi = field(b, lambda x, y, z: (x - 12.5) * H)
    same = field(b, lambda x, y, z: 0.3 * H + 0 * x)
    out1, idl1, _ = O.comparison_box(NGP, GB, H, DELTA, b, phi, [(7, b, same), (3, b, same)], np.zeros(216))
    out2, idl2, _ = O.comparison_box(NGP, GB, H, DELTA, b, phi, [(3, b, same), (7, b, same)], np.zeros(216))
    assert np.all(idl1 == 7) and np.all(idl2 == 3)
    np.testing.assert_array_equal(out1, out2)
    # a strictly larger later value does win
    big = same + H
    _, idl3, _ = O.comparison_box(NGP, GB, H, DELTA, b, phi, [(7, b, same), (3, b, big)], np.zeros(216))
    assert np.all(idl3 == 3)


def test_uncovered_voxels_keep_stale_ids_and_see_minus_one():
    b = [10, 10, 10, 16, 16, 16]
    nbb = [10, 10, 10, 13, 16, 16]  # covers x in [10, 13)
    phi = field(b, lambda x, y, z: 0.2 * H + 0 * x)  # inside the band everywhere
    nphi = field(nbb, lambda x, y, z: -0.4 * H + 0 * x)
    stale = np.arange(216, dtype=np.uint32) + 1000
    out, idl, _ = O.comparison_box(NGP, GB, H, DELTA, b, phi, [(9, nbb, nphi)], stale)
    x = field(b, lambda x, y, z: x).ravel()
    covered = x < 13
    assert np.all(idl.ravel()[covered] == 9)
    np.testing.assert_array_equal(idl.ravel()[~covered], stale[~covered])
    o = out.ravel()
    np.testing.assert_allclose(o[covered], 0.5 * (0.2 * H - (-0.4 * H)), rtol=0, atol=1e-18)
    np.testing.assert_allclose(o[~covered], 0.5 * (0.2 * H - (-1.0)), rtol=0, atol=1e-17)


def test_outside_the_band_the_input_is_copied():
    b = [10, 10, 10, 14, 14, 14]
    phi = field(b, lambda x, y, z: np.where(x % 2 == 0, 0.8 * DELTA, -0.7 * DELTA))
    out, _, _ = O.comparison_box(NGP, GB, H, DELTA, b, phi, [(2, b, phi * 0 + 0.5 * H)], np.zeros(64))
    np.testing.assert_array_equal(out, phi)


def test_wall_distances_by_hand():
    # grid of 20 points, border 3: the domain spans j = 2 .. 16 for the wall rule
    ngp, gb, h = 20, 3, 0.05
    delta = 3 * h
    b = [0, 0, 0, 8, 8, 8]
    phi = np.zeros((8, 8, 8))  # inside the band: out = 0.5*(0 - best) = -best/2
    out, idl, _ = O.comparison_box(ngp, gb, h, delta, b, phi, [], np.full(512, 77, dtype=np.uint32))

    def at(k, i, j):
        return out[k, i, j], idl[k * 64 + i * 8 + j]

    # (5,5,5): every axis distance -3, dist = -3 is not > -gb: no wall value, best stays -1
    assert at(5, 5, 5) == (0.5, 77)
    # (5,5,4): dist = max(-2, -3, -3) = -2 > -3
    v, i = at(5, 5, 4)
    assert i == 0 and v == -0.5 * (-2 * h)
    # (5,5,2): dX = 0, dY = dZ = -3 -> dist = dX = 0
    v, i = at(5, 5, 2)
    assert i == 0 and v == 0.0
    # (5,5,0): dX = 2 -> dist = 2
    assert at(5, 5, 0) == (-0.5 * 2 * h, 0)
    # (5,1,1): dX = dY = 1, dZ < 0 -> sqrt(2)
    assert at(5, 1, 1) == (-0.5 * (np.sqrt(2.0) * h), 0)
    # (1,1,1): corner -> sqrt(3)
    assert at(1, 1, 1) == (-0.5 * (np.sqrt(3.0) * h), 0)
    # (2,2,2): all axis distances 0 -> no branch assigns dist; nor at (2,2,1) (dX=1, dY=dZ=0)
    # or (2,2,0); the last assignment in (k,i,j) order is at (2,1,7): dX=-5, dY=1, dZ=0 ->
    # sqrt(dY^2 + dZ^2) = 1
    assert at(2, 2, 2) == (-0.5 * h, 0)
    assert at(2, 2, 1) == (-0.5 * h, 0)
    assert at(2, 2, 0) == (-0.5 * h, 0)
