"""Box geometry, redistancing and hull reductions of the oracle, pinned by hand-derived
cases and closed forms (CPU)."""
import numpy as np
import pytest

import gragles_3d_amd as G
from oracle import oracle as O


# ---- resizeToCube (src/dimensionalBuffer.h:139-219), cases worked by hand -----------
CUBE_CASES = [
    # bounds {xmin,ymin,zmin,xmax,ymax,zmax}, L, expected
    ((0, 0, 0, 10, 12, 14), 100, (0, 0, 0, 14, 14, 14)),        # grow x by 4 (2/2), y by 2 (1/1), shift up
    ((90, 90, 90, 99, 98, 97), 100, (90, 90, 89, 99, 99, 98)),  # y grows 0 below/1 above, z 1/1
    ((95, 0, 0, 100, 3, 4), 100, (94, 0, 0, 99, 5, 5)),         # x pushed back below L-1, y shifted up
    ((10, 20, 30, 15, 25, 35), 100, (10, 20, 30, 15, 25, 35)),  # already a cube: untouched
    ((5, 5, 5, 12, 8, 6), 50, (5, 3, 2, 12, 10, 9)),            # odd diffs: y 3 -> 1 below 2 above; z 6 -> 3/3
]


@pytest.mark.parametrize("b,L,exp", CUBE_CASES)
def test_resize_to_cube_cases(b, L, exp):
    assert tuple(O.resize_to_cube(b, L)) == exp
    assert tuple(G.resize_to_cube(b, L)) == exp


def test_resize_to_cube_product_matches_oracle_random():
    rng = np.random.default_rng(4)
    for _ in range(2000):
        lo = rng.integers(0, 60, 3)
        ext = rng.integers(1, 40, 3)
        b = np.concatenate([lo, np.minimum(lo + ext, 100)])
        a = O.resize_to_cube(b, 100)
        c = G.resize_to_cube(b, 100)
        assert np.array_equal(a, c)
        assert a[3] - a[0] == a[4] - a[1] == a[5] - a[2]


# ---- redistancing (src/box.cpp:2322-2586) -----------------------------------------
def test_redistancing_planar_interface_exact():
    # sign-only field of the plane x = j0 + 1/2: the sweeps give the exact signed distance
    n, h, delta = 16, 0.01, 0.05
    j0 = 7
    b = np.array([10, 20, 30, 10 + n, 20 + n, 30 + n], dtype=np.int32)
    x = np.arange(n)
    phi = np.where(x <= j0, h, -h)[None, None, :] * np.ones((n, n, 1))
    out = O.redistance_box(b, phi, b, h, delta)
    d = (j0 + 0.5 - x) * h
    exact = np.clip(d, -delta, delta)
    # layers k < izmax-1 get the x/y sweeps (src/box.cpp:2366); the last layer only
    # receives the z-forward step from its neighbour, one h further from zero
    np.testing.assert_allclose(out[:-1], np.broadcast_to(exact, (n - 1, n, n)), atol=1e-15)
    last = np.clip(d + np.sign(d) * h, -delta, delta)
    np.testing.assert_allclose(out[-1], np.broadcast_to(last, (n, n)), atol=1e-15)


def test_redistancing_planar_interface_along_z_and_y():
    n, h, delta = 12, 0.02, 0.1
    b = np.array([0, 0, 0, n, n, n], dtype=np.int32)
    k = np.arange(n)
    phi = np.where(k < 5, h, -h)[:, None, None] * np.ones((1, n, n))
    out = O.redistance_box(b, phi, b, h, delta)
    exact = np.clip((4.5 - k) * h, -delta, delta)
    np.testing.assert_allclose(out, np.broadcast_to(exact[:, None, None], (n, n, n)), atol=1e-15)
    # plane along y: the last z-layer again takes only the z-forward step
    phi_y = np.transpose(phi, (1, 0, 2))
    out_y = O.redistance_box(b, phi_y, b, h, delta)
    np.testing.assert_allclose(out_y[:-1], np.broadcast_to(exact[None, :, None], (n - 1, n, n)), atol=1e-15)
    d = (4.5 - k) * h
    np.testing.assert_allclose(out_y[-1], np.broadcast_to(np.clip(d + np.sign(d) * h, -delta, delta)[:, None],
                                                          (n, n)), atol=1e-15)


def test_redistancing_sphere_is_a_distance_function():
    n, h = 41, 1.0 / 40  # odd n: the centre (n-1)/2 is a grid point
    delta = 6 * h
    b = np.array([0, 0, 0, n, n, n], dtype=np.int32)
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    R = 11.3
    phi = np.clip((R - r) * h, -delta, delta)  # a clamped signed distance, as between timesteps
    out = O.redistance_box(b, phi, b, h, delta)
    exact = (R - r) * h
    band = np.abs(exact) < 3 * h
    # the sweeps measure distance along grid-line paths: never much shorter than the
    # Euclidean distance, never longer than the L1 bound sqrt(3)*|d| (+ the sub-cell
    # interpolation of the first step)
    ao, ae = np.abs(out[band]), np.abs(exact[band])
    assert np.all(ao >= ae - 0.6 * h)
    assert np.all(ao <= np.sqrt(3) * ae + 0.6 * h)
    # along the axes through the centre the path is straight: exact to rounding
    ci = int(round(c))
    line = out[ci, ci, :]
    ref = np.clip((R - np.abs(np.arange(n) - c)) * h, -delta, delta)
    near = np.abs(ref) < 2 * h
    assert np.abs(line[near] - ref[near]).max() < 0.02 * h
    assert np.all(out <= delta) and np.all(out >= -delta)
    # sign preserved wherever the input was not at the interface
    assert np.all(np.sign(out[np.abs(exact) > h]) == np.sign(exact[np.abs(exact) > h]))


def test_redistancing_output_box_larger_than_input():
    # the output box grows by two planes in -x: rows outside the input keep -1 until the
    # x-backward sweep reaches them, exactly the reference's partial coverage
    n, h, delta = 10, 0.05, 0.25
    ib = np.array([2, 0, 0, 2 + n, n, n], dtype=np.int32)
    ob = np.array([0, 0, 0, n, n, n], dtype=np.int32)
    x = np.arange(2, 2 + n)
    phi = np.where(x < 6, h, -h)[None, None, :] * np.ones((n, n, 1))
    out = O.redistance_box(ib, phi, ob, h, delta)
    xo = np.arange(0, n)
    d = (5.5 - xo) * h
    exact = np.clip(d, -delta, delta)
    # inside the input box (x >= 2) every layer below the last is exact
    np.testing.assert_allclose(out[:-1, :, 2:], np.broadcast_to(exact[2:], (n - 1, n, n - 2)), atol=1e-15)
    # the two new columns (x = 1, 0) are reached only by x-backward's outside-branch,
    # cand = out(j) + sgn(out(j-1))*h with out(j-1) = -1 (src/box.cpp:2426-2432): each step
    # moves h towards -1, whatever the sign of the field
    np.testing.assert_allclose(out[:-1, :, 1], out[:-1, :, 2] - h, atol=1e-15)
    np.testing.assert_allclose(out[:-1, :, 0], out[:-1, :, 1] - h, atol=1e-15)


# ---- hull reductions (marching cubes, src/marchingCubes.cpp:364-499, src/grainHull.cpp:40-131)
def test_hull_all_outside_does_not_exist():
    ex, v, a = O.hull_box(-np.ones((8, 8, 8)), (3, 4, 5), 0.1)
    assert not ex and v == 0.0 and a == 0.0


def test_hull_half_space_area():
    # plane z = 3.5 cuts every cell of one layer: (n-1)^2 unit squares
    n, h = 9, 0.01
    z = np.arange(n)
    phi = np.broadcast_to((3.5 - z)[:, None, None], (n, n, n)).copy()
    ex, v, a = O.hull_box(phi, (0, 0, 0), h)
    assert ex
    assert a == pytest.approx((n - 1) ** 2 * h * h, rel=1e-14)


@pytest.mark.parametrize("R", [6.3, 9.7, 14.2])
def test_hull_sphere_volume_and_area_converge(R):
    n = int(2 * R) + 6
    h = 1.0 / 100
    c = (n - 1) / 2.0 + 0.123
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    phi = R - np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    ex, v, a = O.hull_box(phi, (0, 0, 0), h)
    assert ex
    assert v == pytest.approx(4 / 3 * np.pi * (R * h) ** 3, rel=3.0 / R ** 2)
    assert a == pytest.approx(4 * np.pi * (R * h) ** 2, rel=6.0 / R)


def test_hull_volume_translation_invariant():
    # the signed-tetrahedron sum of a closed surface does not depend on the origin;
    # a sign or index slip in any of its six terms breaks this
    n = 20
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    phi = 6.5 - np.sqrt((x - 9.3) ** 2 + 0.7 * (y - 10.1) ** 2 + 1.3 * (z - 9.8) ** 2)
    vols = [O.hull_box(phi, org, 1.0)[1] for org in [(0, 0, 0), (37, 5, 11), (200, 300, 150)]]
    assert vols[1] == pytest.approx(vols[0], rel=1e-9)
    assert vols[2] == pytest.approx(vols[0], rel=1e-9)


def test_triangle_table_consistent():
    from gragles_3d_amd import levelset  # noqa: F401
    corner_edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for c in range(256):
        mask, row = O.tri_row(c)
        exp_mask = 0
        for e, (p, q) in enumerate(corner_edges):
            if ((c >> p) & 1) != ((c >> q) & 1):
                exp_mask |= 1 << e
        assert mask == exp_mask
        tri = row[row >= 0]
        assert tri.size % 3 == 0
        assert (tri.size == 0) == (mask == 0)
        for e in tri:
            assert mask & (1 << int(e)), (c, e)
        # complementary cases cut the same edges
        assert O.tri_row(255 - c)[0] == mask
