"""The oracle's hull topology (grains the marching-cubes hull touches) on
hand-built boxes whose answers follow from the reference's rules
(src/marchingCubes.cpp:501-602, src/box.cpp:3000-3006): inside corners name
the grain itself, outside corners the IDLocal id there (or nothing outside the
IDLocal box), each cut cell's ids are taken ascending, and the result keeps
first appearances in the (z, y, x) cell order."""
import numpy as np

from oracle import oracle as O

N = 8
ORIGIN = (10, 20, 30)
BOX = np.array([10, 20, 30, 18, 28, 38], dtype=np.int32)


def half_space():
    phi = -np.ones((N, N, N))
    phi[:, :, :4] = 1.0  # inside for local x < 4 (arrays are [z][y][x])
    return phi


def test_half_space_names_self_and_neighbour_ascending():
    idl = np.full(N ** 3, 7, dtype=np.uint32)
    ids, junc = O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl, 3, junctions=True)
    assert list(ids) == [3, 7]
    assert list(junc) == [1, 0, 0, 0]  # one grain boundary, no junction
    assert list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl, 9)) == [7, 9]


def test_all_outside_touches_nothing():
    idl = np.full(N ** 3, 7, dtype=np.uint32)
    assert len(O.hull_neighbours_box(-np.ones((N, N, N)), ORIGIN, BOX, idl, 3)) == 0


def test_triple_line_order_of_first_appearance():
    # outside the grain, IDLocal is 5 for local y < 4 and 6 above: the first cut
    # cells (y = 0) see {3, 5}; cells straddling y = 3..4 add 6 later
    idl = np.empty((N, N, N), dtype=np.uint32)
    idl[:, :4, :] = 5
    idl[:, 4:, :] = 6
    got, junc = O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl.ravel(), 3, junctions=True)
    assert list(got) == [3, 5, 6]
    # lists {3,5}, {3,6} (grain boundaries) and {3,5,6} (the triple line)
    assert list(junc) == [2, 1, 0, 0]
    # with the ids swapped the first cells see {3, 6}: order follows the cells, not the ids
    idl2 = np.where(idl == 5, 6, 5).astype(np.uint32)
    assert list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl2.ravel(), 3)) == [3, 6, 5]


def test_quadruple_and_higher_junctions():
    # three outside regions meeting the grain's face: 11 for y < 4, and for
    # y >= 4 12 below z = 4 and 13 above; the cell at the meeting point sees the
    # grain and all three (4 ids: a quadruple junction)
    idl = np.empty((N, N, N), dtype=np.uint32)
    idl[:, :4, :] = 11
    idl[:4, 4:, :] = 12
    idl[4:, 4:, :] = 13
    ids, junc = O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl.ravel(), 3, junctions=True)
    assert sorted(ids) == [3, 11, 12, 13]
    assert list(junc) == [3, 3, 1, 0]  # {3,11} {3,12} {3,13}; three triple lines; the point
    # four quadrants meeting: the centre cell sees five grains (a higher-order junction)
    idl[:4, :4, :] = 11
    idl[:4, 4:, :] = 12
    idl[4:, :4, :] = 13
    idl[4:, 4:, :] = 14
    ids, junc = O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl.ravel(), 3, junctions=True)
    assert sorted(ids) == [3, 11, 12, 13, 14]
    assert list(junc) == [4, 4, 0, 1]


def test_wall_id_zero_is_a_neighbour():
    idl = np.full((N, N, N), 7, dtype=np.uint32)
    idl[N - 1, :, :] = 0  # the top layer of the outside region reads the wall
    got = list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl.ravel(), 3))
    assert got == [3, 7, 0]


def test_corners_outside_the_idlocal_box_name_nothing():
    # IDLocal covers global x >= 15 only; the cut cells lie at global x 13..14,
    # so their outside corners (x = 14) fall outside it and only the grain remains
    b6 = np.array([15, 20, 30, 18, 28, 38], dtype=np.int32)
    idl = np.full(3 * N * N, 7, dtype=np.uint32)
    ids, junc = O.hull_neighbours_box(half_space(), ORIGIN, b6, idl, 3, junctions=True)
    assert list(ids) == [3]
    assert list(junc) == [0, 0, 0, 1]  # the one-id list {3} falls to the reference's default branch


def test_half_space_element_position_is_the_patch_centre():
    # one boundary element; the plane x = 3.5 (local) is cut into two triangles
    # per cell, whose barycenters average to the cell centre, so the element's
    # position is the centre of the (N-1)^2 cells and it has 2 (N-1)^2 triangles
    idl = np.full(N ** 3, 7, dtype=np.uint32)
    el = O.hull_elements_box(half_space(), ORIGIN, BOX, idl, 3)
    assert len(el) == 1
    ids, ntri, pos = el[0]
    assert ids == (3, 7) and ntri == 2 * (N - 1) ** 2
    np.testing.assert_allclose(pos, (ORIGIN[0] + 3.5, ORIGIN[1] + (N - 1) / 2, ORIGIN[2] + (N - 1) / 2),
                               rtol=0, atol=1e-12)


def test_elements_match_the_junction_classes():
    # the elements are the distinct lists the junction counts classify
    phi = -np.ones((N, N, N))
    phi[:, :, :4] = 1.0
    idl = np.full(N ** 3, 5, dtype=np.uint32).reshape(N, N, N)
    idl[:, 4:, :] = 6  # a second neighbour above local y = 4: a triple line along x = 3.5
    ids, junc = O.hull_neighbours_box(phi, ORIGIN, BOX, idl.reshape(-1), 3, junctions=True)
    el = O.hull_elements_box(phi, ORIGIN, BOX, idl.reshape(-1), 3)
    sizes = [len(e[0]) for e in el]
    assert [sizes.count(2), sizes.count(3)] == list(junc[:2])
    assert sum(e[1] for e in el) == 2 * (N - 1) ** 2


def test_boundary_face_area_closed_forms():
    # GrainBoundary::computeArea (src/InterfacialElement.cpp:81-92): the
    # half-space's one boundary element is the flat (N-1) x (N-1) patch at
    # x = 3.5, area (N-1)^2 in grid units; the oracle's whole-hull surface of
    # the same box is the same number
    idl = np.full(N ** 3, 7, dtype=np.uint32)
    el = O.hull_elements_box(half_space(), ORIGIN, BOX, idl, 3, with_area=True)
    assert len(el) == 1 and abs(el[0][3] - (N - 1) ** 2) < 1e-9
    # a tilted plane x + y = const: area grows by sqrt(2) against its x
    # projection, and the elements of two neighbours split it (their areas add
    # up to the whole patch)
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    phi = 4.3 - (x + y) / np.sqrt(2.0)
    idl2 = np.where(z < 4, 5, 6).astype(np.uint32).reshape(-1)
    el2 = O.hull_elements_box(phi, ORIGIN, BOX, idl2, 3, with_area=True)
    _, _, area_all = hull_area_box(phi)
    assert abs(sum(e[3] for e in el2) - area_all) < 1e-9 * area_all
    gb = {e[0]: e[3] for e in el2 if len(e[0]) == 2}
    assert set(gb) == {(3, 5), (3, 6)}


def hull_area_box(phi):
    # the oracle's whole-hull reduction on the same box (exists, volume, area; h = 1: grid units)
    return O.hull_box(phi, ORIGIN, 1.0)
