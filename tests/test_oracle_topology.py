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
    assert list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl, 3)) == [3, 7]
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
    got = list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl.ravel(), 3))
    assert got == [3, 5, 6]
    # with the ids swapped the first cells see {3, 6}: order follows the cells, not the ids
    idl2 = np.where(idl == 5, 6, 5).astype(np.uint32)
    assert list(O.hull_neighbours_box(half_space(), ORIGIN, BOX, idl2.ravel(), 3)) == [3, 6, 5]


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
    assert list(O.hull_neighbours_box(half_space(), ORIGIN, b6, idl, 3)) == [3]
