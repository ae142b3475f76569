"""Invariants of the oracle's neighbour-list rule and whole-step bookkeeping (CPU).

computeSecondOrderNeighbours (src/box.cpp:1090-1134): the new comparison list is
the previous list's still-existing entries, in order, followed by the first
occurrence of every neighbour-of-neighbour that exists and whose input box
intersects the grain's (inclusive bounds), without the grain itself.
"""
import numpy as np

import gragles_3d_amd as G
from oracle import oracle as O


def make(n, ppg, border, seed):
    p = G.derive_params(n, ppg, border)
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    s = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, n)
    s.set_boxes(bounds, exists)
    s.init_distance(idf)
    s.set_lists(offs, ids)
    return s, p


def inclusive_overlap(a, b):
    return not (a[0] > b[3] or a[3] < b[0] or a[1] > b[4] or a[4] < b[1] or a[2] > b[5] or a[5] < b[2])


def test_comparison_list_rule_over_several_steps():
    s, p = make(64, 7, 3, 21)
    N = 64
    for step in range(12):
        s.switch()
        s.convolution()
        s.switch()
        old = {g: s.neighbour_list(g, 0) for g in range(1, N + 1) if s.state(g) != O.DELETED}
        st = s.states()
        s.update_lists()
        for g in range(1, N + 1):
            if st[g] != O.ALIVE:
                continue
            new = list(s.neighbour_list(g, 1))
            kept = [q for q in old[g] if st[q] == O.ALIVE]
            assert new[:len(kept)] == kept
            bg, _ = s.bounds(g, 0)
            cands = []
            for q in old[g]:
                for r in old[q]:
                    if st[r] == O.ALIVE and inclusive_overlap(bg, s.bounds(int(r), 0)[0]):
                        cands.append(int(r))
            exp = list(kept)
            for r in cands:
                if r != g and r not in exp:
                    exp.append(r)
            assert new == exp, (step, g)
            assert len(set(new)) == len(new) and g not in new
        s.comparison()
        s.switch()
        s.level_set()
        s.redistancing()


def test_grain_count_lags_vanishing_by_one_step():
    s, p = make(64, 6, 3, 3)
    prev_states = s.states()
    for step in range(40):
        s.timestep(1)
        st = s.states()
        # a grain is deleted exactly one level-set stage after it vanished
        assert np.all((st == O.DELETED) <= ((prev_states == O.VANISHED) | (prev_states == O.DELETED)) | (
            np.arange(65) == 0))
        assert np.all(prev_states[st == O.VANISHED] == O.ALIVE)
        assert s.count() == int((st[1:] != O.DELETED).sum())
        prev_states = st
    assert (prev_states[1:] == O.DELETED).sum() > 0
