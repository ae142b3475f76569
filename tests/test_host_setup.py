"""Host-side set-up (no GPU): the C ABI exports, grid/time variables, the seeded Voronoi
generator and the initial neighbour lists."""
import os
import re

import numpy as np
import pytest

import gragles_3d_amd as G
from gragles_3d_amd import levelset as LS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            syms |= set(re.findall(r"^\s*(?:int|void|double|const char\*)\s+(gg_\w+)\s*\(", txt, re.M))
    return syms


def test_library_exports_every_declared_symbol():
    lib = G.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in sorted(syms):
        assert hasattr(lib, s), s
    # and the binding covers exactly the declared surface
    assert syms == set(LS.SIGNATURES)


def test_derive_params_matches_reference_formulas():
    # grainhdl.cpp:70-73 and 1430-1442 with TimeStepSize 0.2, then realDomainSize -= 1 (:237-239)
    for n, ppg, border, rds in [(512, 16, 7, 104), (8000, 20, 3, 323), (20000, 20, 3, 438)]:
        p = G.derive_params(n, ppg, border)
        assert p.ngridpoints == rds + 2 * border
        assert p.h == 1.0 / rds
        assert p.delta == border / rds
        assert p.dt == 0.2 / (rds * rds)
        assert p.conv_domain_size == rds - 1


def test_voronoi_labels_nearest_seed_brute_force():
    p = G.derive_params(40, 6, 2)
    idf, bounds, exists = G.generate_voronoi(p, 17)
    seeds = G.voronoi_seeds(40, 17)
    L, gb, h = p.ngridpoints, p.grid_blowup, p.h
    k, i, j = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    x, y, z = (j - gb) * h, (i - gb) * h, (k - gb) * h
    d2 = ((x[..., None] - seeds[:, 0]) ** 2 + (y[..., None] - seeds[:, 1]) ** 2 + (z[..., None] - seeds[:, 2]) ** 2)
    nearest = np.argmin(d2, axis=-1) + 1  # argmin takes the lowest index on ties
    shell = (i < gb) | (j < gb) | (k < gb) | (i >= L - 1 - gb) | (j >= L - 1 - gb) | (k >= L - 1 - gb)
    ref = np.where(shell, 0, nearest)
    assert np.array_equal(idf, ref)
    for g in range(1, 41):
        vox = np.argwhere(ref == g)
        if vox.size == 0:
            assert not exists[g]
            continue
        assert exists[g]
        lo = vox.min(axis=0)[::-1]  # (x, y, z)
        hi = vox.max(axis=0)[::-1]
        exp = np.concatenate([np.maximum(lo - gb, 0), np.minimum(hi + 1 + gb, L)])
        assert np.array_equal(bounds[g], exp)


def test_voronoi_is_deterministic_and_thread_independent():
    p = G.derive_params(64, 8, 3)
    a = G.generate_voronoi(p, 5, True, 1)
    b = G.generate_voronoi(p, 5, True, 7)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_initial_neighbour_lists_brute_force():
    p = G.derive_params(125, 8, 3)
    _, bounds, exists = G.generate_voronoi(p, 3)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    cb = np.array([G.resize_to_cube(b, p.ngridpoints) if exists[g] else b for g, b in enumerate(bounds)])
    for g in range(1, 126):
        lst = ids[offs[g]:offs[g + 1]]
        exp = []
        for q in range(1, 126):
            if q == g or not exists[q] or not exists[g]:
                continue
            a, b = cb[g], cb[q]
            if not (a[0] > b[3] or a[3] < b[0] or a[1] > b[4] or a[4] < b[1] or a[2] > b[5] or a[5] < b[2]):
                exp.append(q)
        assert list(lst) == exp


def test_context_without_gpu_reports_device_error():
    # on a machine without a GPU gg_create must fail cleanly, never fall back
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    p = G.derive_params(8, 6, 2)
    with pytest.raises(G.GGError):
        G.LevelSetContext(p)
