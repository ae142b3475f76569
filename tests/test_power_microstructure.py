"""Uneven (bimodal) microstructure input: seeds by random sequential addition
and the power-diagram labelling (host code, CPU tests).

Pins: with every radius 0 the power diagram is the Voronoi generator's field
bit for bit (same seeds, same border, same ties); on small grids with random
radii every voxel carries the brute-force argmin of |x - s|^2 - r^2 (lowest
id on ties), so the block pruning drops no winner; the bimodal seeds have the
requested mode fraction and size ratio, their balls are disjoint, and so every
power cell is non-empty."""
import numpy as np
import pytest

import gragles_3d_amd as G


def brute_power(p, xyz, radius):
    L, gb, h = p.ngridpoints, p.grid_blowup, p.h
    idf = np.zeros((L, L, L), dtype=np.int32)
    lo, hi = gb, L - 1 - gb
    k, i, j = np.mgrid[lo:hi, lo:hi, lo:hi]
    x, y, z = (j - gb) * h, (i - gb) * h, (k - gb) * h
    best = np.full(x.shape, np.inf)
    bid = np.zeros(x.shape, dtype=np.int32)
    for s in range(p.n_grains):  # ascending id, strict '<': ties keep the lowest id
        dx, dy, dz = x - xyz[s, 0], y - xyz[s, 1], z - xyz[s, 2]
        d = dx * dx + dy * dy + dz * dz - radius[s] * radius[s]
        m = d < best
        best[m] = d[m]
        bid[m] = s + 1
    idf[lo:hi, lo:hi, lo:hi] = bid
    return idf


def test_zero_radii_is_the_voronoi_generator():
    p = G.derive_params(300, 10, 3)
    xyz = G.voronoi_seeds(300, 13).reshape(-1, 3)
    a = G.generate_voronoi(p, 13)
    b = G.generate_power(p, xyz, np.zeros(300))
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("n,seed", [(40, 1), (90, 2), (25, 3)])
def test_power_labels_equal_brute_force(n, seed):
    p = G.derive_params(n, 9, 2)
    rng = np.random.default_rng(seed)
    xyz = rng.random((n, 3))
    radius = rng.random(n) * 0.12 * rng.integers(0, 2, n)  # half the grains weightless
    idf, bounds, exists = G.generate_power(p, xyz, radius)
    ref = brute_power(p, xyz, radius)
    np.testing.assert_array_equal(idf, ref)
    for g in range(1, n + 1):
        zz, yy, xx = np.nonzero(ref == g)
        assert exists[g] == (zz.size > 0)
        if zz.size:
            L, gb = p.ngridpoints, p.grid_blowup
            want = [max(0, xx.min() - gb), max(0, yy.min() - gb), max(0, zz.min() - gb),
                    min(L, xx.max() + 1 + gb), min(L, yy.max() + 1 + gb), min(L, zz.max() + 1 + gb)]
            assert list(bounds[g]) == want


def test_bimodal_seeds_modes_disjoint_balls_and_no_empty_cell():
    n = 2000
    xyz, rho = G.bimodal_seeds(n, 4, frac_large=0.1, size_ratio=4.0, sigma=0.2)
    assert xyz.min() >= 0 and xyz.max() < 1
    big = rho > np.median(rho) * 2.5
    frac = big.mean()
    assert abs(frac - 0.1) < 4 * np.sqrt(0.1 * 0.9 / n)
    # the placement factor is common to all but a few late small grains, so the
    # modes' median radii keep the size ratio
    ratio = np.median(rho[big]) / np.median(rho[~big])
    assert 3.4 < ratio < 4.6, ratio
    # lognormal spread within the large mode (placed first, no shrinking)
    s = np.std(np.log(rho[big]))
    assert 0.15 < s < 0.25, s
    d = np.sqrt(((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    assert (d - rho[:, None] - rho[None, :]).min() >= -1e-12
    p = G.derive_params(n, 14, 3)
    idf, bounds, exists = G.generate_power(p, xyz, rho, with_idfield=False)
    assert exists[1:].all()
    sides = bounds[1:, 3] - bounds[1:, 0]
    assert np.median(sides[big]) > 1.5 * np.median(sides[~big])
