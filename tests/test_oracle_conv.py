"""The oracle's heat-kernel convolution, pinned by numpy pocketfft golden vectors
and by closed forms (tests run on the CPU)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")
SIZES = [4, 5, 6, 7, 8, 9, 10, 11, 12, 16, 27, 31, 37, 41, 64, 70]


def pattern(n):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (((7 * x + 13 * y + 29 * z + x * y * z) % 23) / 23.0 - 0.5) * 0.1


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("tag", ["a", "b"])
@pytest.mark.parametrize("n", SIZES)
def test_matches_pocketfft_golden(gold, n, tag):
    key = "%s%d" % (tag, n)
    nn = int(gold[key + "_nn"][0])
    dt = float(gold[key + "_dt"][0])
    res = O.convolve_box(pattern(n), nn, dt)
    scale = np.abs(pattern(n)).max()
    if n <= 16:
        ref = gold[key + "_out"]
        assert np.abs(res - ref).max() <= 1e-14 * scale * np.log2(n) + 1e-17
    else:
        idx = gold[key + "_idx"]
        ref = gold[key + "_sample"]
        assert np.abs(res.ravel()[idx] - ref).max() <= 4e-15 * scale * np.log2(n)
    w = np.cos(np.arange(n ** 3) * 0.001)
    mom = np.array([res.sum(), (res * res).sum(), (res.ravel() * w).sum()])
    ref_m = gold[key + "_moments"]
    assert np.allclose(mom, ref_m, rtol=1e-11, atol=1e-13 * n ** 3 * scale)


@pytest.mark.parametrize("n", [6, 9, 13, 24])
def test_single_mode_is_eigenfunction(n):
    # cos(2 pi (a x + b y + c z)/n) is scaled by exactly n^3 G(s), s = a2^2 + b2^2 + c2^2
    nn, dt = 103, 0.2 / 104 ** 2
    g = O.gauss_table(n, nn, dt)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for a, b, c in [(1, 0, 0), (0, 2, 1), (3, 1, 2), (n - 1, 1, 0)]:
        phi = np.cos(2 * np.pi * (a * x + b * y + c * z) / n)
        s = min(a, n - a) ** 2 + min(b, n - b) ** 2 + min(c, n - c) ** 2
        out = O.convolve_box(phi, nn, dt)
        np.testing.assert_allclose(out, phi * n ** 3 * g[s], atol=1e-13)


def test_gauss_table_closed_form():
    n, nn, dt = 20, 322, 0.2 / 323 ** 2
    g = O.gauss_table(n, nn, dt)
    s = np.arange(g.size)
    ref = np.exp(-np.sqrt(2 / np.pi) * s * 4.0 * dt * nn * nn / (n * n) * np.pi ** 2) / n ** 3
    np.testing.assert_allclose(g, ref, rtol=1e-13)
    assert g[0] == 1.0 / n ** 3


def test_constant_preserved_and_zero_dt_identity():
    n = 11
    c = np.full((n, n, n), -0.0123)
    np.testing.assert_allclose(O.convolve_box(c, 103, 1e-5), c, rtol=1e-13)
    phi = pattern(n)
    np.testing.assert_allclose(O.convolve_box(phi, 103, 0.0), phi, atol=1e-15)


def test_linearity():
    n = 10
    a = pattern(n)
    b = np.sin(np.arange(n ** 3).reshape(n, n, n) * 0.37)
    dt = 3e-4
    lhs = O.convolve_box(2.0 * a - 3.0 * b, 29, dt)
    rhs = 2.0 * O.convolve_box(a, 29, dt) - 3.0 * O.convolve_box(b, 29, dt)
    np.testing.assert_allclose(lhs, rhs, atol=1e-14)
