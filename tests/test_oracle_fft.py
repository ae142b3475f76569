"""The oracle's FFT convolution follows its documented operation order.

`convolve_box` in oracle/gg_oracle.cpp is pinned for correctness by the
pocketfft golden vectors and closed forms (test_oracle_conv.py).  The GPU's
FP64 path must reproduce it bit for bit, so its operation order is a
specification of its own (header of convolve_box: Stockham passes with radices
4, 2, 3, 5 as butterflies and any other prime as conjugate-pair sums, twiddles
only where j > 0, two real rows per complex FFT).  Here numpy restates that specification pass by pass, vectorised
over whole lines (elementwise IEEE double operations, no fused multiply-add),
and must equal the oracle bit for bit for lengths that exercise every radix and
both row-count parities.  A plausible slip in the C++ (a swapped butterfly
output, a twiddle index off by a factor, the wrong half of a row pair) breaks
the equality even where it would stay inside the golden-vector tolerance.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

C1, C2 = 0.309016994374947424102293417182819059, -0.809016994374947424102293417182819059
S1, S2 = 0.951056516295153572116439333379382143, 0.587785252292473129168705954639072769
S3 = 0.866025403784438646763723170752936183


def radices(n):
    r, m = [], n
    while m % 4 == 0:
        r.append(4)
        m //= 4
    if m % 2 == 0:
        r.append(2)
        m //= 2
    p = 3
    while m > 1:
        while m % p == 0:
            r.append(p)
            m //= p
        p += 2
    return r


def tables(n):
    # the C library's cos/sin of (2.0 * pi * t) / n, as the oracle computes them
    tc = np.array([math.cos(2.0 * math.pi * t / n) for t in range(n)])
    ts = np.array([math.sin(2.0 * math.pi * t / n) for t in range(n)])
    return tc, ts


def twid(v, c, s, inv):
    a, b = v
    return (a * c - b * s, a * s + b * c) if inv else (a * c + b * s, b * c - a * s)


def butterfly(p, v, inv, tc, ts, T):
    add = lambda a, b: (a[0] + b[0], a[1] + b[1])
    sub = lambda a, b: (a[0] - b[0], a[1] - b[1])
    if p == 2:
        return [add(v[0], v[1]), sub(v[0], v[1])]
    if p == 4:
        a0, a1, a2, a3 = add(v[0], v[2]), sub(v[0], v[2]), add(v[1], v[3]), sub(v[1], v[3])
        m1, p1 = (a1[0] + a3[1], a1[1] - a3[0]), (a1[0] - a3[1], a1[1] + a3[0])
        return [add(a0, a2), p1 if inv else m1, sub(a0, a2), m1 if inv else p1]
    if p == 3:
        t1 = add(v[1], v[2])
        t2 = (v[0][0] - 0.5 * t1[0], v[0][1] - 0.5 * t1[1])
        t3 = (S3 * (v[1][0] - v[2][0]), S3 * (v[1][1] - v[2][1]))
        m1, p1 = (t2[0] + t3[1], t2[1] - t3[0]), (t2[0] - t3[1], t2[1] + t3[0])
        return [add(v[0], t1), p1 if inv else m1, m1 if inv else p1]
    if p == 5:
        t1, t2, t3, t4 = add(v[1], v[4]), add(v[2], v[3]), sub(v[1], v[4]), sub(v[2], v[3])
        y0 = ((v[0][0] + t1[0]) + t2[0], (v[0][1] + t1[1]) + t2[1])
        b1 = ((v[0][0] + C1 * t1[0]) + C2 * t2[0], (v[0][1] + C1 * t1[1]) + C2 * t2[1])
        b2 = ((v[0][0] + C2 * t1[0]) + C1 * t2[0], (v[0][1] + C2 * t1[1]) + C1 * t2[1])
        d1 = (S1 * t3[0] + S2 * t4[0], S1 * t3[1] + S2 * t4[1])
        d2 = (S2 * t3[0] - S1 * t4[0], S2 * t3[1] - S1 * t4[1])
        m1, p1 = (b1[0] + d1[1], b1[1] - d1[0]), (b1[0] - d1[1], b1[1] + d1[0])
        m2, p2 = (b2[0] + d2[1], b2[1] - d2[0]), (b2[0] - d2[1], b2[1] + d2[0])
        return [y0, p1 if inv else m1, p2 if inv else m2, m2 if inv else p2, m1 if inv else p1]
    h = (p - 1) // 2
    a = [add(v[r], v[p - r]) if r else None for r in range(h + 1)]
    b = [sub(v[r], v[p - r]) if r else None for r in range(h + 1)]
    y0 = v[0]
    for r in range(1, h + 1):
        y0 = add(y0, a[r])
    out = [y0] + [None] * (p - 1)
    for k in range(1, h + 1):
        Pr = Pi = Qr = Qi = 0.0
        for r in range(1, h + 1):
            t = (r * k % p) * T
            Pr = Pr + a[r][0] * tc[t]
            Pi = Pi + a[r][1] * tc[t]
            Qr = Qr + b[r][0] * ts[t]
            Qi = Qi + b[r][1] * ts[t]
        m = (v[0][0] + Pr, v[0][1] + Pi)
        lo, hi = (m[0] + Qi, m[1] - Qr), (m[0] - Qi, m[1] + Qr)
        out[k], out[p - k] = (hi, lo) if inv else (lo, hi)
    return out


def fft(xr, xi, inv):
    """FFT along the last axis of (xr, xi), as the specification orders it."""
    n = xr.shape[-1]
    tc, ts = tables(n)
    Ns = 1
    for p in radices(n):
        T = n // p
        i = np.arange(T)
        j = i % Ns
        v = [(xr[..., i + r * T], xi[..., i + r * T]) for r in range(p)]
        for r in range(1, p):
            t = r * j * (n // (Ns * p))
            w = twid(v[r], tc[t], ts[t], inv)
            v[r] = (np.where(j > 0, w[0], v[r][0]), np.where(j > 0, w[1], v[r][1]))
        y = butterfly(p, v, inv, tc, ts, T)
        outr, outi = np.empty_like(xr), np.empty_like(xi)
        base = (i - j) * p + j
        for k in range(p):
            outr[..., base + k * Ns] = y[k][0]
            outi[..., base + k * Ns] = y[k][1]
        xr, xi = outr, outi
        Ns *= p
    return xr, xi


def fft_axis(Sr, Si, axis, inv):
    r, i = fft(np.moveaxis(Sr, axis, -1), np.moveaxis(Si, axis, -1), inv)
    return np.ascontiguousarray(np.moveaxis(r, -1, axis)), np.ascontiguousarray(np.moveaxis(i, -1, axis))


def convolve(phi, g):
    n = phi.shape[0]
    n2, h = n // 2 + 1, n // 2
    a = phi[:, 0::2, :]
    b = np.zeros_like(a)
    b[:, :h, :] = phi[:, 1::2, :]
    Zr, Zi = fft(a.copy(), b, False)
    k = np.arange(n2)
    m = (n - k) % n
    Sr = np.zeros((n, n, n2))
    Si = np.zeros((n, n, n2))
    Sr[:, 0::2, :] = 0.5 * (Zr[..., k] + Zr[..., m])
    Si[:, 0::2, :] = 0.5 * (Zi[..., k] - Zi[..., m])
    Sr[:, 1::2, :] = (0.5 * (Zi[..., k] + Zi[..., m]))[:, :h, :]
    Si[:, 1::2, :] = (0.5 * (Zr[..., m] - Zr[..., k]))[:, :h, :]
    Sr, Si = fft_axis(Sr, Si, 1, False)
    Sr, Si = fft_axis(Sr, Si, 0, False)
    kk = np.minimum(np.arange(n), n - np.arange(n))
    s = kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :n2] ** 2
    Sr, Si = Sr * g[s], Si * g[s]
    Sr, Si = fft_axis(Sr, Si, 0, True)
    Sr, Si = fft_axis(Sr, Si, 1, True)
    Ar, Ai = Sr[:, 0::2, :], Si[:, 0::2, :]
    Br, Bi = np.zeros_like(Ar), np.zeros_like(Ai)
    Br[:, :h, :], Bi[:, :h, :] = Sr[:, 1::2, :], Si[:, 1::2, :]
    nr = Ar.shape[1]
    Zr, Zi = np.zeros((n, nr, n)), np.zeros((n, nr, n))
    Zr[..., :n2], Zi[..., :n2] = Ar - Bi, Ai + Br
    src = n - np.arange(n2, n)
    Zr[..., n2:], Zi[..., n2:] = Ar[..., src] + Bi[..., src], Br[..., src] - Ai[..., src]
    zr, zi = fft(Zr, Zi, True)
    out = np.zeros_like(phi)
    out[:, 0::2, :] = zr
    out[:, 1::2, :] = zi[:, :h, :]
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 15, 16, 20, 21, 25, 30, 35, 37, 40, 45, 48])
def test_oracle_follows_the_specified_operation_order(n):
    nn, dt = 103, 0.2 / 104 ** 2 * 30
    rng = np.random.default_rng(n)
    phi = rng.random((n, n, n)) - 0.5
    g = O.gauss_table(n, nn, dt)
    res = O.convolve_box(phi, nn, dt)
    ref = convolve(phi, g)
    np.testing.assert_array_equal(res, ref)
