"""Write tests/golden/fft_golden.npz: heat-kernel convolution golden vectors.

The reference convolves each n^3 box with a Gaussian by an FFT library
(src/box.cpp:685-795): forward real-to-complex 3-D DFT, multiply by
G(s) = exp(-sqrt(2/pi)*s*4*dt*nn^2/n^2*pi^2)/n^3 with s = kx2^2+ky2^2+kz2^2,
k2 = min(k, n-k) (src/box.cpp:751-767), inverse complex-to-real DFT.  Here the
FFT is numpy's pocketfft (numpy 2.2.6), i.e. an independent FFT library in
place of MKL/FFTW.  G is evaluated with the reference's expression in numpy.

Inputs are a deterministic integer pattern (no RNG), so the test can rebuild
them; for n <= 16 the whole output is stored, for larger n a fixed sample of
output voxels plus whole-array moments.
Run: python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

PI = 3.14159265358979323846
SIZES = [4, 5, 6, 7, 8, 9, 10, 11, 12, 16, 27, 31, 37, 41, 64, 70]
NN, DT = 103, 0.2 / 104 ** 2     # the 512-grain configuration's conv_domain_size and dt
NN2, DT2 = 322, 0.2 / 323 ** 2   # the 8000-grain configuration's


def pattern(n):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (((7 * x + 13 * y + 29 * z + x * y * z) % 23) / 23.0 - 0.5) * 0.1


def gfactor(n, nn, dt):
    k = np.arange(n)
    k2 = np.minimum(k, n - k)
    kx = np.arange(n // 2 + 1)
    kx2 = np.minimum(kx, n - kx)
    s = k2[:, None, None] ** 2 + k2[None, :, None] ** 2 + kx2[None, None, :] ** 2
    return np.exp(-np.sqrt(2 / PI) * s * 4.0 * dt * (nn * nn) / (n * n) * PI * PI) / (n * n * n)


def conv(phi, nn, dt):
    n = phi.shape[0]
    F = np.fft.rfftn(phi, axes=(0, 1, 2))
    return np.fft.irfftn(F * gfactor(n, nn, dt), s=(n, n, n), axes=(0, 1, 2), norm="forward")


def sample_idx(n, m=256):
    # fixed pseudo-random voxel sample from a linear congruence (no RNG dependency)
    t = np.arange(m, dtype=np.int64)
    return (t * 2654435761 + 12345) % (n ** 3)


def main():
    out = {}
    for n in SIZES:
        for tag, nn, dt in (("a", NN, DT), ("b", NN2, DT2)):
            phi = pattern(n)
            res = conv(phi, nn, dt)
            key = "%s%d" % (tag, n)
            out[key + "_nn"] = np.array([nn])
            out[key + "_dt"] = np.array([dt])
            if n <= 16:
                out[key + "_out"] = res
            else:
                idx = sample_idx(n)
                out[key + "_idx"] = idx
                out[key + "_sample"] = res.ravel()[idx]
            w = np.cos(np.arange(n ** 3) * 0.001)
            out[key + "_moments"] = np.array([res.sum(), (res * res).sum(), (res.ravel() * w).sum()])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
