"""The FP64 FFT convolution on chosen box sides, bit for bit against the
oracle's FFT: large primes (one direct conjugate-pair pass of length p), two
times a prime, powers and products of 2, 3, 4, 5, up to the largest box side
of 96.  A ball inside a matrix grain gives the matrix a box of side
rds + 2 border - 1 (here the chosen n) and the ball a smaller one; one whole
timestep is compared stage by stage (after the convolution both distance
buffers, then the rest of the step).

Each case prints its box side before touching the GPU.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
import physics_cases as PC
from parity import compare_all, step_both
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIDES = [13, 23, 29, 37, 41, 47, 53, 59, 64, 67, 71, 72, 79, 81, 83, 86, 88, 89, 94, 96]


@pytest.mark.parametrize("n", SIDES)
def test_fft_convolution_box_side_bitexact(n):
    border = 3
    rds = n + 1 - 2 * border
    print("FFT box side %d (rds %d)" % (n, rds), flush=True)
    case = PC.sphere_in_matrix(rds=rds, border=border, radius=max(2.5, rds / 5.0))
    ctx = PC.gpu_context(case, G.FP64_PLAIN)
    orc = PC.oracle_sim(O, case)
    b = ctx.bounds(2, 1)
    assert b[3] - b[0] == n, (b, n)
    step_both(ctx, orc, compare_every_stage=True, tag="box side %d" % n)
    compare_all(ctx, orc, "box side %d, after the step" % n)
    ctx.close()
    orc.close()
