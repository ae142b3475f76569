"""A run that starts from voxelized microstructure files (the IMM input the
reference reads, src/grainhdl.cpp:404-578): GPU FP64 against the oracle from
the same files, bit-exact over timesteps, with a count-0 record (no box).

Each case prints its name, seed and shape before touching the GPU.
"""
import sys

import pytest

import gragles_3d_amd as G
from gragles_3d_amd.microstructure_io import read_voxelized_microstructure
from oracle import oracle as O
from parity import compare_all, step_both
from test_microstructure_io import _generated, write_imm

pytestmark = pytest.mark.gpu


def test_run_from_voxelized_files_matches_oracle(tmp_path):
    print("voxelized input n=27 ppg=12 seed=11 border=4 offset=3", flush=True)
    sys.stdout.flush()
    interior = _generated()
    hp, rp, n = write_imm(tmp_path, interior, offset=3, border=4, zero_count=(5,))
    p, idf, bounds, exists, _ = read_voxelized_microstructure(hp, rp, 4)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    ctx = G.LevelSetContext(p, precision=G.FP64_PLAIN)
    ctx.set_boxes(bounds, exists)
    ctx.init_distance(idf)
    ctx.set_neighbour_lists(offs, ids)
    orc = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, n)
    orc.set_boxes(bounds, exists)
    orc.init_distance(idf)
    orc.set_lists(offs, ids)
    compare_all(ctx, orc, "after init", check_phi=(1,), check_idlocal=True, check_lists=True)
    for s in range(3):
        step_both(ctx, orc)
    assert compare_all(ctx, orc, "after 3 steps") > 20
    ctx.close()
