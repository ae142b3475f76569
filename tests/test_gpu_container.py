"""Voxel container (gg_voxel_container, saveNetworkAsVoxelContainer) on the
GPU: FP64 identical to the oracle after timesteps and after a grid coarsening;
FP32 against FP64; the three files the reference writes.

Each case prints its name, seed and shape before touching the GPU.
"""
import os
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from parity import make_pair, step_both

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_container_matches_oracle_after_steps_and_coarsening():
    _say("container n=64 ppg=10 border=3 seed=5")
    ctx, orc, p = make_pair(64, 10, 3, 5)
    np.testing.assert_array_equal(ctx.voxel_container(), orc.voxel_container())
    for _ in range(3):
        step_both(ctx, orc)
    np.testing.assert_array_equal(ctx.voxel_container(), orc.voxel_container())
    new = int((p.ngridpoints - 2 * p.grid_blowup) * 0.75)
    ctx.coarsen_grid(new, 0.2 / (new * new))
    orc.coarsen(new, 0.2 / (new * new))
    step_both(ctx, orc)
    c = ctx.voxel_container()
    assert c.shape == (new, new, new)
    np.testing.assert_array_equal(c, orc.voxel_container())
    ctx.close()


def test_container_fp32_follows_fp64():
    _say("container fp32 vs fp64 n=125 ppg=8 border=3 seed=7")
    c64, _, _ = make_pair(125, 8, 3, 7)
    c32, _, _ = make_pair(125, 8, 3, 7, precision=G.FP32_BATCHED)
    c64.timestep(4)
    c32.timestep(4)
    a, b = c64.voxel_container(), c32.voxel_container()
    assert (a == b).mean() > 0.999
    c64.close()
    c32.close()


def test_container_files(tmp_path):
    _say("container files n=27 ppg=12 border=3 seed=11")
    ctx, _, p = make_pair(27, 12, 3, 11)
    ctx.timestep(2)
    p_uds, p_raw, p_en = ctx.save_network_as_voxel_container(2, str(tmp_path))
    L = p.conv_domain_size
    assert os.path.basename(p_raw) == "Container_size_%d_t_2.raw" % L
    raw = np.fromfile(p_raw, dtype="<u4")
    np.testing.assert_array_equal(raw.reshape(L, L, L), ctx.voxel_container())
    st, vol, _ = ctx.stats()
    rows = [ln.split("\t") for ln in open(p_uds).read().splitlines()]
    assert [int(r[0]) for r in rows] == [g for g in range(1, 28) if st[g] == G.ALIVE]
    b = ctx.all_bounds(1)
    for r in rows:
        g = int(r[0])
        assert [int(x) for x in r[7:13]] == [b[g, 0], b[g, 3], b[g, 1], b[g, 4], b[g, 2], b[g, 5]]
        assert abs(float(r[13]) - vol[g]) <= 1e-5 * vol[g]
    assert os.path.getsize(p_en) == 0
    ctx.close()
