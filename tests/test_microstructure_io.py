"""Voxelized microstructure input (read_HeaderCPG + read_voxelized_microstructure,
src/grainhdl.cpp:159-223, 404-578; box rule src/box.cpp:227-300) on files
written here from a generated microstructure: the id field comes back with the
offset removed inside a zero border, the boxes follow the reference's rule
(record bounds, +2*border on the upper side, clamped to the grid), records
with count 0 have no box, and the grid parameters follow NX.
"""
import numpy as np
import pytest

import gragles_3d_amd as G
from gragles_3d_amd.microstructure_io import MicrostructureError, read_voxelized_microstructure


def write_imm(tmp_path, interior, offset, border, zero_count=(), bad_id=None):
    """Header + raw files of an interior label field [z][y][x] (ids 1..n)."""
    n_cells = interior.shape[0]
    ids = np.unique(interior)
    ids = ids[ids > 0]
    n = int(ids.max())
    hdr = ["IMM microstructure", "IDOffset \t %d" % offset, "DX \t 1", "DY \t 1", "DZ \t 1",
           "NX \t %d" % n_cells, "NY \t %d" % n_cells, "NZ \t %d" % n_cells, "NumberOfGrains \t %d" % n, "", "",
           "ID x y z phi1 Phi phi2 xmin xmax ymin ymax zmin zmax count stored"]
    for g in range(1, n + 1):
        zz, yy, xx = np.nonzero(interior == g)
        cnt = 0 if g in zero_count else len(xx)
        gid = g + offset - 1 if g != bad_id else g + offset + 5
        hdr.append("%d\t %d\t %d\t %d\t %.3f\t %.3f\t %.3f\t %d\t %d\t %d\t %d\t %d\t %d\t %d \t %g" % (
            gid, xx.mean(), yy.mean(), zz.mean(), 0.1 * g, 0.2, 0.3, xx.min(), xx.max(), yy.min(), yy.max(),
            zz.min(), zz.max(), cnt, 1e-13 * g))
    hp, rp = tmp_path / "grains.txt", tmp_path / "voxels.raw"
    hp.write_text("\n".join(hdr) + "\n")
    (interior.astype(np.int64) + offset - 1).astype("<u4").tofile(str(rp))
    return str(hp), str(rp), n


def _generated(n=27, ppg=12, border=3, seed=11):
    p = G.derive_params(n, ppg, border)
    idf, _, _ = G.generate_voronoi(p, seed)
    ngp, gb, L = p.ngridpoints, p.grid_blowup, p.conv_domain_size
    cube = idf.reshape(ngp, ngp, ngp)
    out = cube[gb:gb + L, gb:gb + L, gb:gb + L].copy()  # the labelled cells
    assert (out > 0).all()
    return out


def test_id_field_boxes_and_parameters(tmp_path):
    interior = _generated()
    border = 4
    hp, rp, n = write_imm(tmp_path, interior, offset=5, border=border, zero_count=(7,))
    p, idf, bounds, exists, extra = read_voxelized_microstructure(hp, rp, border, time_step_size=0.2)
    L = interior.shape[0]
    ngp = L + 2 * border
    assert (p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.n_grains) == (ngp, border, L, n)
    assert p.h == 1.0 / L and p.delta == border / float(L) and p.dt == 0.2 / float(L * L)
    cube = idf.reshape(ngp, ngp, ngp)
    np.testing.assert_array_equal(cube[border:border + L, border:border + L, border:border + L], interior)
    inner = np.zeros(cube.shape, bool)
    inner[border:border + L, border:border + L, border:border + L] = True
    assert (cube[~inner] == 0).all()
    assert exists[7] == 0 and exists[0] == 0 and exists[1:].sum() == n - 1
    for g in (1, 2, n):
        zz, yy, xx = np.nonzero(interior == g)
        lo = [xx.min(), yy.min(), zz.min()]
        hi = [min(v + 2 * border, ngp) for v in (xx.max(), yy.max(), zz.max())]
        assert list(bounds[g]) == lo + hi
    assert np.isclose(extra["bunge"][2, 0], 0.2) and np.isclose(extra["stored"][3], 3e-13)


def test_record_bounds_beyond_the_grid_are_clamped(tmp_path):
    interior = _generated()
    L = interior.shape[0]
    hp, rp, n = write_imm(tmp_path, interior, offset=1, border=3)
    lines = open(hp).read().split("\n")
    rec = lines[12].split("\t")
    rec[7:13] = [" -2", " %d" % (L + 5), " -1", " %d" % (L + 5), " 0", " %d" % (L + 5)]
    lines[12] = "\t".join(rec)
    open(hp, "w").write("\n".join(lines))
    p, _, bounds, exists, _ = read_voxelized_microstructure(hp, rp, 3)
    assert exists[1] == 1
    assert list(bounds[1]) == [0, 0, 0, p.ngridpoints, p.ngridpoints, p.ngridpoints]


def test_ids_must_run_consecutively(tmp_path):
    interior = _generated()
    hp, rp, _ = write_imm(tmp_path, interior, offset=2, border=3, bad_id=4)
    with pytest.raises(MicrostructureError):
        read_voxelized_microstructure(hp, rp, 3)
