"""Voxelized microstructure input (the IMM generator's files), host side.

Restates grainhdl::read_HeaderCPG (src/grainhdl.cpp:159-223) and
grainhdl::read_voxelized_microstructure (src/grainhdl.cpp:404-578) with the box
rule of LSbox(id, vertices, ori, owner, storedEnergy) (src/box.cpp:227-300),
and returns what the C ABI's set-up takes: grid parameters (gg_params), the
ngridpoints^3 id field for gg_init_distance and the per-grain boxes for
gg_set_boxes.

Header file: a skipped first line; eight "name value" records (id offset, DX,
DY, DZ, NX, NY, NZ, number of grains); blank lines; one more skipped line
(column titles); then one record per grain: id x y z bunge1 bunge2 bunge3 xmin
xmax ymin ymax zmin zmax count storedEnergy.  Binary file: NX*NY*NZ uint32 ids
(x fastest), each id - (offset - 1) in the grid, inside a border of
grid_blowup zero voxels.
"""
import numpy as np

from .levelset import GGParams


class MicrostructureError(ValueError):
    pass


def _header(path):
    with open(path, "r") as f:
        lines = f.read().split("\n")
    pos = 1  # line 1 is skipped
    vals = []
    while len(vals) < 8:
        if pos >= len(lines):
            raise MicrostructureError("%s: header ends before its eight records" % path)
        tok = lines[pos].split()
        pos += 1
        if not tok:
            continue
        vals.append(int(tok[-1]))
    while pos < len(lines) and not lines[pos].strip():
        pos += 1
    pos += 1  # the column-title line
    offset, dx, dy, dz, nx, ny, nz, ngrains = vals
    tokens = " ".join(lines[pos:]).split()
    if len(tokens) < 15 * ngrains:
        raise MicrostructureError("%s: %d grain records announced, %d found" % (path, ngrains, len(tokens) // 15))
    recs = np.array(tokens[:15 * ngrains], dtype=object).reshape(ngrains, 15)
    return offset, (dx, dy, dz), (nx, ny, nz), ngrains, recs


def read_voxelized_microstructure(header_path, raw_path, border, time_step_size=0.2):
    """Returns (params, idfield, bounds, exists, extras).

    params: GGParams of the run (realDomainSize = NX; ngridpoints = NX + 2*border,
    h = 1/NX, delta = border/NX, dt = time_step_size/NX^2, the convolution's
    domain size NX).  idfield: ngridpoints^3 int32 (x fastest).  bounds:
    (n+1) x 6 int32 {xmin,ymin,zmin,xmax,ymax,zmax} before resizeToCube;
    exists: (n+1) uint8.  extras: 'bunge' (n+1) x 3 and 'stored' (n+1), per id.
    Grain records with count 0 have no box (the reference skips them).  The
    record of grain n must carry id n + offset - 1: the reference builds grain
    n's box from record n under the id in that record, so ids that are not
    consecutive would misplace boxes.
    """
    offset, _, (nx, ny, nz), ngrains, recs = _header(header_path)
    if not (nx == ny == nz):
        raise MicrostructureError("the grid must be cubic (NX = NY = NZ), got %d %d %d" % (nx, ny, nz))
    rds = nx
    ngp = rds + 2 * border
    p = GGParams()
    p.ngridpoints = ngp
    p.grid_blowup = border
    p.conv_domain_size = rds
    p.boundary_grain_tube = 0
    p.h = 1.0 / rds
    p.delta = border / float(rds)
    p.dt = time_step_size / float(rds * rds)
    p.n_grains = ngrains

    bounds = np.zeros((ngrains + 1, 6), dtype=np.int32)
    exists = np.zeros(ngrains + 1, dtype=np.uint8)
    bunge = np.zeros((ngrains + 1, 3))
    stored = np.zeros(ngrains + 1)
    for n in range(1, ngrains + 1):
        r = recs[n - 1]
        gid = int(r[0])
        count = int(r[13])
        if count == 0:
            continue
        if gid - (offset - 1) != n:
            raise MicrostructureError("record %d carries id %d; ids must run consecutively from the offset %d" %
                                      (n, gid, offset))
        xmin, xmax, ymin, ymax, zmin, zmax = (int(v) for v in r[7:13])
        b = [xmin, ymin, zmin, xmax + 2 * border, ymax + 2 * border, zmax + 2 * border]
        for a in range(3):
            b[a + 3] = min(b[a + 3], ngp)
            b[a] = max(b[a], 0)
        if b[0] == ngp:
            continue  # "no bounding box could be found"
        bounds[n] = b
        exists[n] = 1
        bunge[n] = [float(v) for v in r[4:7]]
        stored[n] = float(r[14])

    raw = np.fromfile(raw_path, dtype="<u4")
    if raw.size < rds ** 3:
        raise MicrostructureError("%s holds %d ids, %d needed" % (raw_path, raw.size, rds ** 3))
    idf = np.zeros((ngp, ngp, ngp), dtype=np.int32)
    idf[border:border + rds, border:border + rds, border:border + rds] = (
        raw[:rds ** 3].astype(np.int64) - (offset - 1)).reshape(rds, rds, rds).astype(np.int32)
    return p, idf.reshape(-1), bounds, exists, {"bunge": bunge, "stored": stored}
