"""ctypes wrapper of the CPU FP64 oracle (oracle/gg_oracle.cpp).

Test infrastructure only: tests/, bench.py's cpu_baseline leg and
__graft_entry__.smoke() use it as the checker.  The gragles_3d_amd package never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libggoracle.so")

DELETED, ALIVE, VANISHED = 0, 1, 2


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ngridpoints", ctypes.c_int32),
        ("grid_blowup", ctypes.c_int32),
        ("conv_domain_size", ctypes.c_int32),
        ("boundary_grain_tube", ctypes.c_int32),
        ("h", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("n_grains", ctypes.c_int32),
    ]


def build():
    """Compile the oracle with its Makefile (g++ -O3 -fopenmp -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        i32p, u8p, f64p, u32p = P(ctypes.c_int32), P(ctypes.c_uint8), P(ctypes.c_double), P(ctypes.c_uint32)
        vp = ctypes.c_void_p
        sig = {
            "orc_last_error": (ctypes.c_char_p, [vp]),
            "orc_create": (vp, [P(OrcParams)]),
            "orc_destroy": (None, [vp]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_set_sample_stride": (None, [ctypes.c_int]),
            "orc_set_physics": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, f64p, f64p]),
            "orc_get_elements": (ctypes.c_long, [vp, ctypes.c_int, i32p, i32p, i32p, i32p, f64p, f64p, f64p,
                                                 ctypes.c_long]),
            "orc_get_element_barycenters": (ctypes.c_long, [vp, ctypes.c_int, ctypes.c_long, f64p, ctypes.c_long]),
            "orc_get_element_ids": (ctypes.c_long, [vp, ctypes.c_int, ctypes.c_long, i32p, ctypes.c_long]),
            "orc_misorientation": (ctypes.c_double, [f64p, f64p]),
            "orc_read_shockley": (ctypes.c_double, [ctypes.c_double, ctypes.c_int]),
            "orc_mobility": (ctypes.c_double, [ctypes.c_double, ctypes.c_int]),
            "orc_set_boxes": (ctypes.c_int, [vp, i32p, u8p]),
            "orc_init_distance": (ctypes.c_int, [vp, i32p]),
            "orc_set_lists": (ctypes.c_int, [vp, i32p, i32p]),
            "orc_switch": (ctypes.c_int, [vp]),
            "orc_convolution": (ctypes.c_int, [vp]),
            "orc_update_lists": (ctypes.c_int, [vp]),
            "orc_comparison": (ctypes.c_int, [vp]),
            "orc_level_set": (ctypes.c_int, [vp]),
            "orc_redistancing": (ctypes.c_int, [vp]),
            "orc_count": (ctypes.c_int, [vp]),
            "orc_timestep": (ctypes.c_int, [vp, ctypes.c_int, i32p]),
            "orc_coarsen": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
            "orc_voxel_container": (ctypes.c_long, [vp, u32p, ctypes.c_long]),
            "orc_get_hull_elements": (ctypes.c_long, [vp, ctypes.c_int, i32p, i32p, i32p, f64p, ctypes.c_long]),
            "orc_get_hull_element_areas": (ctypes.c_long, [vp, ctypes.c_int, f64p, ctypes.c_long]),
            "orc_hull_elements_box": (ctypes.c_long, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p,
                                                      i32p, u32p, ctypes.c_int, i32p, i32p, i32p, f64p, f64p,
                                                      ctypes.c_long]),
            "orc_resize_grid_box": (ctypes.c_long, [i32p, f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, i32p, i32p, f64p,
                                                    ctypes.c_long]),
            "orc_get_state": (ctypes.c_int, [vp, ctypes.c_int]),
            "orc_get_bounds": (ctypes.c_long, [vp, ctypes.c_int, ctypes.c_int, i32p]),
            "orc_get_phi": (ctypes.c_long, [vp, ctypes.c_int, ctypes.c_int, f64p, ctypes.c_long]),
            "orc_get_idlocal": (ctypes.c_long, [vp, ctypes.c_int, i32p, u32p, ctypes.c_long]),
            "orc_get_list": (ctypes.c_long, [vp, ctypes.c_int, ctypes.c_int, i32p, ctypes.c_long]),
            "orc_get_hull_neighbours": (ctypes.c_long, [vp, ctypes.c_int, i32p, ctypes.c_long]),
            "orc_hull_neighbours_box": (ctypes.c_long, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p,
                                                        i32p, u32p, ctypes.c_int, i32p, ctypes.c_long, i32p]),
            "orc_get_hull_junctions": (None, [vp, ctypes.c_int, i32p]),
            "orc_get_stats": (None, [vp, f64p, f64p, i32p]),
            "orc_resize_to_cube": (None, [i32p, ctypes.c_int]),
            "orc_convolve_box": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, f64p, f64p]),
            "orc_gauss_table": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, f64p]),
            "orc_hull_box": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p,
                                            ctypes.c_double, f64p, f64p]),
            "orc_redistance_box": (None, [i32p, f64p, i32p, f64p, ctypes.c_double, ctypes.c_double]),
            "orc_tri_row": (ctypes.c_int, [ctypes.c_int, i32p]),
            "orc_comparison_box": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_double, i32p, f64p, ctypes.c_int, i32p, i32p, f64p,
                                                  u32p, f64p, i32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def set_threads(n):
    lib().orc_set_threads(int(n))


def misorientation(q1, q2):
    """LSbox::computeMisorientation of two unit quaternions (q0, q1, q2, q3), radians, at least 1 degree."""
    a = np.ascontiguousarray(q1, dtype=np.float64)
    b = np.ascontiguousarray(q2, dtype=np.float64)
    return float(lib().orc_misorientation(_p(a, ctypes.c_double), _p(b, ctypes.c_double)))


def read_shockley(misori, unique=False):
    return float(lib().orc_read_shockley(float(misori), 1 if unique else 0))


def mobility(misori, model=True):
    return float(lib().orc_mobility(float(misori), 1 if model else 0))


def set_sample_stride(stride):
    """Timing only: stages visit just the grains whose id is a multiple of stride."""
    lib().orc_set_sample_stride(int(stride))


# ---- standalone helpers ---------------------------------------------------
def resize_to_cube(bounds, L):
    b = np.ascontiguousarray(bounds, dtype=np.int32).copy()
    lib().orc_resize_to_cube(_p(b, ctypes.c_int32), int(L))
    return b


def convolve_box(phi, nn, dt):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0]
    assert phi.shape == (n, n, n)
    out = np.empty_like(phi)
    lib().orc_convolve_box(n, int(nn), float(dt), _p(phi, ctypes.c_double), _p(out, ctypes.c_double))
    return out


def gauss_table(n, nn, dt):
    g = np.empty(3 * (n // 2) ** 2 + 1, dtype=np.float64)
    lib().orc_gauss_table(n, int(nn), float(dt), _p(g, ctypes.c_double))
    return g


def hull_box(phi, origin, h):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0]
    v = np.zeros(1)
    a = np.zeros(1)
    ex = lib().orc_hull_box(n, int(origin[0]), int(origin[1]), int(origin[2]), _p(phi, ctypes.c_double),
                            float(h), _p(v, ctypes.c_double), _p(a, ctypes.c_double))
    return bool(ex), float(v[0]), float(a[0])


def hull_neighbours_box(phi, origin, idl_bounds, idl, self_id, junctions=False):
    """Grains the hull of a cube box touches (own id for inside corners, IDLocal
    ids elsewhere), in first-appearance order; with junctions=True also the
    counts of distinct cell id lists of 2, 3, 4 and other sizes."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0]
    b6 = np.ascontiguousarray(idl_bounds, dtype=np.int32)
    idl = np.ascontiguousarray(idl, dtype=np.uint32)
    assert idl.size == (b6[3] - b6[0]) * (b6[4] - b6[1]) * (b6[5] - b6[2])
    out = np.empty(1024, dtype=np.int32)
    j4 = np.zeros(4, dtype=np.int32)
    k = lib().orc_hull_neighbours_box(n, int(origin[0]), int(origin[1]), int(origin[2]), _p(phi, ctypes.c_double),
                                      _p(b6, ctypes.c_int32), _p(idl, ctypes.c_uint32), int(self_id),
                                      _p(out, ctypes.c_int32), out.size, _p(j4, ctypes.c_int32))
    assert k >= 0
    return (out[:k].copy(), j4) if junctions else out[:k].copy()


def hull_elements_box(phi, origin, idl_bounds, idl, self_id, with_area=False):
    """Interfacial elements of a cube box's hull: [(ids, ntri, mean barycenter)]
    sorted by ids (with_area: a fourth field, the summed triangle areas in grid
    units)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0]
    b6 = np.ascontiguousarray(idl_bounds, dtype=np.int32)
    idl = np.ascontiguousarray(idl, dtype=np.uint32)
    cap = 1024
    nids, ids8, ntri = np.zeros(cap, np.int32), np.zeros(8 * cap, np.int32), np.zeros(cap, np.int32)
    pos = np.zeros(3 * cap)
    area = np.zeros(cap)
    k = lib().orc_hull_elements_box(n, int(origin[0]), int(origin[1]), int(origin[2]), _p(phi, ctypes.c_double),
                                    _p(b6, ctypes.c_int32), _p(idl, ctypes.c_uint32), int(self_id),
                                    _p(nids, ctypes.c_int32), _p(ids8, ctypes.c_int32), _p(ntri, ctypes.c_int32),
                                    _p(pos, ctypes.c_double), _p(area, ctypes.c_double), cap)
    assert k >= 0
    out = [(tuple(ids8[8 * r:8 * r + min(8, nids[r])].tolist()), int(ntri[r]), tuple(pos[3 * r:3 * r + 3]))
           for r in range(k)]
    if with_area:
        out = [e + (float(area[r]),) for r, e in enumerate(out)]
    return out


def redistance_box(in_bounds, phi_in, out_bounds, h, delta):
    ib = np.ascontiguousarray(in_bounds, dtype=np.int32)
    ob = np.ascontiguousarray(out_bounds, dtype=np.int32)
    phi_in = np.ascontiguousarray(phi_in, dtype=np.float64)
    out = np.empty((ob[5] - ob[2], ob[4] - ob[1], ob[3] - ob[0]), dtype=np.float64)
    lib().orc_redistance_box(_p(ib, ctypes.c_int32), _p(phi_in, ctypes.c_double), _p(ob, ctypes.c_int32),
                             _p(out, ctypes.c_double), float(h), float(delta))
    return out


def resize_grid_box(src_bounds, phi, gb, ngp_new, h_old, h_new, delta_new):
    """LSbox::resizeGrid on one box (src/box.cpp:2595-2697): returns (scaled
    bounds m_new*, cubed bounds, resampled values [z][y][x])."""
    sb = np.ascontiguousarray(src_bounds, dtype=np.int32)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    nb = np.zeros(6, dtype=np.int32)
    db = np.zeros(6, dtype=np.int32)
    cap = int(max(ngp_new, 1)) ** 3
    out = np.empty(cap, dtype=np.float64)
    n = lib().orc_resize_grid_box(_p(sb, ctypes.c_int32), _p(phi, ctypes.c_double), int(gb), int(ngp_new),
                                  float(h_old), float(h_new), float(delta_new), _p(nb, ctypes.c_int32),
                                  _p(db, ctypes.c_int32), _p(out, ctypes.c_double), cap)
    if n < 0:
        raise RuntimeError("resized box larger than the new grid")
    shape = (db[5] - db[2], db[4] - db[1], db[3] - db[0])
    return nb, db, out[:n].reshape(shape)


def comparison_box(ngp, gb, h, delta, own_bounds, phi_in, neighbours, idl_stale, tube=0):
    """Comparison + wall + set-comparison of one grain.

    neighbours: list of (id, bounds6, phi) in comparison-list order.
    Returns (phi_out, idlocal, new_bounds)."""
    ob = np.ascontiguousarray(own_bounds, dtype=np.int32)
    phi_in = np.ascontiguousarray(phi_in, dtype=np.float64)
    nbb = np.ascontiguousarray([b for _, b, _ in neighbours] or [[0] * 6], dtype=np.int32)
    ids = np.ascontiguousarray([i for i, _, _ in neighbours] or [0], dtype=np.int32)
    phis = np.ascontiguousarray(np.concatenate([np.ravel(p) for _, _, p in neighbours]) if neighbours
                                else np.zeros(1), dtype=np.float64)
    idl = np.ascontiguousarray(idl_stale, dtype=np.uint32).copy()
    out = np.empty_like(phi_in)
    nb = np.zeros(6, dtype=np.int32)
    rc = lib().orc_comparison_box(int(ngp), int(gb), int(tube), float(h), float(delta), _p(ob, ctypes.c_int32),
                                  _p(phi_in, ctypes.c_double), len(neighbours), _p(nbb, ctypes.c_int32),
                                  _p(ids, ctypes.c_int32), _p(phis, ctypes.c_double), _p(idl, ctypes.c_uint32),
                                  _p(out, ctypes.c_double), _p(nb, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error(None).decode())
    return out, idl, nb


def tri_row(c):
    r = np.empty(16, dtype=np.int32)
    m = lib().orc_tri_row(int(c), _p(r, ctypes.c_int32))
    return m, r


# ---- whole simulation -----------------------------------------------------
class OracleSim:
    """All grain boxes of one simulation, stage by stage, on the CPU in FP64."""

    def __init__(self, ngridpoints, grid_blowup, conv_domain_size, h, delta, dt, n_grains,
                 boundary_grain_tube=0):
        p = OrcParams(ngridpoints, grid_blowup, conv_domain_size, boundary_grain_tube, h, delta, dt, n_grains)
        self.n_grains = n_grains
        self.h = lib().orc_create(ctypes.byref(p))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())

    def set_boxes(self, bounds, exists):
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        e = np.ascontiguousarray(exists, dtype=np.uint8)
        self._chk(lib().orc_set_boxes(self.h, _p(b, ctypes.c_int32), _p(e, ctypes.c_uint8)))

    def init_distance(self, idfield):
        f = np.ascontiguousarray(idfield, dtype=np.int32)
        self._chk(lib().orc_init_distance(self.h, _p(f, ctypes.c_int32)))

    def set_lists(self, offsets, ids):
        o = np.ascontiguousarray(offsets, dtype=np.int32)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        if i.size == 0:
            i = np.zeros(1, dtype=np.int32)
        self._chk(lib().orc_set_lists(self.h, _p(o, ctypes.c_int32), _p(i, ctypes.c_int32)))

    def switch(self):
        self._chk(lib().orc_switch(self.h))

    def convolution(self):
        self._chk(lib().orc_convolution(self.h))

    def update_lists(self):
        self._chk(lib().orc_update_lists(self.h))

    def comparison(self):
        self._chk(lib().orc_comparison(self.h))

    def level_set(self):
        self._chk(lib().orc_level_set(self.h))

    def redistancing(self):
        self._chk(lib().orc_redistancing(self.h))

    def count(self):
        return lib().orc_count(self.h)

    def timestep(self, nsteps=1):
        c = np.zeros(max(nsteps, 1), dtype=np.int32)
        self._chk(lib().orc_timestep(self.h, int(nsteps), _p(c, ctypes.c_int32)))
        return c[:nsteps]

    def coarsen(self, new_size, new_dt):
        """gridCoarsement's coarsening branch; a coarsening step is coarsen()
        followed by timestep()."""
        self._chk(lib().orc_coarsen(self.h, int(new_size), float(new_dt)))

    def set_physics(self, correction=True, mobility_model=False, unique_energies=True, stored_energy=False,
                    tl_drag=0.0, quat=None, see=None):
        """Convolution-correction settings, orientations (rows 0..n_grains of (q0, q1, q2, q3); row 0
        the domain wall's pseudo grain) and stored energies (0..n_grains)."""
        q = None if quat is None else np.ascontiguousarray(quat, dtype=np.float64)
        e = None if see is None else np.ascontiguousarray(see, dtype=np.float64)
        self._chk(lib().orc_set_physics(self.h, int(bool(correction)), int(bool(mobility_model)),
                                        int(bool(unique_energies)), int(bool(stored_energy)), float(tl_drag),
                                        None if q is None else _p(q, ctypes.c_double),
                                        None if e is None else _p(e, ctypes.c_double)))

    def correction_elements(self, gid):
        """The interfacial elements of the grain's last hull in projection order: [(class 3 HJ /
        2 QJ / 1 TL / 0 GB, ids tuple, barycenter count, position, energy, mobility)]."""
        n = lib().orc_get_elements(self.h, int(gid), None, None, None, None, None, None, None, 0)
        k = max(n, 1)
        cls, nnb, nbary = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(k, np.int32)
        nb8 = np.zeros(8 * k, np.int32)
        pos, en, mob = np.zeros(3 * k), np.zeros(k), np.zeros(k)
        lib().orc_get_elements(self.h, int(gid), _p(cls, ctypes.c_int32), _p(nnb, ctypes.c_int32),
                               _p(nb8, ctypes.c_int32), _p(nbary, ctypes.c_int32), _p(pos, ctypes.c_double),
                               _p(en, ctypes.c_double), _p(mob, ctypes.c_double), k)
        return [(int(cls[r]), tuple(nb8[8 * r:8 * r + min(8, nnb[r])].tolist()), int(nbary[r]),
                 tuple(pos[3 * r:3 * r + 3]), float(en[r]), float(mob[r])) for r in range(n)]

    def element_barycenters(self, gid, e):
        """(n, 3) barycenters (x, y, z) of element e of correction_elements(gid)."""
        n = lib().orc_get_element_barycenters(self.h, int(gid), int(e), None, 0)
        if n < 0:
            raise IndexError("no element %d" % e)
        xyz = np.zeros(3 * max(n, 1))
        lib().orc_get_element_barycenters(self.h, int(gid), int(e), _p(xyz, ctypes.c_double), n)
        return xyz[:3 * n].reshape(n, 3)

    def element_ids(self, gid, e):
        """All neighbour ids of element e of correction_elements(gid), repeats included."""
        n = lib().orc_get_element_ids(self.h, int(gid), int(e), None, 0)
        if n < 0:
            raise IndexError("no element %d" % e)
        ids = np.zeros(max(n, 1), np.int32)
        lib().orc_get_element_ids(self.h, int(gid), int(e), _p(ids, ctypes.c_int32), n)
        return ids[:n].copy()

    def hull_elements(self, gid):
        """The grain's interfacial elements sorted by ids: [(ids tuple, ntri,
        position (x, y, z) = mean triangle barycenter)]."""
        cap = 4096
        nids = np.zeros(cap, np.int32)
        ids8 = np.zeros(8 * cap, np.int32)
        ntri = np.zeros(cap, np.int32)
        pos = np.zeros(3 * cap)
        n = lib().orc_get_hull_elements(self.h, int(gid), _p(nids, ctypes.c_int32), _p(ids8, ctypes.c_int32),
                                        _p(ntri, ctypes.c_int32), _p(pos, ctypes.c_double), cap)
        if n < 0:
            raise RuntimeError("too many elements")
        return [(tuple(ids8[8 * r:8 * r + min(8, nids[r])].tolist()), int(ntri[r]), tuple(pos[3 * r:3 * r + 3]))
                for r in range(n)]

    def hull_element_areas(self, gid):
        """Summed triangle areas (grid units) of hull_elements(gid), same order."""
        cap = 4096
        a = np.zeros(cap)
        n = lib().orc_get_hull_element_areas(self.h, int(gid), _p(a, ctypes.c_double), cap)
        if n < 0:
            raise RuntimeError("too many elements")
        return a[:n].copy()

    def voxel_container(self):
        """saveNetworkAsVoxelContainer's id container, [z][y][x] of side
        realDomainSize (the convolution's domain size)."""
        L = int(lib().orc_voxel_container(self.h, None, 0))
        out = np.zeros(L ** 3, dtype=np.uint32)
        n = lib().orc_voxel_container(self.h, _p(out, ctypes.c_uint32), out.size)
        if n < 0:
            raise RuntimeError("container buffer too small")
        return out.reshape(L, L, L)

    def state(self, gid):
        return lib().orc_get_state(self.h, int(gid))

    def bounds(self, gid, which):
        b = np.zeros(6, dtype=np.int32)
        n = lib().orc_get_bounds(self.h, int(gid), int(which), _p(b, ctypes.c_int32))
        return (b, n) if n >= 0 else (None, -1)

    def phi(self, gid, which):
        """Distance buffer of grain gid (which: 0 = input, 1 = output) as its linear storage."""
        b, n = self.bounds(gid, which)
        if b is None:
            return None, None
        buf = np.empty(max(n, 1), dtype=np.float64)
        m = lib().orc_get_phi(self.h, int(gid), int(which), _p(buf, ctypes.c_double), buf.size)
        return b, buf[:m]

    def idlocal(self, gid):
        b = np.zeros(6, dtype=np.int32)
        n = lib().orc_get_idlocal(self.h, int(gid), _p(b, ctypes.c_int32), None, 0)
        buf = np.empty(max(n, 1), dtype=np.uint32)
        lib().orc_get_idlocal(self.h, int(gid), _p(b, ctypes.c_int32), _p(buf, ctypes.c_uint32), buf.size)
        return b, buf[:n]

    def neighbour_list(self, gid, which=0):
        buf = np.empty(4096, dtype=np.int32)
        n = lib().orc_get_list(self.h, int(gid), int(which), _p(buf, ctypes.c_int32), buf.size)
        assert n >= 0
        return buf[:n].copy()

    def hull_neighbours(self, gid):
        """Grains the last hull of gid touched, in first-appearance order."""
        buf = np.empty(4096, dtype=np.int32)
        n = lib().orc_get_hull_neighbours(self.h, int(gid), _p(buf, ctypes.c_int32), buf.size)
        assert n >= 0
        return buf[:n].copy()

    def hull_junctions(self, gid):
        """Distinct cell id lists of the last hull of gid with 2, 3, 4 and other numbers of ids."""
        c4 = np.zeros(4, dtype=np.int32)
        lib().orc_get_hull_junctions(self.h, int(gid), _p(c4, ctypes.c_int32))
        return c4

    def stats(self):
        v = np.zeros(self.n_grains + 1)
        a = np.zeros(self.n_grains + 1)
        nb = np.zeros((self.n_grains + 1) * 6, dtype=np.int32)
        lib().orc_get_stats(self.h, _p(v, ctypes.c_double), _p(a, ctypes.c_double), _p(nb, ctypes.c_int32))
        return v, a, nb.reshape(-1, 6)

    def states(self):
        return np.array([self.state(i) for i in range(self.n_grains + 1)], dtype=np.int32)
