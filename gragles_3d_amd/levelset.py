"""ctypes binding of libgg_levelset.so (include/gg_levelset.h, include/gg_grainhdl.h).

LevelSetContext mirrors the reference's per-grain operator surface as driven by
grainhdl (src/grainhdl.cpp:836-919): convolution, switch, neighbour lists,
comparison, level set, redistancing, grain count.  Every call goes to the HIP
library; there is no other implementation behind it.  A missing or unloadable
library raises, it never falls back.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgg_levelset.so")

GG_OK, GG_ERR_ARG, GG_ERR_HIP, GG_ERR_CAPACITY, GG_ERR_STATE, GG_ERR_DEVICE = range(6)
FP64_PLAIN, FP32_BATCHED = 0, 1
ORDER_SPATIAL, ORDER_VOLUME = 0, 1
CMP_ORDER_GRAIN, CMP_ORDER_SPACE = 0, 1
CONV_DEFAULT, CONV_SPLIT_BF16 = 0, 1
DELETED, ALIVE, VANISHED = 0, 1, 2
# comparison tile shapes of gg_autotune: (layers per wavefront, gathers in flight, run-major order)
CMP_SHAPES = ((8, 8, 0), (16, 8, 0), (16, 16, 0), (32, 8, 0), (8, 8, 1), (16, 8, 1), (16, 16, 1), (32, 8, 1))
CMP_VARIANTS = len(CMP_SHAPES)
SEGMENTS = ("conv_xy", "conv_z", "conv_xy_inv", "idl_resize", "lists", "comparison_n32", "hull", "finalize", "plan",
            "redist_xy", "redist_z", "post_redist", "count", "comparison_n48", "comparison_n64", "comparison_n96",
            "resample", "container", "correction", "elements")


class GGParams(ctypes.Structure):
    _fields_ = [
        ("ngridpoints", ctypes.c_int32),
        ("grid_blowup", ctypes.c_int32),
        ("conv_domain_size", ctypes.c_int32),
        ("boundary_grain_tube", ctypes.c_int32),
        ("h", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("n_grains", ctypes.c_int32),
        ("max_list", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("arena_voxels", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GGError(RuntimeError):
    pass


_lib = None

_P = ctypes.POINTER
_i32p, _u8p, _f64p, _u32p, _i64p = (_P(ctypes.c_int32), _P(ctypes.c_uint8), _P(ctypes.c_double),
                                    _P(ctypes.c_uint32), _P(ctypes.c_int64))
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32

# name -> (restype, argtypes): every exported symbol of include/*.h
SIGNATURES = {
    "gg_create": (ctypes.c_int, [_P(GGParams), _P(_vp)]),
    "gg_destroy": (ctypes.c_int, [_vp]),
    "gg_last_error": (ctypes.c_char_p, [_vp]),
    "gg_set_boxes": (ctypes.c_int, [_vp, _i32p, _u8p]),
    "gg_init_distance": (ctypes.c_int, [_vp, _i32p]),
    "gg_init_distance_noalloc": (ctypes.c_int, [_vp, _i32p]),
    "gg_set_neighbour_lists": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "gg_switch_buffers": (ctypes.c_int, [_vp]),
    "gg_convolution": (ctypes.c_int, [_vp]),
    "gg_update_neighbour_lists": (ctypes.c_int, [_vp]),
    "gg_comparison": (ctypes.c_int, [_vp]),
    "gg_level_set": (ctypes.c_int, [_vp]),
    "gg_coarsen_grid": (ctypes.c_int, [_vp, _i32, ctypes.c_double]),
    "gg_voxel_container": (ctypes.c_int, [_vp, _u32p, ctypes.c_int64, _i32p]),
    "gg_download_hull_elements": (ctypes.c_int, [_vp, _i32, _i32p, _i32p, _i32p, _f64p, _i32, _i32p]),
    "gg_set_physics": (ctypes.c_int, [_vp, ctypes.c_void_p, _f64p, _f64p]),
    "gg_download_elements": (ctypes.c_int, [_vp, _i32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _i32,
                                            _i32p]),
    "gg_redistancing": (ctypes.c_int, [_vp]),
    "gg_count_grains": (ctypes.c_int, [_vp, _i32p]),
    "gg_timestep": (ctypes.c_int, [_vp, _i32, _i32]),
    "gg_synchronize": (ctypes.c_int, [_vp]),
    "gg_autotune": (ctypes.c_int, [_vp, _i32, _i32p, _P(ctypes.c_float)]),
    "gg_get_tuning": (ctypes.c_int, [_vp, _i32p]),
    "gg_set_tuning": (ctypes.c_int, [_vp, _i32p]),
    "gg_set_work_order": (ctypes.c_int, [_vp, _i32]),
    "gg_set_convolution_form": (ctypes.c_int, [_vp, _i32]),
    "gg_get_work_order": (ctypes.c_int, [_vp, _i32p]),
    "gg_set_comparison_order": (ctypes.c_int, [_vp, _i32, _i32]),
    "gg_get_comparison_order": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "gg_get_count_history": (ctypes.c_int, [_vp, _i32p, _i32, _i32p]),
    "gg_get_grain_stats": (ctypes.c_int, [_vp, _i32p, _f64p, _f64p]),
    "gg_get_bounds": (ctypes.c_int, [_vp, _i32, _i32, _i32p]),
    "gg_download_distance": (ctypes.c_int, [_vp, _i32, _i32, _f64p, ctypes.c_int64, _i64p]),
    "gg_download_idlocal": (ctypes.c_int, [_vp, _i32, _i32p, _u32p, ctypes.c_int64, _i64p]),
    "gg_upload_phi": (ctypes.c_int, [_vp, _i32, _i32, _f64p, ctypes.c_int64]),
    "gg_download_list": (ctypes.c_int, [_vp, _i32, _i32, _i32p, _i32, _i32p]),
    "gg_get_all_bounds": (ctypes.c_int, [_vp, _i32, _i32p]),
    "gg_enable_hull_topology": (ctypes.c_int, [_vp, _i32]),
    "gg_build_neighbour_lists": (ctypes.c_int, [_vp]),
    "gg_get_hull_neighbour_counts": (ctypes.c_int, [_vp, _i32p, _u8p]),
    "gg_download_hull_neighbours": (ctypes.c_int, [_vp, _i32, _i32p, _i32, _i32p]),
    "gg_get_hull_junction_counts": (ctypes.c_int, [_vp, _i32p]),
    "gg_get_faces": (ctypes.c_int, [_vp, _i32p, _i32p, _f64p, ctypes.c_int64, _i64p]),
    "gg_get_element_counts": (ctypes.c_int, [_vp, _i32p]),
    "gg_download_all_lists": (ctypes.c_int, [_vp, _i32, _i32p, _i32p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "gg_get_new_bounds": (ctypes.c_int, [_vp, _i32p]),
    "gg_box_voxels": (ctypes.c_int, [_vp, _i64p]),
    "gg_enable_stage_timing": (ctypes.c_int, [_vp, _i32]),
    "gg_get_stage_times": (ctypes.c_int, [_vp, _f64p, _i64p, _i32, _i32]),
    "gg_derive_params": (ctypes.c_int, [_i32, _i32, _i32, ctypes.c_double, _P(GGParams)]),
    "gg_resize_to_cube": (None, [_i32p, _i32]),
    "gg_generate_voronoi": (ctypes.c_int, [_P(GGParams), ctypes.c_uint64, _i32p, _i32p, _u8p, _i32]),
    "gg_voronoi_seeds": (ctypes.c_int, [_i32, ctypes.c_uint64, _f64p]),
    "gg_bimodal_seeds": (ctypes.c_int, [_i32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        _f64p, _f64p]),
    "gg_generate_power": (ctypes.c_int, [_P(GGParams), _f64p, _f64p, _i32p, _i32p, _u8p, _i32]),
    "gg_misorientation": (ctypes.c_double, [_f64p, _f64p]),
    "gg_gb_pair_tables": (ctypes.c_int, [_i32, _f64p, _i32, _i32, _f64p, _f64p, _i32]),
    "gg_euler_to_quaternion": (None, [_f64p, _f64p]),
    "gg_random_orientations": (ctypes.c_int, [_i32, ctypes.c_uint64, _i32, _f64p]),
    "gg_lognormal_energies": (ctypes.c_int, [_i32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _f64p]),
    "gg_initial_neighbour_lists": (ctypes.c_int, [_P(GGParams), _i32p, _u8p, _i32p, _i32p, ctypes.c_int64,
                                                  _i64p]),
}


def lib():
    """Load the in-tree HIP library (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GGError("libgg_levelset.so is not built: run __graft_entry__.build() or "
                          "python gragles_3d_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ---- host set-up (no GPU) --------------------------------------------------
def derive_params(n_grains, points_per_grain, border, time_step_size=0.2):
    p = GGParams()
    rc = lib().gg_derive_params(int(n_grains), int(points_per_grain), int(border), float(time_step_size),
                                ctypes.byref(p))
    if rc != GG_OK:
        raise GGError("gg_derive_params failed (%d)" % rc)
    return p


def coarsening_plan(current_grains, initial_grains, loop, points_per_grain, real_domain_size, enabled=True,
                    gradient=0.95, time_step_size=0.2):
    """grainhdl::gridCoarsement's test (src/grainhdl.cpp:1447-1451): the new
    grid size sized for points_per_grain points per grain diameter at the
    current grain count, used when the count has fallen below gradient times
    the count at the last coarsening (or the start), after the first step, when
    it is smaller than the current grid.  Returns (new_size, new_dt) with the
    Gaussian mode's dt (src/grainhdl.cpp:1430-1434), or None."""
    import math
    new_size = int(math.pow(current_grains, 1 / 3.0) * math.pow(math.pi * 4 / 3, 1. / 3) / 2 * points_per_grain)
    if (float(current_grains) / float(initial_grains) < gradient and loop != 0 and enabled
            and new_size < real_domain_size):
        return new_size, time_step_size / float(new_size * new_size)
    return None


def uds_text(ids, bounds, h, volume, bunge=None, stored=None):
    """The Microstructure_<loop>.uds lines of saveNetworkAsVoxelContainer
    (src/grainhdl.cpp:1088-1107): per grain, tab-separated, ID, the box centre
    ((max - min) / 2 + min, integer division) times h for x, y, z, the three
    Bunge angles, xmin xmax ymin ymax zmin zmax of the output buffer, volume and
    stored elastic energy; doubles as C++ streams print them by default (six
    significant digits, %g).  bounds rows are {xmin,ymin,zmin,xmax,ymax,zmax};
    bunge (n x 3) and stored (n) default to 0, as this path carries no
    orientations or stored energy."""
    lines = []
    for r, gid in enumerate(ids):
        b = [int(v) for v in bounds[r]]
        cx, cy, cz = ((b[3] - b[0]) // 2 + b[0]) * h, ((b[4] - b[1]) // 2 + b[1]) * h, ((b[5] - b[2]) // 2 + b[2]) * h
        eul = (0.0, 0.0, 0.0) if bunge is None else tuple(float(x) for x in bunge[r])
        see = 0.0 if stored is None else float(stored[r])
        f = ["%d" % int(gid), "%g" % cx, "%g" % cy, "%g" % cz, "%g" % eul[0], "%g" % eul[1], "%g" % eul[2],
             "%d" % b[0], "%d" % b[3], "%d" % b[1], "%d" % b[4], "%d" % b[2], "%d" % b[5], "%g" % float(volume[r]),
             "%g" % see]
        lines.append("\t".join(f) + "\n")
    return "".join(lines)


def resize_to_cube(bounds6, ngridpoints):
    b = np.array(bounds6, dtype=np.int32)
    lib().gg_resize_to_cube(_p(b, ctypes.c_int32), int(ngridpoints))
    return b


def generate_voronoi(params, seed, with_idfield=True, nthreads=0):
    """Seeded Voronoi microstructure: (idfield[L,L,L] or None, bounds[N+1,6], exists[N+1])."""
    N, L = params.n_grains, params.ngridpoints
    idf = np.empty((L, L, L), dtype=np.int32) if with_idfield else None
    bounds = np.zeros((N + 1, 6), dtype=np.int32)
    exists = np.zeros(N + 1, dtype=np.uint8)
    rc = lib().gg_generate_voronoi(ctypes.byref(params), ctypes.c_uint64(seed),
                                   _p(idf, ctypes.c_int32) if idf is not None else None,
                                   _p(bounds, ctypes.c_int32), _p(exists, ctypes.c_uint8), int(nthreads))
    if rc != GG_OK:
        raise GGError("gg_generate_voronoi failed (%d)" % rc)
    return idf, bounds, exists


def bimodal_seeds(n_grains, seed, frac_large=0.1, size_ratio=4.0, sigma=0.2):
    """Seeds (n, 3) and radii (n,) of an uneven grain population: a fraction
    frac_large of the grains size_ratio times larger, each radius lognormal
    (sigma) about its mode, placed largest first by random sequential addition
    (gg_bimodal_seeds)."""
    xyz = np.zeros((n_grains, 3))
    r = np.zeros(n_grains)
    rc = lib().gg_bimodal_seeds(int(n_grains), ctypes.c_uint64(seed), float(frac_large), float(size_ratio),
                                float(sigma), _p(xyz, ctypes.c_double), _p(r, ctypes.c_double))
    if rc != GG_OK:
        raise GGError("gg_bimodal_seeds failed (%d)" % rc)
    return xyz, r


def generate_power(params, xyz, radius, with_idfield=True, nthreads=0):
    """Power-diagram microstructure of the given seeds and radii:
    (idfield[L,L,L] or None, bounds[N+1,6], exists[N+1]), conventions as generate_voronoi."""
    N, L = params.n_grains, params.ngridpoints
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    if xyz.shape != (N, 3) or radius.shape != (N,):
        raise GGError("need %d seeds and radii" % N)
    idf = np.empty((L, L, L), dtype=np.int32) if with_idfield else None
    bounds = np.zeros((N + 1, 6), dtype=np.int32)
    exists = np.zeros(N + 1, dtype=np.uint8)
    rc = lib().gg_generate_power(ctypes.byref(params), _p(xyz, ctypes.c_double), _p(radius, ctypes.c_double),
                                 _p(idf, ctypes.c_int32) if idf is not None else None,
                                 _p(bounds, ctypes.c_int32), _p(exists, ctypes.c_uint8), int(nthreads))
    if rc != GG_OK:
        raise GGError("gg_generate_power failed (%d)" % rc)
    return idf, bounds, exists


def voronoi_seeds(n_grains, seed):
    xyz = np.zeros((n_grains, 3))
    rc = lib().gg_voronoi_seeds(int(n_grains), ctypes.c_uint64(seed), _p(xyz, ctypes.c_double))
    if rc != GG_OK:
        raise GGError("gg_voronoi_seeds failed (%d)" % rc)
    return xyz


class GGPhysics(ctypes.Structure):
    """gg_physics: settings of the convolution correction."""
    _fields_ = [("correction", ctypes.c_int32), ("mobility_model", ctypes.c_int32),
                ("unique_energies", ctypes.c_int32), ("stored_energy", ctypes.c_int32),
                ("triple_line_drag", ctypes.c_double)]


def misorientation(qa, qb):
    """Cubic misorientation of two unit quaternions, floored at 1 degree (radians)."""
    a = np.ascontiguousarray(qa, dtype=np.float64)
    b = np.ascontiguousarray(qb, dtype=np.float64)
    return float(lib().gg_misorientation(_p(a, ctypes.c_double), _p(b, ctypes.c_double)))


def gb_pair_tables(quat, mobility_model=True, unique_energies=False, nthreads=0):
    """(energy, mobility) of every ordered id pair, each (n, n) for quat of shape (n, 4)."""
    q = np.ascontiguousarray(quat, dtype=np.float64)
    n = q.shape[0]
    gam, mob = np.zeros((n, n)), np.zeros((n, n))
    rc = lib().gg_gb_pair_tables(int(n), _p(q, ctypes.c_double), int(mobility_model), int(unique_energies),
                                 _p(gam, ctypes.c_double), _p(mob, ctypes.c_double), int(nthreads))
    if rc != GG_OK:
        raise GGError("gg_gb_pair_tables failed (%d)" % rc)
    return gam, mob


def euler_to_quaternion(euler):
    e = np.ascontiguousarray(euler, dtype=np.float64)
    q = np.zeros(4)
    lib().gg_euler_to_quaternion(_p(e, ctypes.c_double), _p(q, ctypes.c_double))
    return q


def random_orientations(n_grains, seed, reference_abs=False):
    """(n_grains+1, 4) unit quaternions: row 0 the wall (Euler 0, 35 deg, 0), then Shoemake's."""
    q = np.zeros((n_grains + 1, 4))
    rc = lib().gg_random_orientations(int(n_grains), ctypes.c_uint64(seed), int(reference_abs),
                                      _p(q, ctypes.c_double))
    if rc != GG_OK:
        raise GGError("gg_random_orientations failed (%d)" % rc)
    return q


def textured_orientations(n_grains, seed, fraction=0.45, scatter_deg=7.5):
    """(n_grains+1, 4) unit quaternions: random_orientations(n_grains, seed), then a
    fraction of the grains (drawn with a generator seeded by `seed`) replaced by
    rotations of one texture component (the first grain's orientation) by an
    angle uniform in [0, scatter_deg] about a uniformly random axis, as the
    reference's UseTexture scatters orientations about a texture component
    (src/grainhdl.cpp:93-96).  Two such grains are at most 2 * scatter_deg
    apart, so with 7.5 degrees their boundary is low-angle (below 15 degrees):
    about fraction^2 of the boundaries."""
    q = random_orientations(n_grains, seed)
    rng = np.random.default_rng(seed)
    ref = q[1].copy()
    pick = rng.random(n_grains) < fraction
    for g in np.nonzero(pick)[0] + 1:
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = np.deg2rad(scatter_deg) * rng.random()
        d = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])
        w1, x1, y1, z1 = ref
        w2, x2, y2, z2 = d
        q[g] = [w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2]
    return q


def lognormal_energies(n_grains, seed, median=0.1, sigma=1.0):
    """n_grains+1 stored energies (entry 0, the wall, is 0)."""
    v = np.zeros(n_grains + 1)
    rc = lib().gg_lognormal_energies(int(n_grains), ctypes.c_uint64(seed), float(median), float(sigma),
                                     _p(v, ctypes.c_double))
    if rc != GG_OK:
        raise GGError("gg_lognormal_energies failed (%d)" % rc)
    return v


def initial_neighbour_lists(params, bounds, exists):
    N = params.n_grains
    b = np.ascontiguousarray(bounds, dtype=np.int32)
    e = np.ascontiguousarray(exists, dtype=np.uint8)
    offs = np.zeros(N + 2, dtype=np.int32)
    n = ctypes.c_int64(0)
    lib().gg_initial_neighbour_lists(ctypes.byref(params), _p(b, ctypes.c_int32), _p(e, ctypes.c_uint8),
                                     _p(offs, ctypes.c_int32), None, 0, ctypes.byref(n))
    ids = np.zeros(max(int(n.value), 1), dtype=np.int32)
    rc = lib().gg_initial_neighbour_lists(ctypes.byref(params), _p(b, ctypes.c_int32), _p(e, ctypes.c_uint8),
                                          _p(offs, ctypes.c_int32), _p(ids, ctypes.c_int32), ids.size,
                                          ctypes.byref(n))
    if rc != GG_OK:
        raise GGError("gg_initial_neighbour_lists failed (%d)" % rc)
    return offs, ids[: int(n.value)]


# ---- device context -----------------------------------------------------------
class LevelSetContext:
    """All grain boxes of one simulation on one GPU."""

    def __init__(self, params, precision=FP64_PLAIN, max_list=512, device=0, arena_voxels=0):
        p = GGParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(params), ctypes.sizeof(GGParams))
        p.precision = int(precision)
        p.max_list = int(max_list)
        p.device = int(device)
        p.arena_voxels = int(arena_voxels)
        self.params = p
        self.n_grains = p.n_grains
        h = _vp()
        rc = lib().gg_create(ctypes.byref(p), ctypes.byref(h))
        if rc != GG_OK:
            raise GGError("gg_create failed with code %d" % rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != GG_OK:
            raise GGError("code %d: %s" % (rc, lib().gg_last_error(self.h).decode()))

    # set-up
    def set_boxes(self, bounds, exists):
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        e = np.ascontiguousarray(exists, dtype=np.uint8)
        self._chk(lib().gg_set_boxes(self.h, _p(b, ctypes.c_int32), _p(e, ctypes.c_uint8)))

    def init_distance(self, idfield):
        f = np.ascontiguousarray(idfield, dtype=np.int32)
        self._chk(lib().gg_init_distance(self.h, _p(f, ctypes.c_int32)))

    def init_distance_noalloc(self, idfield):
        """init_distance through the library's static scratch: no device allocation."""
        f = np.ascontiguousarray(idfield, dtype=np.int32)
        self._chk(lib().gg_init_distance_noalloc(self.h, _p(f, ctypes.c_int32)))

    def set_neighbour_lists(self, offsets, ids):
        o = np.ascontiguousarray(offsets, dtype=np.int32)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        if i.size == 0:
            i = np.zeros(1, dtype=np.int32)
        self._chk(lib().gg_set_neighbour_lists(self.h, _p(o, ctypes.c_int32), _p(i, ctypes.c_int32)))

    def build_neighbour_lists(self):
        """Initial neighbour lists from the boxes, built on the device (same lists as
        initial_neighbour_lists on the host)."""
        self._chk(lib().gg_build_neighbour_lists(self.h))

    # stages
    def switch(self):
        self._chk(lib().gg_switch_buffers(self.h))

    def convolution(self):
        self._chk(lib().gg_convolution(self.h))

    def update_lists(self):
        self._chk(lib().gg_update_neighbour_lists(self.h))

    def comparison(self):
        self._chk(lib().gg_comparison(self.h))

    def level_set(self):
        self._chk(lib().gg_level_set(self.h))

    def hull_elements(self, gid):
        """Interfacial elements of the grain's last hull (topology on): [(ids,
        triangles, mean barycenter (x, y, z) in grid coordinates)] sorted by ids."""
        n = ctypes.c_int32(0)
        self._chk(lib().gg_download_hull_elements(self.h, int(gid), None, None, None, None, 0, ctypes.byref(n)))
        k = max(n.value, 1)
        nids, ids8, ntri = np.zeros(k, np.int32), np.zeros(8 * k, np.int32), np.zeros(k, np.int32)
        pos = np.zeros(3 * k)
        self._chk(lib().gg_download_hull_elements(self.h, int(gid), _p(nids, ctypes.c_int32), _p(ids8, ctypes.c_int32),
                                                  _p(ntri, ctypes.c_int32), _p(pos, ctypes.c_double), k,
                                                  ctypes.byref(n)))
        return [(tuple(ids8[8 * r:8 * r + min(8, nids[r])].tolist()), int(ntri[r]), tuple(pos[3 * r:3 * r + 3]))
                for r in range(n.value)]

    def set_physics(self, correction=True, mobility_model=False, unique_energies=True, stored_energy=False,
                    triple_line_drag=0.0, quat=None, see=None):
        """gg_set_physics: the convolution correction for the following steps."""
        ph = GGPhysics(int(correction), int(mobility_model), int(unique_energies), int(stored_energy),
                       float(triple_line_drag))
        q = np.ascontiguousarray(quat, dtype=np.float64) if quat is not None else None
        e = np.ascontiguousarray(see, dtype=np.float64) if see is not None else None
        self._keep_phys = (q, e)
        self._chk(lib().gg_set_physics(self.h, ctypes.byref(ph), _p(q, ctypes.c_double) if q is not None else None,
                                       _p(e, ctypes.c_double) if e is not None else None))

    def correction_elements(self, gid):
        """The interfacial elements of grain gid's last hull in projection order: dict of arrays
        cls (3 high-order junction, 2 quadruple, 1 triple line, 0 boundary), nb (tuples of the
        bordering ids, ascending), nbary, pos (n, 3), energy, mobility."""
        n = ctypes.c_int32(0)
        self._chk(lib().gg_download_elements(self.h, int(gid), None, None, None, None, None, None, None, 0,
                                             ctypes.byref(n)))
        k = max(n.value, 1)
        cls, nnb, nb8, nbary = (np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(8 * k, np.int32),
                                np.zeros(k, np.int32))
        pos, en, mo = np.zeros(3 * k), np.zeros(k), np.zeros(k)
        self._chk(lib().gg_download_elements(self.h, int(gid), _p(cls, ctypes.c_int32), _p(nnb, ctypes.c_int32),
                                             _p(nb8, ctypes.c_int32), _p(nbary, ctypes.c_int32),
                                             _p(pos, ctypes.c_double), _p(en, ctypes.c_double),
                                             _p(mo, ctypes.c_double), k, ctypes.byref(n)))
        m = n.value
        return {"cls": cls[:m], "nb": [tuple(nb8[8 * e:8 * e + min(8, nnb[e])].tolist()) for e in range(m)],
                "nbary": nbary[:m], "pos": pos[:3 * m].reshape(m, 3), "energy": en[:m], "mobility": mo[:m]}

    def voxel_container(self):
        """saveNetworkAsVoxelContainer's id container (Container_size_*.raw
        data), [z][y][x] of side realDomainSize; call after a whole timestep."""
        L = ctypes.c_int32(0)
        self._chk(lib().gg_voxel_container(self.h, None, 0, ctypes.byref(L)))
        n = L.value
        out = np.zeros(n ** 3, dtype=np.uint32)
        self._chk(lib().gg_voxel_container(self.h, _p(out, ctypes.c_uint32), out.size, ctypes.byref(L)))
        return out.reshape(n, n, n)

    def save_network_as_voxel_container(self, loop, directory=".", bunge=None, stored=None):
        """grainhdl::saveNetworkAsVoxelContainer (src/grainhdl.cpp:1074-1154):
        writes Microstructure_<loop>.uds (alive grains, see uds_text),
        Container_size_<size>_t_<loop>.raw (the id container, uint32, x fastest)
        and ContainerVolumeEnergy_size_<size>_t_<loop>.raw, which the reference
        leaves empty (its energy values go to the already closed container
        file).  bunge / stored: per grain id (rows 0..n_grains), optional.
        Returns the three paths."""
        import os
        cont = self.voxel_container()
        size = cont.shape[0]
        st, vol, _ = self.stats()
        ids = np.nonzero(st == ALIVE)[0]
        ids = ids[ids > 0]
        b = self.all_bounds(1)[ids]
        h = float(self.params.h)
        text = uds_text(ids, b, h, vol[ids], None if bunge is None else np.asarray(bunge)[ids],
                        None if stored is None else np.asarray(stored)[ids])
        p_uds = os.path.join(directory, "Microstructure_%d.uds" % int(loop))
        p_raw = os.path.join(directory, "Container_size_%d_t_%d.raw" % (size, int(loop)))
        p_en = os.path.join(directory, "ContainerVolumeEnergy_size_%d_t_%d.raw" % (size, int(loop)))
        with open(p_uds, "w") as f:
            f.write(text)
        cont.astype("<u4").tofile(p_raw)
        open(p_en, "wb").close()
        return p_uds, p_raw, p_en

    def coarsen_grid(self, new_size, new_dt):
        """gridCoarsement's coarsening branch (src/grainhdl.cpp:1446-1492): the
        new grid (ngridpoints = new_size + 2*grid_blowup, h = 1/new_size,
        delta = grid_blowup/new_size, dt = new_dt), every grain resampled onto
        it, IDLocal recalculated and the contour extracted.  A coarsening step
        is coarsen_grid() followed by timestep()."""
        self._chk(lib().gg_coarsen_grid(self.h, int(new_size), float(new_dt)))
        # the host copy of the grid variables follows the library's
        p = self.params
        p.ngridpoints = int(new_size) + 2 * p.grid_blowup
        p.conv_domain_size = int(new_size)
        p.h = 1.0 / int(new_size)
        p.delta = p.grid_blowup / float(int(new_size))
        p.dt = float(new_dt)

    def redistancing(self):
        self._chk(lib().gg_redistancing(self.h))

    def count(self):
        c = ctypes.c_int32(0)
        self._chk(lib().gg_count_grains(self.h, ctypes.byref(c)))
        return c.value

    def timestep(self, nsteps=1, use_graph=False):
        self._chk(lib().gg_timestep(self.h, int(nsteps), 1 if use_graph else 0))

    def autotune(self, reps=3):
        """One timestep that also times the comparison's tile shapes per box-size class on its
        own inputs and keeps the fastest; returns (choice per class, ms[class][shape])."""
        ch = np.zeros(4, dtype=np.int32)
        ms = np.zeros((4, CMP_VARIANTS), dtype=np.float32)
        self._chk(lib().gg_autotune(self.h, int(reps), _p(ch, ctypes.c_int32), _p(ms, ctypes.c_float)))
        return ch, ms

    def tuning(self):
        v = np.zeros(4, dtype=np.int32)
        self._chk(lib().gg_get_tuning(self.h, _p(v, ctypes.c_int32)))
        return v

    def set_tuning(self, variants):
        v = np.ascontiguousarray(variants, dtype=np.int32)
        self._chk(lib().gg_set_tuning(self.h, _p(v, ctypes.c_int32)))

    def set_convolution_form(self, form):
        """CONV_DEFAULT or CONV_SPLIT_BF16 (FP32 batched path only): the circulant
        convolution as split-BF16 (3-term) matrix products on the matrix cores."""
        self._chk(lib().gg_set_convolution_form(self.h, int(form)))

    def set_work_order(self, order):
        """ORDER_SPATIAL (Morton order of the initial boxes) or ORDER_VOLUME (a
        device-side queue sorted by box side, largest first, rebuilt every plan)."""
        self._chk(lib().gg_set_work_order(self.h, int(order)))

    def work_order(self):
        v = ctypes.c_int32(0)
        self._chk(lib().gg_get_work_order(self.h, ctypes.byref(v)))
        return v.value

    def set_comparison_order(self, order, shape=None):
        """CMP_ORDER_GRAIN (one launch per size class, a grain's tiles together)
        or CMP_ORDER_SPACE (one launch, tiles sorted by the cell of their
        centre, so grains that share neighbour voxels read them together);
        shape 0..3 of the spatial launch (None: keep the current one)."""
        if shape is None:
            shape = self.comparison_order()[1]
        self._chk(lib().gg_set_comparison_order(self.h, int(order), int(shape)))

    def comparison_order(self):
        """(order, shape of the spatial order's launch)."""
        v, sh = ctypes.c_int32(0), ctypes.c_int32(0)
        self._chk(lib().gg_get_comparison_order(self.h, ctypes.byref(v), ctypes.byref(sh)))
        return v.value, sh.value

    def synchronize(self):
        self._chk(lib().gg_synchronize(self.h))

    # inspection
    def count_history(self):
        n = ctypes.c_int32(0)
        self._chk(lib().gg_get_count_history(self.h, None, 0, ctypes.byref(n)))
        buf = np.zeros(max(n.value, 1), dtype=np.int32)
        self._chk(lib().gg_get_count_history(self.h, _p(buf, ctypes.c_int32), buf.size, ctypes.byref(n)))
        return buf[: n.value]

    def stats(self):
        N = self.n_grains
        st = np.zeros(N + 1, dtype=np.int32)
        v = np.zeros(N + 1)
        a = np.zeros(N + 1)
        self._chk(lib().gg_get_grain_stats(self.h, _p(st, ctypes.c_int32), _p(v, ctypes.c_double),
                                           _p(a, ctypes.c_double)))
        return st, v, a

    def bounds(self, gid, which):
        b = np.zeros(6, dtype=np.int32)
        self._chk(lib().gg_get_bounds(self.h, int(gid), int(which), _p(b, ctypes.c_int32)))
        return b

    def phi(self, gid, which):
        n = ctypes.c_int64(0)
        self._chk(lib().gg_download_distance(self.h, int(gid), int(which), None, 0, ctypes.byref(n)))
        buf = np.empty(max(n.value, 1), dtype=np.float64)
        self._chk(lib().gg_download_distance(self.h, int(gid), int(which), _p(buf, ctypes.c_double), buf.size,
                                             ctypes.byref(n)))
        return self.bounds(gid, which), buf[: n.value]

    def upload_phi(self, gid, which, values):
        """Overwrite grain gid's distance buffer (which: 0 input, 1 output) with
        values (its box volume of doubles, x fastest)."""
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._chk(lib().gg_upload_phi(self.h, int(gid), int(which), _p(v, ctypes.c_double), v.size))

    def idlocal(self, gid):
        b = np.zeros(6, dtype=np.int32)
        n = ctypes.c_int64(0)
        self._chk(lib().gg_download_idlocal(self.h, int(gid), _p(b, ctypes.c_int32), None, 0, ctypes.byref(n)))
        buf = np.empty(max(n.value, 1), dtype=np.uint32)
        self._chk(lib().gg_download_idlocal(self.h, int(gid), _p(b, ctypes.c_int32), _p(buf, ctypes.c_uint32),
                                            buf.size, ctypes.byref(n)))
        return b, buf[: n.value]

    def neighbour_list(self, gid, which=0):
        n = ctypes.c_int32(0)
        self._chk(lib().gg_download_list(self.h, int(gid), int(which), None, 0, ctypes.byref(n)))
        buf = np.zeros(max(n.value, 1), dtype=np.int32)
        self._chk(lib().gg_download_list(self.h, int(gid), int(which), _p(buf, ctypes.c_int32), buf.size,
                                         ctypes.byref(n)))
        return buf[: n.value].copy()

    def enable_hull_topology(self, on=True):
        """Collect, at every level set, the grains each hull touches (off by default)."""
        self._chk(lib().gg_enable_hull_topology(self.h, 1 if on else 0))

    def hull_neighbour_counts(self):
        """Per grain: number of grains its last hull touched (itself included) and
        whether the domain wall was among them."""
        cnt = np.zeros(self.params.n_grains + 1, dtype=np.int32)
        wall = np.zeros(self.params.n_grains + 1, dtype=np.uint8)
        self._chk(lib().gg_get_hull_neighbour_counts(self.h, _p(cnt, ctypes.c_int32), _p(wall, ctypes.c_uint8)))
        return cnt, wall.astype(bool)

    def element_counts(self):
        """(n_grains+1, 4): boundaries, triple lines, quadruple and high-order
        junctions after mergeJunction (correction on)."""
        out = np.zeros((self.params.n_grains + 1, 4), dtype=np.int32)
        self._chk(lib().gg_get_element_counts(self.h, _p(out, ctypes.c_int32)))
        return out

    def faces(self):
        """get_Faces of every alive grain (what Faces_<loop>.bin holds): arrays
        grain_a, grain_b (grain_a > grain_b), area (physical units)."""
        n = ctypes.c_int64(0)
        self._chk(lib().gg_get_faces(self.h, None, None, None, 0, ctypes.byref(n)))
        k = max(n.value, 1)
        a, b, ar = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(k)
        self._chk(lib().gg_get_faces(self.h, _p(a, ctypes.c_int32), _p(b, ctypes.c_int32), _p(ar, ctypes.c_double),
                                     k, ctypes.byref(n)))
        m = n.value
        return a[:m], b[:m], ar[:m]

    def hull_junction_counts(self):
        """Per grain (rows 0..n_grains): distinct cell id lists of its last hull with
        2, 3, 4 and other numbers of grains (boundaries, triple lines, quadruple
        and higher-order junctions)."""
        c4 = np.zeros((self.params.n_grains + 1, 4), dtype=np.int32)
        self._chk(lib().gg_get_hull_junction_counts(self.h, _p(c4, ctypes.c_int32)))
        return c4

    def hull_neighbours(self, gid):
        """The grains the last hull of gid touched, in the reference's order."""
        buf = np.zeros(64, dtype=np.int32)
        n = ctypes.c_int32(0)
        self._chk(lib().gg_download_hull_neighbours(self.h, int(gid), _p(buf, ctypes.c_int32), buf.size,
                                                    ctypes.byref(n)))
        return buf[: n.value].copy()

    def all_bounds(self, which):
        """Rows 0..n_grains of every grain's buffer bounds (which=0 input, 1 output)."""
        b = np.zeros((self.params.n_grains + 1, 6), dtype=np.int32)
        self._chk(lib().gg_get_all_bounds(self.h, int(which), _p(b, ctypes.c_int32)))
        return b

    def all_neighbour_lists(self, which=0):
        """Every grain's list as CSR (offsets[g]..offsets[g+1] = grain g)."""
        offs = np.zeros(self.params.n_grains + 2, dtype=np.int32)
        n = ctypes.c_int64(0)
        self._chk(lib().gg_download_all_lists(self.h, int(which), _p(offs, ctypes.c_int32), None, 0, ctypes.byref(n)))
        ids = np.zeros(max(n.value, 1), dtype=np.int32)
        self._chk(lib().gg_download_all_lists(self.h, int(which), _p(offs, ctypes.c_int32), _p(ids, ctypes.c_int32),
                                              ids.size, ctypes.byref(n)))
        return offs, ids[:n.value]

    def new_bounds(self):
        b = np.zeros((self.n_grains + 1, 6), dtype=np.int32)
        self._chk(lib().gg_get_new_bounds(self.h, _p(b, ctypes.c_int32)))
        return b

    def box_voxels(self):
        v = ctypes.c_int64(0)
        self._chk(lib().gg_box_voxels(self.h, ctypes.byref(v)))
        return int(v.value)

    def enable_stage_timing(self, on=True):
        self._chk(lib().gg_enable_stage_timing(self.h, 1 if on else 0))

    def stage_times(self, reset=True):
        """{segment: (total ms, timed passes)} since the last reset."""
        ms = np.zeros(len(SEGMENTS))
        n = np.zeros(len(SEGMENTS), dtype=np.int64)
        self._chk(lib().gg_get_stage_times(self.h, _p(ms, ctypes.c_double), _p(n, ctypes.c_int64), len(SEGMENTS),
                                           1 if reset else 0))
        return {k: (float(a), int(b)) for k, a, b in zip(SEGMENTS, ms, n)}


def setup_simulation(n_grains, points_per_grain, border, seed, precision=FP64_PLAIN, time_step_size=0.2,
                     max_list=512, nthreads=0):
    """Voronoi microstructure -> boxes -> initial distance -> neighbour lists, on the GPU.

    Mirrors grainhdl::initializeSimulation + run_sim's prologue
    (src/grainhdl.cpp:58-127, 832-834).  Returns (ctx, params, inputs) where
    inputs holds the host arrays used, so an independent checker can start from
    the same state.
    """
    p = derive_params(n_grains, points_per_grain, border, time_step_size)
    idf, bounds, exists = generate_voronoi(p, seed, True, nthreads)
    offs, ids = initial_neighbour_lists(p, bounds, exists)
    ctx = LevelSetContext(p, precision=precision, max_list=max_list)
    ctx.set_boxes(bounds, exists)
    ctx.init_distance(idf)
    ctx.set_neighbour_lists(offs, ids)
    return ctx, p, dict(idfield=idf, bounds=bounds, exists=exists, offsets=offs, ids=ids)
