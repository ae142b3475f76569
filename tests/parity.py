"""GPU-vs-oracle parity helpers shared by the GPU tests and __graft_entry__.smoke().

Both sides start from the same host inputs (seeded Voronoi microstructure,
boxes, initial neighbour lists) and run the same stage sequence; compare_all()
checks every grain's state, buffer bounds, distance buffers, IDLocal and
neighbour lists.  FP64 comparisons are bit-exact (np.array_equal on the raw
doubles, so -0.0 == 0.0 but NaN would fail); volume and area, whose summation
order differs, are compared to a relative 1e-11.
"""
import numpy as np

import gragles_3d_amd as G
from oracle import oracle as O

STAGE_SEQUENCE = ("switch", "convolution", "switch", "update_lists", "comparison", "switch", "level_set",
                  "redistancing")


def make_pair(n_grains, ppg, border, seed, precision=G.FP64_PLAIN, max_list=512):
    p = G.derive_params(n_grains, ppg, border)
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    ctx = G.LevelSetContext(p, precision=precision, max_list=max_list)
    ctx.set_boxes(bounds, exists)
    ctx.init_distance(idf)
    ctx.set_neighbour_lists(offs, ids)
    orc = O.OracleSim(p.ngridpoints, p.grid_blowup, p.conv_domain_size, p.h, p.delta, p.dt, n_grains)
    orc.set_boxes(bounds, exists)
    orc.init_distance(idf)
    orc.set_lists(offs, ids)
    return ctx, orc, p


def run_stage(sim, name):
    getattr(sim, name)()


class Mismatch(AssertionError):
    pass


def compare_all(ctx, orc, where, check_lists=True, check_idlocal=True, check_phi=(1,), rtol_stats=1e-11):
    """Defaults fit the state after a whole timestep (or after init): the
    output buffer holds the redistanced field, the input buffer is stale."""
    st_g, vol_g, area_g = ctx.stats()
    st_o = orc.states()
    if not np.array_equal(st_g, st_o):
        bad = np.nonzero(st_g != st_o)[0]
        raise Mismatch("%s: grain states differ at ids %s (gpu %s, oracle %s)" % (where, bad[:10], st_g[bad[:10]],
                                                                               st_o[bad[:10]]))
    vol_o, area_o, _ = orc.stats()
    alive = np.nonzero(st_o == O.ALIVE)[0]
    for arr_g, arr_o, nm in ((vol_g, vol_o, "volume"), (area_g, area_o, "area")):
        a, b = arr_g[alive], arr_o[alive]
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        if err.size and err.max() > rtol_stats:
            i = int(np.argmax(err))
            raise Mismatch("%s: %s of grain %d differs: gpu %r oracle %r" % (where, nm, alive[i], a[i], b[i]))
    for gid in alive:
        for which in (0, 1):
            bo, _ = orc.bounds(gid, which)
            bg = ctx.bounds(gid, which)
            if not np.array_equal(bo, bg):
                raise Mismatch("%s: grain %d buffer %d bounds gpu %s oracle %s" % (where, gid, which, bg, bo))
        for which in check_phi:
            bo, po = orc.phi(gid, which)
            bg, pg = ctx.phi(gid, which)
            if not np.array_equal(bo, bg):
                raise Mismatch("%s: grain %d buffer %d bounds gpu %s oracle %s" % (where, gid, which, bg, bo))
            if po.size != pg.size or not np.array_equal(po, pg):
                d = np.nonzero(po != pg)[0] if po.size == pg.size else []
                raise Mismatch("%s: grain %d buffer %d values differ (%d of %d, first at %s: gpu %r oracle %r)" %
                               (where, gid, which, len(d), po.size, d[:1],
                                pg[d[0]] if len(d) else None, po[d[0]] if len(d) else None))
        if check_idlocal:
            bo, io = orc.idlocal(gid)
            bg, ig = ctx.idlocal(gid)
            if not np.array_equal(bo, bg) or not np.array_equal(io, ig):
                d = np.nonzero(io != ig)[0] if io.size == ig.size else []
                raise Mismatch("%s: grain %d IDLocal differs (bounds %s/%s, %d voxels)" % (where, gid, bg, bo,
                                                                                           len(d)))
        if check_lists:
            lo = orc.neighbour_list(gid, 0)
            lg = ctx.neighbour_list(gid, 0)
            if not np.array_equal(lo, lg):
                raise Mismatch("%s: grain %d neighbour list differs: gpu %s oracle %s" % (where, gid, lg[:12],
                                                                                        lo[:12]))
    return len(alive)


# What is defined after each stage.  A resized buffer keeps stale contents in
# the reference (std::vector resize), which no later stage reads, so it is not
# compared: the output buffer after level_set and the input buffer after
# redistancing.  The comparison list becomes the neighbour list only at
# level_set on the GPU (one selector flip per grain), so lists are compared
# from there on.
_CHECKS = {
    "convolution": dict(check_phi=(0, 1), check_idlocal=False, check_lists=False),
    "update_lists": dict(check_phi=(), check_idlocal=False, check_lists=False),
    "comparison": dict(check_phi=(0, 1), check_idlocal=True, check_lists=False),
    "level_set": dict(check_phi=(0,), check_idlocal=True, check_lists=True),
    "redistancing": dict(check_phi=(1,), check_idlocal=True, check_lists=True),
}


def compare_lists_pending(ctx, orc, where):
    """Comparison lists just built by update_lists (which=1 on both sides)."""
    st = orc.states()
    for gid in np.nonzero(st == O.ALIVE)[0]:
        lo = orc.neighbour_list(gid, 1)
        lg = ctx.neighbour_list(gid, 1)
        if not np.array_equal(lo, lg):
            raise Mismatch("%s: grain %d comparison list differs: gpu %s oracle %s" % (where, gid, lg[:12], lo[:12]))


def step_both(ctx, orc, compare_every_stage=False, tag=""):
    for name in STAGE_SEQUENCE:
        run_stage(ctx, name)
        run_stage(orc, name)
        if compare_every_stage and name != "switch":
            compare_all(ctx, orc, "%s after %s" % (tag, name), **_CHECKS[name])
            if name == "update_lists":
                compare_lists_pending(ctx, orc, "%s after %s" % (tag, name))
