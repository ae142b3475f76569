"""The convolution correction (src/box.cpp:428-536) on the GPU against the
oracle: the plain FP64 path bit for bit (distance buffers, IDLocal, lists,
states after every stage or step), the interfacial elements grain by grain
(class, bordering grains, triangle count, energy and mobility bit for bit;
junction positions to 1e-9 grid units, as their barycenter sums are fixed
point), the
graph-replayed step equal to the eager one, and the FP32 path within its
tolerance.

Each case prints its name, seed and shape before touching the GPU.
"""
import sys

import numpy as np
import pytest

import gragles_3d_amd as G
from oracle import oracle as O
import physics_cases as PC
from parity import compare_all, step_both

pytestmark = pytest.mark.gpu

TL_DRAG = 1e6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(case, phys, precision=G.FP64_PLAIN):
    ctx = PC.gpu_context(case, precision)
    orc = PC.oracle_sim(O, case)
    ctx.set_physics(**phys)
    o = dict(phys)
    o["tl_drag"] = o.pop("triple_line_drag")
    orc.set_physics(**o)
    return ctx, orc


def _compare_elements(ctx, orc, where):
    st = orc.states()
    n_el = 0
    counts = ctx.element_counts()
    for g in np.nonzero(st == O.ALIVE)[0]:
        eg = ctx.correction_elements(g)
        eo = orc.correction_elements(g)
        assert len(eo) == len(eg["cls"]), (where, g, len(eo), len(eg["cls"]))
        want = [sum(1 for e in eo if e[0] == k) for k in (0, 1, 2, 3)]  # GB, TL, QJ, HJ after merging
        assert list(counts[g]) == want, (where, g, list(counts[g]), want)
        for e, (c, nb, nbary, pos, en, mo) in enumerate(eo):
            ids = sorted(set(orc.element_ids(g, e).tolist()))
            assert eg["cls"][e] == c, (where, g, e)
            assert list(eg["nb"][e]) == ids[:8], (where, g, e, eg["nb"][e], ids)
            assert eg["nbary"][e] == nbary, (where, g, e)
            assert eg["energy"][e] == en and eg["mobility"][e] == mo, (where, g, e, eg["energy"][e], en)
            if c >= 2:
                np.testing.assert_allclose(eg["pos"][e], pos, rtol=0, atol=1e-9, err_msg="%s g%d e%d" % (where, g, e))
        n_el += len(eo)
    return n_el


def _orientation_physics(p, see=True):
    q = PC.low_angle_orientations(p.n_grains, 11)
    s = G.lognormal_energies(p.n_grains, 3, median=0.1, sigma=1.0)
    return dict(correction=True, mobility_model=True, unique_energies=False, stored_energy=see,
                triple_line_drag=TL_DRAG, quat=q, see=s)


def test_sphere_in_matrix_stagewise():
    _say("correction: ball in a matrix, FP64, stage by stage")
    case = PC.sphere_in_matrix()
    p = case[0]
    q = np.zeros((3, 4))
    q[0] = [1, 0, 0, 0]
    q[1] = PC.axis_angle([1, 2, 2], 0.2)
    q[2] = PC.quat_mul(q[1], PC.axis_angle([0, 1, 0], np.deg2rad(9)))
    phys = dict(correction=True, mobility_model=True, unique_energies=False, stored_energy=True,
                triple_line_drag=TL_DRAG, quat=q, see=np.array([0.0, 0.3, 0.05]))
    ctx, orc = _pair(case, phys)
    for s in range(4):
        step_both(ctx, orc, compare_every_stage=True, tag="step %d" % s)
        assert _compare_elements(ctx, orc, "step %d" % s) >= 3
    ctx.close()
    orc.close()


@pytest.mark.parametrize("n_grains,ppg,seed,see", [(27, 10, 5, True), (64, 8, 2, False)])
def test_voronoi_anisotropic_fp64(n_grains, ppg, seed, see):
    _say("correction: Voronoi n=%d ppg=%d seed=%d, Read-Shockley + mobility model, SEE %s, FP64"
         % (n_grains, ppg, seed, see))
    case = PC.voronoi_case(n_grains, ppg, 3, seed)
    ctx, orc = _pair(case, _orientation_physics(case[0], see))
    for s in range(5):
        step_both(ctx, orc, compare_every_stage=(s < 2), tag="step %d" % s)
        compare_all(ctx, orc, "step %d" % s)
        assert _compare_elements(ctx, orc, "step %d" % s) > 0
    ctx.close()
    orc.close()


def test_isotropic_stored_energy_fp64():
    # configuration 5's physics: unit energies and mobilities, lognormal stored
    # energies, the reference's default TripleLineDrag (0: triple lines and
    # junctions get zero mobility)
    _say("correction: Voronoi n=27 ppg=10 seed=7, isotropic + SEE, TripleLineDrag 0, FP64")
    case = PC.voronoi_case(27, 10, 3, 7)
    p = case[0]
    phys = dict(correction=True, mobility_model=False, unique_energies=True, stored_energy=True,
                triple_line_drag=0.0, quat=None, see=G.lognormal_energies(p.n_grains, 5, 0.1, 1.0))
    ctx, orc = _pair(case, phys)
    for s in range(4):
        step_both(ctx, orc, tag="step %d" % s)
        compare_all(ctx, orc, "step %d" % s)
        _compare_elements(ctx, orc, "step %d" % s)
    ctx.close()
    orc.close()


def test_graph_replay_equals_eager():
    _say("correction: graph replay vs eager, Voronoi n=27 ppg=10 seed=5, FP64")
    case = PC.voronoi_case(27, 10, 3, 5)
    phys = _orientation_physics(case[0])
    a = PC.gpu_context(case)
    b = PC.gpu_context(case)
    a.set_physics(**phys)
    b.set_physics(**phys)
    a.timestep(6, use_graph=False)
    for _ in range(3):
        b.timestep(2, use_graph=True)
        b.synchronize()
    sa, sb = a.stats(), b.stats()
    assert np.array_equal(sa[0], sb[0])
    for g in np.nonzero(sa[0] == G.ALIVE)[0]:
        assert np.array_equal(a.phi(g, 1)[1], b.phi(g, 1)[1]), g
        assert np.array_equal(a.idlocal(g)[1], b.idlocal(g)[1]), g
    a.close()
    b.close()


def test_correction_off_again_restores_plain_path():
    _say("correction: switched off again, Voronoi n=27 ppg=10 seed=5, FP64")
    case = PC.voronoi_case(27, 10, 3, 5)
    ctx = PC.gpu_context(case)
    orc = PC.oracle_sim(O, case)
    ctx.set_physics(**_orientation_physics(case[0]))
    ctx.set_physics(correction=False)
    for s in range(2):
        step_both(ctx, orc, tag="step %d" % s)
        compare_all(ctx, orc, "step %d" % s)
    ctx.close()
    orc.close()


def test_fp32_with_correction_near_fp64():
    # the FP64 path with the correction equals the oracle (tests above)
    _say("correction: FP32 batched vs FP64, Voronoi n=64 ppg=8 seed=2, 5 steps")
    from fp32_error import phi_volume_errors
    case = PC.voronoi_case(64, 8, 3, 2)
    p = case[0]
    phys = _orientation_physics(p)
    c64 = PC.gpu_context(case, G.FP64_PLAIN)
    c32 = PC.gpu_context(case, G.FP32_BATCHED)
    c64.set_physics(**phys)
    c32.set_physics(**phys)
    c64.timestep(5)
    c32.timestep(5)
    e = phi_volume_errors(c32, c64, p.delta)
    _say("fp32 errors: %s" % e)
    assert e["grains_with_equal_boxes"] >= 0.9 * e["grains_compared"]
    assert e["phi_rms_err_in_band"] < 1e-4 * p.delta
    assert e["volume_rel_err_max"] < 1e-3
    c64.close()
    c32.close()
