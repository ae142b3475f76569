"""The oracle's convolution correction (src/box.cpp:428-536) and its
grain-boundary model, pinned without the reference:

* misorientation: against a brute force over the 24 rotations of the cube and
  textbook cases (rotations about <100> by multiples of 90 degrees are
  symmetry operations; 60 degrees about <111> is the twin misorientation);
* Read-Shockley energy and sigmoidal mobility: closed forms;
* the library's host tables (gg_gb_pair_tables) equal the oracle's functions;
* interfacial elements of a Voronoi run: classes by list size, counts against
  the hull's independent list census, grain-boundary / triple-line /
  quadruple-junction energies and mobilities from the brute-force
  misorientation, the merge rule's invariant;
* the correction itself: on a ball inside a matrix every band voxel is
  corrected with its pair's mobility times energy, and the stored-energy shift
  uses the energy (GBInfo's swapped fields), both compared bit for bit with
  the closed form; on a Voronoi run, sampled band voxels against a brute-force
  nearest-barycenter search over the elements.
"""
import numpy as np
import pytest

from oracle import oracle as O
import physics_cases as PC


def test_misorientation_matches_24_operator_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b = rng.normal(size=4), rng.normal(size=4)
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        assert abs(O.misorientation(a, b) - PC.misorientation_brute_force(a, b)) < 1e-9


def test_misorientation_textbook_cases():
    one_deg = np.pi / 180
    base = PC.axis_angle([0.3, -0.5, 0.8], 0.7)
    cases = [
        (PC.axis_angle([0, 0, 1], np.pi / 2), one_deg),          # a cube symmetry: floored at 1 degree
        (PC.axis_angle([1, 0, 0], np.pi), one_deg),
        (PC.axis_angle([0, 0, 1], np.deg2rad(30)), np.deg2rad(30)),
        (PC.axis_angle([1, 1, 1], np.deg2rad(60)), np.deg2rad(60)),  # Sigma-3 twin
        (PC.axis_angle([0, 0, 1], np.deg2rad(50)), np.deg2rad(40)),  # 50 about <100> = 40 by symmetry
    ]
    for d, want in cases:
        got = O.misorientation(base, PC.quat_mul(base, d))
        assert abs(got - want) < 1e-7, (d, got, want)


def test_read_shockley_and_mobility_closed_forms():
    ref = 15 * np.pi / 180
    assert O.read_shockley(ref, False) == 1.0
    assert O.read_shockley(0.3, True) == 1.0
    assert abs(O.read_shockley(ref / np.e, False) - 2 / np.e) < 1e-15
    assert O.read_shockley(2 * ref, False) == 1.0
    assert abs(O.mobility(ref, True) - (1 - 0.99 * np.exp(-5))) < 1e-15
    assert O.mobility(np.pi / 180, True) == 0.1  # 1 - 0.99 exp(-tiny) < 0.1: floored
    assert O.mobility(0.5, False) == 1.0
    for th in np.linspace(np.pi / 180, 1.0, 37):
        assert abs(O.read_shockley(th, False) - PC.read_shockley(th)) < 1e-14
        assert abs(O.mobility(th, True) - PC.sigmoid_mobility(th)) < 1e-14


def test_library_pair_tables_equal_the_oracle():
    G = pytest.importorskip("gragles_3d_amd")
    q = G.random_orientations(12, 9)
    gam, mob = G.gb_pair_tables(q, mobility_model=True, unique_energies=False, nthreads=2)
    for a in range(13):
        for b in range(13):
            th = O.misorientation(q[a], q[b])
            assert G.misorientation(q[a], q[b]) == th
            assert gam[a, b] == O.read_shockley(th, False)
            assert mob[a, b] == O.mobility(th, True)
    g1, m1 = G.gb_pair_tables(q, mobility_model=False, unique_energies=True)
    assert (g1 == 1).all() and (m1 == 1).all()


def test_orientation_helpers():
    G = pytest.importorskip("gragles_3d_amd")
    q = G.random_orientations(50, 4)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    # the wall: Euler (0, 35 deg, 0) is a 35 degree rotation about x
    assert np.allclose(q[0], PC.axis_angle([1, 0, 0], np.deg2rad(35)))
    qa = G.random_orientations(50, 4, reference_abs=True)
    assert (qa[1:] >= 0).all()
    see = G.lognormal_energies(2000, 7, median=0.1, sigma=1.0)
    assert see[0] == 0.0 and (see[1:] > 0).all()
    assert abs(np.median(np.log(see[1:])) - np.log(0.1)) < 0.1
    assert abs(np.std(np.log(see[1:])) - 1.0) < 0.08


# ---- elements --------------------------------------------------------------
TL_DRAG = 1e6


def _triple(q, me, a, b, ds):
    mis = lambda u, v: PC.misorientation_brute_force(q[u], q[v])
    avg = (PC.sigmoid_mobility(mis(me, a)) + PC.sigmoid_mobility(mis(a, b)) + PC.sigmoid_mobility(mis(me, b))) / 3
    mobility = 1 / (1 / (3 * np.sqrt(3 * ds * ds) * TL_DRAG) + 1 / avg)
    sigma = PC.read_shockley(mis(me, a)) - PC.read_shockley(mis(a, b)) + PC.read_shockley(mis(me, b))
    return min(max(sigma, 0.01), 1.0), mobility


@pytest.fixture(scope="module")
def voronoi_run():
    case = PC.voronoi_case()
    p = case[0]
    q = PC.low_angle_orientations(p.n_grains, 11)
    see = np.linspace(0.0, 0.2, p.n_grains + 1)
    see[0] = 0.0
    orc = PC.oracle_sim(O, case)
    orc.set_physics(True, mobility_model=True, unique_energies=False, stored_energy=True, tl_drag=TL_DRAG,
                    quat=q, see=see)
    ref = PC.oracle_sim(O, case)  # the same run without the correction
    for sim in (orc, ref):
        sim.timestep(1)
    yield case, q, see, orc, ref
    orc.close()
    ref.close()


def test_elements_follow_the_hull_lists(voronoi_run):
    case, q, _, orc, _ = voronoi_run
    p = case[0]
    ds = 1.0 / p.conv_domain_size
    n_tl = n_qj = n_hj = 0
    for gid in range(1, p.n_grains + 1):
        if orc.state(gid) != O.ALIVE:
            continue
        els = orc.correction_elements(gid)
        lists = orc.hull_elements(gid)  # the hull's distinct id lists (independent census)
        cls = [e[0] for e in els]
        assert cls == sorted(cls, reverse=True)  # HJ, QJ, TL, GB
        assert sum(e[2] for e in els) == sum(l[1] for l in lists)
        by_size = {}
        for ids, _, _ in lists:
            by_size[len(ids)] = by_size.get(len(ids), 0) + 1
        assert cls.count(0) == by_size.get(2, 0)
        assert cls.count(1) == by_size.get(3, 0)
        assert cls.count(2) + cls.count(3) <= by_size.get(4, 0) + sum(v for k, v in by_size.items() if k not in (2, 3, 4))
        qpos = []
        for c, nb, nbary, pos, en, mo in els:
            if c == 0:
                th = PC.misorientation_brute_force(q[gid], q[nb[0]])
                assert abs(en - PC.read_shockley(th)) < 1e-9 and abs(mo - PC.sigmoid_mobility(th)) < 1e-9
            elif c == 1:
                e_, m_ = _triple(q, gid, nb[0], nb[1], ds)
                assert abs(en - e_) < 1e-9 and abs(mo - m_) < 1e-9
                n_tl += 1
            elif c == 2 and len(nb) == 3:
                t = [_triple(q, gid, nb[0], nb[1], ds), _triple(q, gid, nb[1], nb[2], ds),
                     _triple(q, gid, nb[0], nb[2], ds)]
                assert abs(en - sum(x[0] for x in t) / 3) < 1e-9 and abs(mo - sum(x[1] for x in t) / 3) < 1e-9
                qpos.append(np.array(pos))
                # an unmerged quadruple junction sits at its list's mean barycenter
                match = [l for l in lists if tuple(sorted(l[0])) == tuple(sorted(nb + (gid,)))]
                assert len(match) == 1 and np.allclose(pos, match[0][2], rtol=0, atol=1e-9)
                n_qj += 1
            else:
                assert en == 1.0 and mo == 1.0
                n_hj += 1
        for i in range(len(qpos)):  # mergeJunction leaves no two quadruple junctions closer than 2
            for j in range(i + 1, len(qpos)):
                assert np.linalg.norm(qpos[i] - qpos[j]) >= 2
    assert n_tl > 0 and n_qj + n_hj > 0  # the run exercises junctions


def _band_expected(orc, ref, gid, p, q_pair, see, brute=None, limit=None):
    """(expected, got) corrected output values of grain gid's band voxels after
    the convolution of orc (correction on) and ref (off), from the pair weights
    q_pair(gid, id) or the nearest-element search brute(gid, id, point)."""
    b, phi_in = orc.phi(gid, 0)
    _, out_c = orc.phi(gid, 1)
    _, out_n = ref.phi(gid, 1)
    ib, idl = orc.idlocal(gid)
    assert np.array_equal(ib, b)  # IDLocal was resized to the output bounds
    nx, ny = b[3] - b[0], b[4] - b[1]
    band = np.nonzero(np.abs(phi_in) < p.delta)[0]
    if limit is not None:
        band = band[:: max(1, band.size // limit)]
    exp = np.empty(band.size)
    for n_, li in enumerate(band):
        idv = int(idl[li])
        if brute is None:
            w_e, w_m = q_pair(gid, idv)
        else:
            z, rem = divmod(int(li), nx * ny)
            y, x = divmod(rem, nx)
            w_e, w_m = brute(gid, idv, np.array([y + b[1], x + b[0], z + b[2]], dtype=np.float64))
        val = phi_in[li]
        o = val + (out_n[li] - val) * w_e * w_m
        f = 0.0
        if see is not None and (idv == 0 or orc.state(idv) != O.DELETED):
            f = (see[gid] - see[idv]) * w_m * p.dt / ((1 / 1.21) * (1 / 1.21))
        exp[n_] = o - f
    return exp, out_c[band]


def test_sphere_correction_closed_form():
    case = PC.sphere_in_matrix()
    p = case[0]
    q = np.zeros((3, 4))
    q[0] = [1, 0, 0, 0]
    q[1] = PC.axis_angle([1, 2, 2], 0.2)
    q[2] = PC.quat_mul(q[1], PC.axis_angle([0, 1, 0], np.deg2rad(9)))  # a 9 degree boundary
    see = np.array([0.0, 0.3, 0.05])
    for stored in (False, True):
        orc, ref = PC.oracle_sim(O, case), PC.oracle_sim(O, case)
        orc.set_physics(True, mobility_model=True, unique_energies=False, stored_energy=stored, tl_drag=TL_DRAG,
                        quat=q, see=see)
        for sim in (orc, ref):
            sim.timestep(1)
            sim.switch()
            sim.convolution()
        th12 = O.misorientation(q[1], q[2])
        assert abs(th12 - np.deg2rad(9)) < 1e-9
        for gid in (1, 2):
            els = orc.correction_elements(gid)
            assert all(e[0] == 0 for e in els)  # grain boundaries only
            assert sorted(e[1][0] for e in els) == ([2] if gid == 1 else [0, 1])

            def pair(g, idv):
                if idv not in {e[1][0] for e in els}:
                    return 1.0, 1.0
                th = O.misorientation(q[g], q[idv])
                return O.mobility(th, True), O.read_shockley(th, False)  # (w_energy, w_mobility) = (m, gamma)

            exp, got = _band_expected(orc, ref, gid, p, pair, see if stored else None)
            assert exp.size > 100
            assert np.array_equal(exp, got)
            _, out_c = orc.phi(gid, 1)
            _, out_n = ref.phi(gid, 1)
            assert (out_c != out_n).sum() > 100  # the low-angle weights (< 1) changed the band
        orc.close()
        ref.close()


def test_voronoi_correction_against_brute_force_projection(voronoi_run):
    case, q, see, orc, ref = voronoi_run
    p = case[0]
    for sim in (orc, ref):
        sim.switch()
        sim.convolution()
    checked = 0
    for gid in range(1, p.n_grains + 1, 4):
        if orc.state(gid) != O.ALIVE:
            continue
        els = orc.correction_elements(gid)
        bary = [orc.element_barycenters(gid, e) for e in range(len(els))]
        members = [set(orc.element_ids(gid, e).tolist()) for e in range(len(els))]

        def brute(g, idv, point):
            best, w = 10000000.0, (1.0, 1.0)
            for (c, nb, _, _, en, mo), B, ids in zip(els, bary, members):
                if idv not in ids:
                    continue
                d = ((point - B) ** 2).sum(axis=1)
                k = int(np.argmin(d))
                if d[k] < best:
                    best, w = d[k], (mo, en)
            return w

        exp, got = _band_expected(orc, ref, gid, p, None, see, brute=brute, limit=300)
        # squared distances are summed in numpy's order here: equal to the oracle's
        # to an ulp, which can only matter at an exact tie between elements
        assert np.allclose(exp, got, rtol=0, atol=1e-15)
        checked += exp.size
    assert checked > 500
