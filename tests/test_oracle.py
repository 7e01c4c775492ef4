"""Pins the C oracle (oracle/ls1_oracle.c) before anything is compared to it.

Checks, each chosen so a plausible restatement mistake fails one of them:
* closed-form KATs of VectorizedCellProcessorTest (U0: F = +-24, U = 0,
  W = 96; F0: F = 0, U = -4, W = 0);
* golden vectors written by the reference's own code (tests/golden/
  ref_golden.npz from oracle/_ref: potforce.h + FullMolecule::calcFM on the
  reference's Vectorization*.inp fixtures and on seeded periodic systems);
* the reference's Comp2Param tables and Component masses / inertia;
* FullMolecule::upd_preF / upd_postF outputs;
* finite differences (F = -dU/dr, torque = -e x dU/de) for all 7 bodies;
* brute-force minimum-image sums against the C08 + halo traversal;
* exhaustive point-in-cell search for the cell index, and a
  CellProcessorStub-style count that C08 visits every neighbour pair once.
"""
import itertools
import pathlib

import numpy as np
import pytest

import scenarios
from ls1inp import read_inp
from oracle_lib import Oracle, RefLegacy, macro_to_upot_virial

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
G = np.load(GOLD / "ref_golden.npz")

FIXTURES = [
    ("VectorizationLennardJones", 35.0),
    ("VectorizationLennardJones1CLJ", 35.0),
    ("VectorizationCharge", 35.0),
    ("VectorizationChargeDipole", 35.0),
    ("VectorizationChargeQuadrupole", 35.0),
    ("VectorizationDipole", 35.0),
    ("VectorizationDipoleQuadrupole", 35.0),
    ("VectorizationQuadrupole", 35.0),
    ("VectorizationWater", 6.16),
    ("VectorizationMultiComponentMultiPotentials", 35.0),
    ("VectorizationMultiComponentMultiPotentials_50_molecules", 35.0),
]

TOL = 1e-12  # relative to the largest reference magnitude of each quantity


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def test_kat_u0():
    ps = read_inp(GOLD / "inp" / "ForceCalculationTestU0.inp")
    o = Oracle.from_phase_space(ps, 1.1, periodic=False)
    res = o.forces(ps.r, ps.q, ps.cid)
    expect = np.array([[-24, -24, 0], [24, -24, 0], [-24, 24, 0], [24, 24, 0]], float)
    np.testing.assert_allclose(res["F"], expect, atol=1e-8)
    u, w = macro_to_upot_virial(res["macro"])
    assert abs(u) < 1e-8 and abs(w - 96.0) < 1e-8


def test_kat_f0():
    ps = read_inp(GOLD / "inp" / "ForceCalculationTestF0.inp")
    o = Oracle.from_phase_space(ps, 1.3, periodic=False)
    res = o.forces(ps.r, ps.q, ps.cid)
    assert np.abs(res["F"]).max() < 1e-7
    u, w = macro_to_upot_virial(res["macro"])
    assert abs(u + 4.0) < 1e-8 and abs(w) < 1e-6


@pytest.mark.parametrize("name,rc", FIXTURES)
def test_fixture_vs_reference_legacy(name, rc):
    ps = read_inp(GOLD / "inp" / f"{name}.inp")
    o = Oracle.from_phase_space(ps, rc, periodic=False)
    res = o.forces(ps.r, ps.q, ps.cid)
    for k in ("F", "M", "Vi"):
        ref = G[f"{name}/{k}"]
        if np.abs(ref).max() == 0:
            assert np.abs(res[k]).max() == 0
        else:
            assert rel_err(res[k], ref) < TOL, k
    u, w = macro_to_upot_virial(res["macro"])
    ur, wr = G[f"{name}/upot_virial"]
    assert abs(u - ur) <= TOL * max(abs(ur), 1e-300)
    assert abs(w - wr) <= TOL * max(abs(wr), 1e-300)
    # brute force over all pairs agrees with the C08 traversal
    bf = o.forces(ps.r, ps.q, ps.cid, brute=True)
    for k in ("F", "M", "Vi"):
        if np.abs(res[k]).max() > 0:
            assert rel_err(bf[k], res[k]) < TOL


@pytest.mark.parametrize("name,rc", FIXTURES)
def test_tables_and_inertia_vs_reference(name, rc):
    ps = read_inp(GOLD / "inp" / f"{name}.inp")
    o = Oracle.from_phase_space(ps, rc, periodic=False)
    m, I, rot = o.component_props()
    np.testing.assert_array_equal(m, G[f"{name}/mass"])
    np.testing.assert_array_equal(I, G[f"{name}/inertia"])
    np.testing.assert_array_equal(rot, G[f"{name}/rotdof"])
    e, s, sh = o.lj_tables()
    np.testing.assert_array_equal(e, G[f"{name}/eps24"])
    np.testing.assert_array_equal(s, G[f"{name}/sig2"])
    np.testing.assert_array_equal(sh, G[f"{name}/shift6"])


@pytest.mark.parametrize("tag", ["periodic_lj", "periodic_mix"])
def test_periodic_traversal_vs_reference(tag):
    desc = G[f"{tag}/comp_desc"]
    rc, eps_rf = G[f"{tag}/rc_epsrf"]
    ncomp = len(scenarios.parse_components(desc))
    mix = np.array([[1.0, 1.0]] * (ncomp * (ncomp - 1) // 2))
    if ncomp == 3:
        mix[1] = (0.97, 1.02)
    o = Oracle(ncomp, desc, mix, eps_rf, rc, rc, G[f"{tag}/box"], periodic=True)
    res = o.forces(G[f"{tag}/r"], G[f"{tag}/q"], G[f"{tag}/cid"])
    for k in ("F", "M", "Vi"):
        assert rel_err(res[k], G[f"{tag}/{k}"]) < TOL, k
    u, w = macro_to_upot_virial(res["macro"])
    ur, wr = G[f"{tag}/upot_virial"]
    assert abs(u - ur) <= TOL * abs(ur) and abs(w - wr) <= TOL * abs(wr)


def test_leapfrog_vs_reference():
    desc = G["leapfrog/comp_desc"]
    ncomp = len(scenarios.parse_components(desc))
    mix = np.array([[1.0, 1.0]] * (ncomp * (ncomp - 1) // 2))
    o = Oracle(ncomp, desc, mix, 1.0, 2.5, 2.5, np.full(3, 100.0), periodic=True)
    cid = G["leapfrog/cid"]
    r1, v1, q1, L1 = o.leapfrog_pre(0.002, G["leapfrog/r0"], G["leapfrog/v0"], G["leapfrog/q0"],
                                    G["leapfrog/L0"], G["leapfrog/F"], G["leapfrog/M"], cid)
    for a, k in ((r1, "r1"), (v1, "v1"), (q1, "q1"), (L1, "L1")):
        np.testing.assert_allclose(a, G[f"leapfrog/{k}"], rtol=0, atol=1e-14)
    v2, L2, sums = o.leapfrog_post(0.002, v1, L1, q1, G["leapfrog/F"], G["leapfrog/M"], cid)
    np.testing.assert_allclose(v2, G["leapfrog/v2"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(L2, G["leapfrog/L2"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(sums[:2], G["leapfrog/sums"], rtol=1e-13)


# ------------------------------------------------------- finite differences

def _energy(kind, out):
    return out[9] / 6.0 if kind == 0 else out[9]


@pytest.mark.parametrize("kind", range(7))
def test_site_pair_finite_differences(kind):
    rng = np.random.default_rng(100 + kind)
    for _ in range(5):
        ri = rng.normal(size=3)
        rj = ri + rng.normal(size=3)
        rj = ri + (rj - ri) / np.linalg.norm(rj - ri) * rng.uniform(0.9, 2.0)
        ei = rng.normal(size=3); ei /= np.linalg.norm(ei)
        ej = rng.normal(size=3); ej /= np.linalg.norm(ej)
        pi, pj = rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5)
        args = dict(eps24=24.0, sig2=1.0, shift6=0.0, eps_rf_invrc3=0.0)
        base = Oracle.site_pair(kind, ri, ei, pi, rj, ej, pj, **args)
        h = 1e-6
        grad = np.zeros(3)
        for d in range(3):
            dp, dm = ri.copy(), ri.copy()
            dp[d] += h
            dm[d] -= h
            up = _energy(kind, Oracle.site_pair(kind, dp, ei, pi, rj, ej, pj, **args))
            um = _energy(kind, Oracle.site_pair(kind, dm, ei, pi, rj, ej, pj, **args))
            grad[d] = (up - um) / (2 * h)
        np.testing.assert_allclose(base[:3], -grad, rtol=2e-6, atol=1e-7 * np.abs(base[:3]).max())
        # torques: first site oriented for DD, DQ, QQ; second for CD, DD, CQ, DQ, QQ
        for which, e, slot in ((0, ei, 3), (1, ej, 6)):
            oriented = (kind in (3, 5, 6)) if which == 0 else (kind in (2, 3, 4, 5, 6))
            if not oriented:
                continue
            g = np.zeros(3)
            for d in range(3):
                ep, em = e.copy(), e.copy()
                ep[d] += h
                em[d] -= h
                if which == 0:
                    up = _energy(kind, Oracle.site_pair(kind, ri, ep, pi, rj, ej, pj, **args))
                    um = _energy(kind, Oracle.site_pair(kind, ri, em, pi, rj, ej, pj, **args))
                else:
                    up = _energy(kind, Oracle.site_pair(kind, ri, ei, pi, rj, ep, pj, **args))
                    um = _energy(kind, Oracle.site_pair(kind, ri, ei, pi, rj, em, pj, **args))
                g[d] = (up - um) / (2 * h)
            tau = -np.cross(e, g)
            np.testing.assert_allclose(base[slot:slot + 3], tau, rtol=2e-6,
                                       atol=1e-7 * max(np.abs(tau).max(), 1e-3))


def test_reaction_field_term():
    """D-D reaction-field energy term rffac*e1e2 = p1 p2 epsRFInvrc3 e1.e2."""
    e1 = np.array([0.0, 0.0, 1.0]); e2 = np.array([0.0, 0.6, 0.8])
    out = Oracle.site_pair(3, np.zeros(3), e1, 1.5, np.array([1.0, 0.2, 0.1]), e2, 0.7,
                           eps_rf_invrc3=0.01)
    assert abs(out[10] - 1.5 * 0.7 * 0.01 * 0.8) < 1e-15


# ----------------------------------------------------- enumeration / cells

@pytest.mark.parametrize("sysname", ["small_lj", "small_mixture"])
def test_periodic_brute_force(sysname):
    s = getattr(scenarios, sysname)(seed=21)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    a = o.forces(s.r, s.q, s.cid)
    b = o.forces(s.r, s.q, s.cid, brute=True)
    for k in ("F", "M", "Vi"):
        assert rel_err(a[k], b[k]) < TOL
    ua, wa = macro_to_upot_virial(a["macro"])
    ub, wb = macro_to_upot_virial(b["macro"])
    assert abs(ua - ub) < TOL * abs(ub) and abs(wa - wb) < TOL * abs(wb)
    # Newton 3: total force vanishes
    assert np.abs(a["F"].sum(0)).max() < 1e-10 * np.abs(a["F"]).max()


def test_cell_index_bruteforce():
    rng = np.random.default_rng(5)
    for box, rc, mode in (((10.0, 7.3, 12.1), 2.5, 0), ((37.6386,) * 3, 2.5, 1), ((9.0, 9.0, 9.0), 3.0, 0)):
        o = Oracle(1, scenarios.comp_1clj(), None, 1.0, rc, rc, np.array(box), periodic=True,
                   cell_mode=mode)
        cpd, cl, hmin, hmax = o.cell_geometry()
        pts = rng.uniform(-cl, np.array(box) + cl, size=(3000, 3))
        # grid-aligned points and box faces are the rounding edge cases
        k = rng.integers(0, cpd - 1, size=(500, 3))
        pts = np.vstack([pts, k * cl, np.array([0.0, 0.0, 0.0])[None], np.array(box)[None] - 1e-15])
        idx = o.cell_index(pts)

        def border(i, d):
            if i == 0:
                return hmin[d]
            if i == 1:
                return 0.0
            if i == cpd[d] - 1:
                return box[d]
            if i == cpd[d]:
                return hmax[d]
            return i * cl[d] + hmin[d]

        for p, ci in zip(pts, idx):
            ijk = (ci % cpd[0], (ci // cpd[0]) % cpd[1], ci // (cpd[0] * cpd[1]))
            for d in range(3):
                inside = border(ijk[d], d) <= p[d] < border(ijk[d] + 1, d)
                clamped = (ijk[d] == 0 and p[d] < border(1, d)) or (
                    ijk[d] == cpd[d] - 1 and p[d] >= border(cpd[d] - 1, d))
                assert inside or clamped, (p, ijk, d)


def test_c08_covers_every_neighbour_pair_once():
    """LinkedCellsTest::testTraversalMethods idea: count visits per pair."""
    o = Oracle(1, scenarios.comp_1clj(), None, 1.0, 2.5, 2.5, np.array([10.0, 12.5, 7.5]))
    cpd, _, _, _ = o.cell_geometry()
    pairs = o.c08_pairs()
    seen = {}
    for c1, c2 in pairs:
        key = (min(c1, c2), max(c1, c2))
        seen[key] = seen.get(key, 0) + 1
    assert all(v == 1 for v in seen.values())

    def ijk(c):
        return (c % cpd[0], (c // cpd[0]) % cpd[1], c // (cpd[0] * cpd[1]))

    def halo(t):
        return any(t[d] == 0 or t[d] == cpd[d] - 1 for d in range(3))

    expect = set()
    cells = list(itertools.product(range(cpd[0]), range(cpd[1]), range(cpd[2])))
    for a in cells:
        for d in itertools.product((-1, 0, 1), repeat=3):
            b = tuple(a[k] + d[k] for k in range(3))
            if any(b[k] < 0 or b[k] >= cpd[k] for k in range(3)):
                continue
            if halo(a) and halo(b):
                continue
            ca = a[0] + cpd[0] * (a[1] + cpd[1] * a[2])
            cb = b[0] + cpd[0] * (b[1] + cpd[1] * b[2])
            expect.add((min(ca, cb), max(ca, cb)))
    assert set(seen) == expect
    for c1, c2 in pairs:  # the non-halo cell is handed over first
        assert not halo(ijk(c1))


def test_wrap_semantics():
    o = Oracle(1, scenarios.comp_1clj(), None, 1.0, 2.5, 2.5, np.array([10.0, 10.0, 10.0]))
    r = np.array([[-1e-17, 10.0, 5.0], [-0.5, 10.5, 9.999999999999998]])
    w = o.wrap(r)
    assert np.all(w >= 0) and np.all(w < 10.0)
    assert w[0, 1] == 0.0 and w[0, 0] == np.nextafter(10.0, 0)


def test_energy_conservation_short_run():
    """Truncated-shifted LJ (continuous energy) from an FCC start: the
    kick-drift-kick step order conserves U + K to a few 1e-5 over 300 steps;
    a dropped half kick or a mis-ordered drift breaks this by orders."""
    s = scenarios.lj_fcc(ncell=5)
    desc = np.array(scenarios.comp_1clj(rc_shift=s.rc, shifted=True))
    o = Oracle(1, desc, None, 1.0, s.rc, s.rc, s.box)
    res = o.simulate(300, 0.002, s.r, s.v, s.q, s.L, s.cid)
    ps = res["per_step"]
    e = ps[:, 0] + 0.5 * (ps[:, 2] + ps[:, 3])
    assert np.abs(e - e[0]).max() / abs(e[0]) < 1e-4


def test_threaded_traversal_matches_serial():
    """The OpenMP C08 walk (CPU baseline timing only) gives the serial walk's
    forces to rounding of the macroscopic sums' order, and equal counters."""
    s = scenarios.small_mixture(n=600, seed=29)
    a = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box).forces(s.r, s.q, s.cid)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    o.set_threads(4)
    b = o.forces(s.r, s.q, s.cid)
    for k in ("F", "M", "Vi"):
        np.testing.assert_array_equal(a[k], b[k])  # per-molecule sums keep their order
    np.testing.assert_allclose(b["macro"], a["macro"], rtol=1e-13)
    np.testing.assert_array_equal(a["counters"], b["counters"])


def test_reference_check_count():
    """ox_count_checks: the checks with macroscopic sums equal the traversal's
    own molecule-pair counter; all checks add the boundary-halo pairs seen
    from the second periodic side (none in an open system)."""
    s = scenarios.small_lj(n=500, seed=31)
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box)
    ref = o.forces(s.r, s.q, s.cid)
    allc, macro = o.count_checks(s.r, s.q, s.cid)
    assert macro == ref["counters"][0] and allc > macro
    n = s.n
    # brute force over minimum images: unordered pairs whose cells are neighbours
    # are all checked once with macroscopic sums
    op = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box, periodic=False)
    a2, m2 = op.count_checks(s.r, s.q, s.cid)
    assert a2 == m2 and m2 <= n * (n - 1) // 2
