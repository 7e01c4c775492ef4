"""The oracle's NVT tail against what fixes it independently of the code:
closed forms of the homogeneous long-range correction, quadrature of the
orientation-averaged site-site integrals it is built from, the reaction-field
self term, and the velocity-scaling thermostat's scaling laws and explosion
capping (Homogeneous.cpp:21-196, Domain.cpp:157-313,
VelocityScalingThermostat.cpp:70-90). CPU only."""
import math
import pathlib
import sys

import numpy as np
import pytest
from scipy import integrate

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "ls1-mardyn_amd"))

import scenarios  # noqa: E402
from oracle_lib import Oracle  # noqa: E402


def lj_comp(sites, shifted=False, rc_shift=2.5):
    """Component record with LJ centres [(x, y, z, m, eps, sigma)]."""
    d = [len(sites), 0, 0, 0]
    for x, y, z, m, e, s in sites:
        d += [x, y, z, m, e, s, rc_shift, 1.0 if shifted else 0.0]
    return d + [0.0, 0.0, 0.0]


def oracle(desc, ncomp=1, rc=2.5, rc_lj=2.5, eps_rf=1.0, mix=None, box=(10.0, 10.0, 10.0)):
    return Oracle(ncomp, np.array(desc, np.float64), mix, eps_rf, rc, rc_lj, np.array(box))


def textbook_lj(N, V, eps, sig, rc):
    """U_lrc = 2 pi N rho int_rc^inf r^2 u(r) dr and W_lrc = 3 V P_lrc for
    u = 4 eps ((s/r)^12 - (s/r)^6) (Allen & Tildesley, eqs. 2.144-2.145)."""
    rho = N / V
    u = 8.0 * math.pi * N * rho * eps * (sig ** 12 / (9 * rc ** 9) - sig ** 6 / (3 * rc ** 3))
    p = 16.0 / 3.0 * math.pi * rho ** 2 * eps * sig ** 3 * (
        2.0 / 3.0 * (sig / rc) ** 9 - (sig / rc) ** 3)
    return u, 3.0 * V * p


def test_lrc_single_centre_closed_form():
    eps, sig, rc, N, box = 1.3, 1.1, 2.7, 5000.0, (20.0, 21.0, 22.0)
    o = oracle(lj_comp([(0, 0, 0, 1.0, eps, sig)]), rc=rc, rc_lj=rc, box=box)
    V = float(np.prod(box))
    out = o.lrc_homogeneous([N], V)
    u, w = textbook_lj(N, V, eps, sig, rc)
    assert abs(out[0] - u) < 1e-13 * abs(u)
    assert abs(out[1] - w) < 1e-13 * abs(w)
    assert abs(out[2] - w / (3 * N)) < 1e-13 * abs(w / (3 * N))


def test_lrc_skips_shifted_and_cross_component_pairs():
    """Only unshifted table entries contribute; cross-component entries are
    always truncated-shifted (Comp2Param.cpp:79-105), so a binary mixture's
    correction is the sum of its two pure-component terms."""
    rc, box = 2.5, (15.0, 15.0, 15.0)
    V = float(np.prod(box))
    desc = lj_comp([(0, 0, 0, 1.0, 1.0, 1.0)]) + lj_comp([(0, 0, 0, 1.0, 0.85, 1.1)])
    o = oracle(desc, ncomp=2, rc=rc, rc_lj=rc, box=box, mix=np.array([[1.0, 1.0]]))
    n1, n2 = 1200.0, 800.0
    out = o.lrc_homogeneous([n1, n2], V)
    # pure-component terms: (pi rho / 3N) N_i^2 24 eps_i (...) = 8 pi N_i^2 / V
    # eps_i (...), the single-component form of N_i molecules in V
    u1, w1 = textbook_lj(n1, V, 1.0, 1.0, rc)
    u2, w2 = textbook_lj(n2, V, 0.85, 1.1, rc)
    assert abs(out[0] - (u1 + u2)) < 1e-12 * abs(u1 + u2)
    assert abs(out[1] - (w1 + w2)) < 1e-12 * abs(w1 + w2)
    oshift = oracle(lj_comp([(0, 0, 0, 1.0, 1.0, 1.0)], shifted=True), rc=rc, rc_lj=rc, box=box)
    assert np.all(oshift.lrc_homogeneous([1000.0], V) == 0.0)


_GL_X, _GL_W = np.polynomial.legendre.leggauss(48)


def _site_average(n, sig2, r, t1, t2):
    """<(s^2/sigma^2)^n> and its r-derivative for s = |r z + a - b|, |a| = t1,
    |b| = t2 with independent uniform orientations (a zero offset is a site
    at the centre). 48-point Gauss-Legendre over the two cosines: the
    integrand is analytic in them for r > t1 + t2, so the rule is exact to
    rounding."""
    if t1 == 0.0 and t2 == 0.0:
        s2 = r * r
        return (s2 / sig2) ** n, n * (s2 / sig2) ** (n - 1) * 2 * r / sig2
    if t1 == 0.0 or t2 == 0.0:
        t = t1 or t2
        s2 = r * r + t * t + 2 * r * t * _GL_X
        ds2 = 2 * r + 2 * t * _GL_X
        f = (s2 / sig2) ** n
        df = n * (s2 / sig2) ** (n - 1) * ds2 / sig2
        return 0.5 * (_GL_W @ f), 0.5 * (_GL_W @ df)
    u1, u2 = _GL_X[:, None], _GL_X[None, :]
    w = _GL_W[:, None] * _GL_W[None, :]
    R2 = r * r + t1 * t1 + 2 * r * t1 * u1
    R = np.sqrt(R2)
    s2 = R2 + t2 * t2 - 2 * R * t2 * u2
    ds2 = (2 * R - 2 * t2 * u2) * (r + t1 * u1) / R
    f = (s2 / sig2) ** n
    df = n * (s2 / sig2) ** (n - 1) * ds2 / sig2
    return 0.25 * (w * f).sum(), 0.25 * (w * df).sum()


def _radial(n, sig2, rc, t1, t2):
    """int_rc^inf r^2 g_n(r) dr and int_rc^inf r^3 g_n'(r) dr (the virial
    form: r . grad on the molecule-centre separation) by adaptive quadrature."""
    tu = integrate.quad(lambda r: r * r * _site_average(n, sig2, r, t1, t2)[0], rc, np.inf,
                        epsabs=0, epsrel=1e-13, limit=200)[0]
    tv = integrate.quad(lambda r: r ** 3 * _site_average(n, sig2, r, t1, t2)[1], rc, np.inf,
                        epsabs=0, epsrel=1e-13, limit=200)[0]
    return tu, tv


@pytest.mark.parametrize("sites", [
    [(0.0, 0.0, 0.3, 0.5, 1.0, 1.0), (0.0, 0.0, -0.3, 0.5, 1.0, 1.0)],   # site-site (TISS)
    [(0.0, 0.0, 0.0, 0.6, 1.2, 0.9), (0.25, 0.0, 0.0, 0.4, 0.8, 1.05)],  # centre-site (TICS)
])
def test_lrc_site_integrals_by_quadrature(sites):
    """The reference's closed-form radial integrals against direct quadrature
    of the orientation-averaged site-site LJ energy and virial beyond rc."""
    rc, N, box = 2.6, 3000.0, (18.0, 18.0, 18.0)
    V = float(np.prod(box))
    o = oracle(lj_comp(sites), rc=rc, rc_lj=rc, box=box)
    out = o.lrc_homogeneous([N], V)
    e24, s2, _ = o.lj_tables()
    taus = [math.sqrt(x * x + y * y + z * z) for x, y, z, *_ in sites]
    U = W = 0.0
    for a in range(len(sites)):
        for b in range(len(sites)):
            u12, v12 = _radial(-6, s2[a, b], rc, taus[a], taus[b])
            u6, v6 = _radial(-3, s2[a, b], rc, taus[a], taus[b])
            U += N * N * e24[a, b] * (u12 - u6)
            W += N * N * e24[a, b] * (v12 - v6)
    fac = math.pi * (N / V) / (3 * N)
    assert abs(out[0] - fac * U) < 1e-11 * abs(fac * U)
    assert abs(out[1] - (-fac * W)) < 1e-11 * abs(fac * W)


def test_lrc_reaction_field_self_term():
    """-1/2 epsRFInvrc3 N mu^2 for the molecule dipole mu = sum q r + point
    dipoles, times 3 in the virial (Homogeneous.cpp:122-136)."""
    rc, eps_rf, N, box = 2.5, 4.0, 1000.0, (12.0, 12.0, 12.0)
    q, d, mu_pt = 0.7, 0.3, 0.5
    # an LJ centre (carries the mass), charges +q / -q at z = +-d/2, a point
    # dipole along x; LJ shifted so only the self term remains
    desc = [1, 2, 1, 0,
            0, 0, 0, 1.0, 1.0, 1.0, rc, 1.0,
            0, 0, d / 2, 0.0, q,
            0, 0, -d / 2, 0.0, -q,
            0, 0, 0, 2.0, 0.0, 0.0, mu_pt,
            0.0, 0.0, 0.0]
    o = oracle(desc, rc=rc, rc_lj=rc, eps_rf=eps_rf, box=box)
    out = o.lrc_homogeneous([N], float(np.prod(box)))
    mu2 = (q * d) ** 2 + mu_pt ** 2
    e = 2 * (eps_rf - 1) / (rc ** 3 * (2 * eps_rf + 1))
    assert abs(out[0] - (-0.5 * e * N * mu2)) < 1e-14
    assert abs(out[1] - 3 * (-0.5 * e * N * mu2)) < 1e-14


def test_lrc_rejects_too_small_cutoff():
    o = oracle(lj_comp([(0, 0, 1.3, 0.5, 1, 1), (0, 0, -1.3, 0.5, 1, 1)]), rc=2.5, rc_lj=2.5)
    with pytest.raises(RuntimeError):
        o.lrc_homogeneous([10.0], 1000.0)


def test_thermostat_scaling_laws():
    Ti = 1.1
    N, rot = 1000.0, 2000.0
    # translational T = Ti, rotational T = Ti: both betas exactly 1
    tp = Oracle.thermostat_params([3 * N * Ti, rot * Ti, N, rot], Ti)
    assert tp[1] == 1.0 and tp[2] == 1.0 and tp[3] == 0
    assert abs(tp[0] - Ti) < 1e-15
    # 25 % too hot: beta = 0.8^0.4 (> 0.9, no explosion)
    tp = Oracle.thermostat_params([3 * N * Ti * 1.25, rot * Ti / 1.25, N, rot], Ti)
    assert abs(tp[1] - 0.8 ** 0.4) < 1e-15 and abs(tp[2] - 1.25 ** 0.4) < 1e-15
    assert tp[3] == 0
    assert abs(tp[0] - (3 * N * 1.25 + rot / 1.25) * Ti / (3 * N + rot)) < 1e-14
    # twice too hot: beta_trans 0.5^0.4 < 0.9 -> explosion: beta_trans 1,
    # beta_rot = raw^0.0091
    tp = Oracle.thermostat_params([3 * N * Ti * 2.0, rot * Ti * 2.0, N, rot], Ti)
    assert tp[3] == 1 and tp[1] == 1.0
    assert abs(tp[2] - (0.5 ** 0.4) ** 0.0091) < 1e-15
    # heuristics off: plain betas
    tp = Oracle.thermostat_params([3 * N * Ti * 2.0, rot * Ti * 2.0, N, rot], Ti, heuristics=False)
    assert tp[3] == 0 and abs(tp[1] - 0.5 ** 0.4) < 1e-15
    # no rotational DOF: sum I w^2 ignored, beta_rot 1
    tp = Oracle.thermostat_params([3 * N * Ti, 123.0, N, 0.0], Ti)
    assert tp[2] == 1.0 and abs(tp[0] - Ti) < 1e-15
    # no target: betas 1
    tp = Oracle.thermostat_params([3 * N * Ti * 5, 0.0, N, 0.0], 0.0)
    assert tp[1] == 1.0 and tp[2] == 1.0 and tp[3] == 0


def test_velocity_scaling_and_explosion_capping():
    """Capping brings an over-hot molecule's kinetic energy to exactly the
    limit (10 Ti translational, 3 rotDOF Ti rotational) and scales F / M by
    the same factor; everyone is then scaled by the betas."""
    Ti = 0.8
    desc = np.array(scenarios.comp_1clj() + scenarios.comp_2cljq(), np.float64)
    o = Oracle(2, desc, np.array([[1.0, 1.0]]), 1.0, 3.0, 3.0, np.array([10.0, 10.0, 10.0]))
    m, I, rotdof = o.component_props()
    cid = np.array([0, 0, 1, 1], np.int32)
    rng = np.random.default_rng(3)
    q = scenarios.random_quaternions(rng, 4)
    v = np.array([[0.5, 0.1, -0.2], [8.0, -3.0, 2.0], [0.2, 0.3, 0.1], [0.1, 0.0, 0.2]])
    L = np.zeros((4, 3))
    # molecule 3: large spin about a body axis with nonzero inertia
    L[3] = scenarios.quat_rotate(q[3:4], np.array([[3.0, 0.0, 0.0]]))[0]
    F = rng.normal(size=(4, 3))
    M = rng.normal(size=(4, 3))
    tp = np.array([0.0, 0.97, 0.95, 1.0])
    v2, L2, F2, M2 = o.velocity_scaling(v, L, F, M, q, cid, tp, Ti)
    ut = 0.5 * m[cid] * (v * v).sum(1)
    assert ut[1] > 10 * Ti and ut[0] < 10 * Ti
    c = math.sqrt(10 * Ti / ut[1])
    np.testing.assert_allclose(v2[1], v[1] * c * tp[1], rtol=1e-15)
    np.testing.assert_allclose(F2[1], F[1] * c, rtol=1e-15)
    np.testing.assert_array_equal(F2[[0, 2, 3]], F[[0, 2, 3]])
    u_after = 0.5 * m[cid[1]] * ((v2[1] / tp[1]) ** 2).sum()
    assert abs(u_after - 10 * Ti) < 1e-13
    # rotation of molecule 3 (2CLJQ, rotDOF 2): capped to 3 * 2 * Ti
    w = scenarios.quat_rotate(q[3:4] * np.array([1, -1, -1, -1]), L[3:4])[0]
    inv = np.where(I[1] > 0, 1.0 / np.where(I[1] > 0, I[1], 1.0), 0.0)
    ur = 0.5 * (I[1] * (w * inv) ** 2).sum()
    assert rotdof[1] == 2 and ur > 3 * 2 * Ti
    dc = math.sqrt(3 * 2 * Ti / ur)
    np.testing.assert_allclose(L2[3], L[3] * dc * tp[2], rtol=1e-14)
    np.testing.assert_allclose(M2[3], M[3] * dc, rtol=1e-14)
    np.testing.assert_allclose(v2[[0, 2, 3]], v[[0, 2, 3]] * tp[1], rtol=1e-15)
    # no explosion flag: only the betas
    v3, L3, F3, _ = o.velocity_scaling(v, L, F, M, q, cid, np.array([0.0, 0.97, 0.95, 0.0]), Ti)
    np.testing.assert_allclose(v3, v * 0.97, rtol=1e-15)
    np.testing.assert_array_equal(F3, F)


def test_nvt_simulation_records_and_final_scaling():
    """An LJ fluid under the velocity-scaling thermostat: each record's
    temperature and betas follow from that step's sums, and the final
    velocities are the last step's post-kick velocities times beta_trans."""
    s = scenarios.lj_fcc(ncell=4, T=1.0, seed=5)
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box)
    res = o.simulate_nvt(30, 0.002, s.r, s.v, s.q, s.L, s.cid, T_target=0.9, lrc=True)
    per = res["per_step"]
    for row in per:
        tp = Oracle.thermostat_params([row[2], row[3], s.n, 0.0], 0.9)
        assert abs(tp[0] - row[6]) <= 1e-15 * row[6]
        assert tp[1] == row[7] and tp[2] == row[8]
    assert np.all(per[:, 7] != 1.0) and np.all(per[:, 7] > 0.9)  # scaled, no explosion
    mv2 = (res["v"] ** 2).sum()  # unit masses
    assert abs(mv2 - per[-1, 7] ** 2 * per[-1, 2]) < 1e-12 * mv2
