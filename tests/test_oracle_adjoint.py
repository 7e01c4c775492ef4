"""Pins the oracle's FP64 adjoint (oracle/ls1_oracle.c ox_adjoint) by central
finite differences of the oracle's own Upot, virial and forces:
* dU/d(eps_ab, sigma_ab), dW/d(...) through rebuilt LJ tables (the shift term
  of shifted pairs moves with its parameters, ls1's Comp2Param rules);
* dU/dq, dW/dq through rebuilt component records;
* the force vector-Jacobian product through a directional derivative of
  c . F(r + h delta).
"""
import numpy as np
import pytest

import scenarios
from oracle_lib import Oracle, macro_to_upot_virial

H = 1e-6


def tables_from_theta(eps, sig, shift_rc, base_shift6):
    eps24 = 24.0 * eps
    sig2 = sig * sig
    shift6 = base_shift6.copy()
    m = shift_rc > 0
    s6 = (sig2[m] / shift_rc[m] ** 2) ** 3
    shift6[m] = eps24[m] * (s6 - s6 * s6)
    return eps24, sig2, shift6


def uw(o, s):
    res = o.forces(s.r, s.q, s.cid)
    return macro_to_upot_virial(res["macro"]), res["F"]


def small_binary(seed=81, n=300):
    rho = 0.7
    box = np.full(3, (n / rho) ** (1 / 3))
    r = scenarios.rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.85)
    desc = np.array(scenarios.comp_1clj(1.0, 1.0, 1.0, 2.5) + scenarios.comp_1clj(0.85, 1.1, 1.0, 2.5))
    rng = np.random.default_rng(seed)
    cid = (rng.random(n) < 0.4).astype(np.int32)
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    return scenarios.System("binary", 2, desc, np.array([[1.02, 0.97]]), 1.0, 2.5, 2.5, box, r,
                            np.zeros((n, 3)), q, np.zeros((n, 3)), cid, 0.9)


def small_charged(seed=82, n=120):
    rho = 0.3
    box = np.full(3, (n / rho) ** (1 / 3))
    r = scenarios.rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.95)
    desc = np.array(scenarios.comp_water(2.5) + scenarios.comp_1clj(0.9, 1.05, 1.0, 2.5, True))
    rng = np.random.default_rng(seed)
    cid = (rng.random(n) < 0.3).astype(np.int32)
    q = scenarios.random_quaternions(rng, n)
    return scenarios.System("charged", 2, desc, np.array([[0.98, 1.03]]), 1.0, 2.5, 2.5, box, r,
                            np.zeros((n, 3)), q, np.zeros((n, 3)), cid, 1.0)


@pytest.mark.parametrize("make", [small_binary, small_charged])
def test_lj_parameter_gradients_fd(make):
    s = make()
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    e24, s2, sh6 = o.lj_tables()
    eps, sig, src = e24 / 24.0, np.sqrt(s2), o.lj_shift_rc()
    # the parameterisation reproduces ls1's own tables
    t = tables_from_theta(eps, sig, src, sh6)
    np.testing.assert_allclose(t[2], sh6, rtol=1e-13, atol=1e-15)
    adj = o.adjoint(s.r, s.q, s.cid, np.zeros((s.n, 3)))
    nl = len(eps)
    for a in range(nl):
        for b in range(a, nl):
            for which, key_u, key_w in ((0, "dU_deps", "dW_deps"), (1, "dU_dsig", "dW_dsig")):
                vals = []
                for sgn in (1, -1):
                    e, g = eps.copy(), sig.copy()
                    tgt = e if which == 0 else g
                    tgt[a, b] += sgn * H
                    if a != b:
                        tgt[b, a] += sgn * H
                    o.set_lj_tables(*tables_from_theta(e, g, src, sh6))
                    vals.append(uw(o, s)[0])
                o.set_lj_tables(e24, s2, sh6)
                fd_u = (vals[0][0] - vals[1][0]) / (2 * H)
                fd_w = (vals[0][1] - vals[1][1]) / (2 * H)
                for got, fd in ((adj[key_u][a, b], fd_u), (adj[key_w][a, b], fd_w)):
                    assert abs(got - fd) <= 1e-6 * max(abs(fd), 1e-3), (a, b, which, got, fd)
                assert adj[key_u][a, b] == adj[key_u][b, a]


def test_charge_gradients_fd():
    s = small_charged()
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    adj = o.adjoint(s.r, s.q, s.cid, np.zeros((s.n, 3)))
    comps = scenarios.parse_components(s.comp_desc)
    # offsets of the charge records in the flat description (water: component 0)
    p = 4 + 8 * comps[0]["nlj"]
    for site in range(comps[0]["nc"]):
        idx = p + 5 * site + 4
        vals = []
        for sgn in (1, -1):
            d = s.comp_desc.copy()
            d[idx] += sgn * H
            oo = Oracle(s.ncomp, d, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
            vals.append(uw(oo, s)[0])
        fd_u = (vals[0][0] - vals[1][0]) / (2 * H)
        fd_w = (vals[0][1] - vals[1][1]) / (2 * H)
        assert abs(adj["dU_dq"][0, site] - fd_u) <= 1e-6 * max(abs(fd_u), 1e-3)
        assert abs(adj["dW_dq"][0, site] - fd_w) <= 1e-6 * max(abs(fd_w), 1e-3)


@pytest.mark.parametrize("make", [small_binary, small_charged])
def test_force_vjp_fd(make):
    s = make()
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    rng = np.random.default_rng(5)
    cot = rng.normal(size=(s.n, 3))
    delta = rng.normal(size=(s.n, 3))
    adj = o.adjoint(s.r, s.q, s.cid, cot)
    r0 = s.r.copy()
    vals = []
    for sgn in (1, -1):
        s.r = r0 + sgn * H * delta
        vals.append((cot * uw(o, s)[1]).sum())
    s.r = r0
    fd = (vals[0] - vals[1]) / (2 * H)
    got = (adj["vjp"] * delta).sum()
    assert abs(got - fd) <= 1e-6 * abs(fd)
