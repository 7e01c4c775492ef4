"""The NVT tail of the step on the GPU (homogeneous long-range correction,
global velocity-scaling thermostat with the explosion heuristic) against the
oracle running the reference step order (ox_simulate_nvt), FP64 form:
positions, velocities, angular momenta, forces and virials within 1e-11 of
the largest reference magnitude after 20 steps, per-step Upot / virial / sums
and the last step's T and betas within 1e-11 relative; batched replicas with
different targets bit-identical to single runs."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP64, Context
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run_gpu(s, steps, T_target, lrc, dt=0.002):
    with Context(0) as c:
        c.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, True)
        c.set_thermostat(T_target, lrc=lrc)
        c.upload(s.r, s.v, s.q, s.L, s.cid)
        c.update_container()
        c.traverse()
        per = c.step(steps, dt, use_graph=True)
        return c.download(), per, c.thermostat_state()[0]


def check(s, steps, T_target, lrc, expect_explosion=False):
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    ref = o.simulate_nvt(steps, 0.002, s.r, s.v, s.q, s.L, s.cid, T_target, lrc=lrc)
    out, per, th = run_gpu(s, steps, T_target, lrc)
    rp = ref["per_step"]
    for k in range(4):
        assert rel_err(per[:, k], rp[:, k]) < TOL, k
    assert abs(th[4] - rp[-1, 6]) < TOL * rp[-1, 6]
    assert abs(th[5] - rp[-1, 7]) < TOL and abs(th[6] - rp[-1, 8]) < TOL
    if expect_explosion is not None:
        assert bool(th[7]) == expect_explosion
    if lrc:
        np.testing.assert_allclose(th[1:4], ref["lrc"], rtol=1e-14)
    d = out["r"] - ref["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < TOL * np.abs(s.box).max()
    for key in ("v", "L", "F", "M", "Vi"):
        if np.abs(ref[key]).max() > 0:
            assert rel_err(out[key], ref[key]) < TOL, key


def test_nvt_lj_lrc_vs_oracle():
    """One-centre LJ (the pair kernel writes molecule virials): LRC on, the
    thermostat cools a lattice start towards its target."""
    say("case nvt lj")
    s = scenarios.lj_fcc(ncell=5, T=1.0, seed=21)
    check(s, 20, 0.9, lrc=True)


def test_nvt_explosion_heuristic_vs_oracle():
    """A target far below the current temperature fires the explosion
    heuristic every step (capping, beta_trans 1)."""
    say("case nvt explosion")
    s = scenarios.lj_fcc(ncell=5, T=1.5, seed=22)
    check(s, 20, 0.3, lrc=False, expect_explosion=True)


def test_nvt_multisite_lrc_vs_oracle():
    """2CLJQ, water and a dipolar component: site-site LRC terms, the
    reaction-field self term, rotational scaling."""
    say("case nvt mixture")
    s = scenarios.small_mixture(n=600, rho=0.25, seed=23)
    check(s, 20, 0.9, lrc=True, expect_explosion=None)


def test_nvt_replicas_equal_single_runs():
    say("case nvt replicas")
    reps = [scenarios.ljts_box(n, rho, T, seed=200 + k)
            for k, (n, rho, T) in enumerate([(1500, 0.6, 1.2), (800, 0.3, 0.9), (2500, 0.8, 1.0)])]
    targets = [1.0, 0.0, 0.85]  # the second replica runs without thermostat
    steps = 10
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP64)
        c.set_thermostat(targets, lrc=True)
        c.update_container()
        c.traverse()
        per = c.step(steps, reps[0].dt, use_graph=True)
        out = c.download()
        th = c.thermostat_state()
    off = np.cumsum([0] + [s.n for s in reps])
    for k, s in enumerate(reps):
        o1, p1, t1 = run_gpu(s, steps, targets[k], True, dt=s.dt)
        sl = slice(off[k], off[k + 1])
        for key in ("r", "v", "F", "Vi"):
            assert np.array_equal(out[key][sl], o1[key]), (k, key)
        np.testing.assert_allclose(per[:, k, :4], p1[:, :4], rtol=1e-11,
                                   atol=1e-11 * np.abs(p1[:, :4]).max())
        # a replica without target has no thermostat record in its single run
        n_cmp = 8 if targets[k] > 0 else 4
        np.testing.assert_allclose(th[k][:n_cmp], t1[:n_cmp], rtol=1e-11)
    assert th[1][5] == 1.0 and th[1][6] == 1.0
