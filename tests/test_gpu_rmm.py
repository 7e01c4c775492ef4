"""Reduced-memory one-centre mode on the GPU (mdx_set_rmm: VCP1CLJRMM +
LeapfrogRMM, the kick inside the pair kernel) against the oracle's
restatement (FP64 form: positions, half-step velocities and the per-step
records within 1e-11), against the GPU's own standard Leapfrog (same
positions), the FP32 form's error, replicas, and its argument checks."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context, MdxError
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def wrap_diff(a, b, box):
    d = a - b
    return d - np.round(d / box) * box


def run(s, steps, dt, rmm, precision=PREC_FP64):
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=precision)
        if rmm:
            c.set_rmm(dt)
        else:
            c.update_container()
            c.traverse()
        per = c.step(steps, dt, use_graph=True)
        return c.download(), per


@pytest.mark.parametrize("name", ["ljts_box", "lj_fcc"])
def test_rmm_vs_oracle(name):
    say("case rmm", name)
    s = (scenarios.ljts_box(3000, 0.7, 0.9, seed=31) if name == "ljts_box"
         else scenarios.lj_fcc(ncell=5, seed=32))
    steps, dt = 50, 0.002
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box)
    ref = o.simulate_rmm(steps, dt, s.r, s.v, s.q, s.cid)
    out, per = run(s, steps, dt, True)
    assert np.abs(wrap_diff(out["r"], ref["r"], s.box)).max() < TOL * s.box.max()
    assert np.abs(out["v"] - ref["v"]).max() < TOL * np.abs(ref["v"]).max()
    rp = ref["per_step"]
    for k in range(3):
        assert np.abs(per[:, k] - rp[:, k]).max() < TOL * np.abs(rp[:, k]).max(), k
    assert np.all(per[:, 3] == 0.0)


def test_rmm_positions_equal_standard_leapfrog():
    say("case rmm vs leapfrog")
    s = scenarios.ljts_box(2000, 0.6, 1.0, seed=33)
    steps, dt = 50, 0.002
    a, pa = run(s, steps, dt, True)
    b, pb = run(s, steps, dt, False)
    assert np.abs(wrap_diff(a["r"], b["r"], s.box)).max() < TOL * s.box.max()
    np.testing.assert_allclose(pa[:, :2], pb[:, :2], rtol=1e-10)


def test_rmm_fp32_form():
    say("case rmm fp32")
    s = scenarios.ljts_box(3000, 0.7, 0.9, seed=34)
    a, _ = run(s, 20, 0.002, True, PREC_FP32_FP64ACC)
    b, _ = run(s, 20, 0.002, True, PREC_FP64)
    assert np.abs(wrap_diff(a["r"], b["r"], s.box)).max() < 1e-6
    assert np.sqrt(((a["v"] - b["v"]) ** 2).mean()) < 1e-5 * np.sqrt((b["v"] ** 2).mean())


def test_rmm_replicas_equal_single_runs():
    say("case rmm replicas")
    reps = [scenarios.ljts_box(n, rho, T, seed=300 + k)
            for k, (n, rho, T) in enumerate([(1200, 0.5, 1.1), (2500, 0.8, 0.9)])]
    steps, dt = 10, 0.002
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP64)
        c.set_rmm(dt)
        per = c.step(steps, dt, use_graph=True)
        out = c.download()
    off = np.cumsum([0] + [s.n for s in reps])
    for k, s in enumerate(reps):
        o1, p1 = run(s, steps, dt, True)
        sl = slice(off[k], off[k + 1])
        for key in ("r", "v"):
            assert np.array_equal(out[key][sl], o1[key]), (k, key)
        np.testing.assert_allclose(per[:, k, :3], p1[:, :3], rtol=1e-11)


def test_rmm_argument_checks():
    s = scenarios.small_mixture(n=100)
    with Context(0) as c:
        c.setup_system(s)
        with pytest.raises(MdxError):
            c.set_rmm(0.001)
    s = scenarios.ljts_box(500, 0.5, 1.0, seed=35)
    with Context(0) as c:
        c.setup_system(s)
        c.set_rmm(0.002)
        with pytest.raises(MdxError):
            c.step(1, 0.001)
        c.step(2, 0.002)
        c.set_rmm(0.0)
        c.step(1, 0.001)
