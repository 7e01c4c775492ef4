"""Per-dimension boundaries on the GPU (mdx_set_boundaries): forces with no
images across reflecting dimensions, and trajectories with reflecting walls,
against the oracle (FP64 form, 1e-12 / 1e-11); replicas bit-identical to
single runs; unsupported combinations rejected."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import (BOUNDARY_OUTFLOW, BOUNDARY_PERIODIC, BOUNDARY_REFLECTING, PREC_FP64, Context,
                    MdxError)
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
P, R = BOUNDARY_PERIODIC, BOUNDARY_REFLECTING


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_for(s, types):
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    o.set_boundaries(types)
    return o


def gpu_run(s, types, steps=0, dt=0.002):
    with Context(0) as c:
        c.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, True)
        c.set_boundaries(types)
        c.upload(s.r, s.v, s.q, s.L, s.cid)
        c.update_container()
        mac = c.traverse()
        f0 = c.download()
        per = c.step(steps, dt, use_graph=True) if steps else None
        return f0, mac, c.download(), per


@pytest.mark.parametrize("types", [(P, P, R), (R, P, R), (R, R, R)])
def test_forces_vs_oracle(types):
    say("case boundaries forces", types)
    s = scenarios.small_mixture(n=400, rho=0.3, seed=51)
    ref = oracle_for(s, types).forces(s.r, s.q, s.cid)
    f0, mac, _, _ = gpu_run(s, types)
    for k in ("F", "M", "Vi"):
        assert np.abs(f0[k] - ref[k]).max() < 1e-12 * np.abs(ref[k]).max(), k
    np.testing.assert_allclose(mac, ref["macro"], rtol=1e-12, atol=1e-12 * np.abs(ref["macro"]).max())


def hot_walls_system(seed):
    """LJ fluid whose molecules near the z walls move towards them."""
    s = scenarios.lj_fcc(ncell=5, T=1.2, seed=seed)
    near = s.r[:, 2] > s.box[2] - s.rc
    s.v[near, 2] = np.abs(s.v[near, 2]) + 0.5
    near = s.r[:, 2] < s.rc
    s.v[near, 2] = -np.abs(s.v[near, 2]) - 0.5
    return s


@pytest.mark.parametrize("types", [(P, P, R), (R, R, R)])
def test_reflecting_trajectory_vs_oracle(types):
    say("case boundaries trajectory", types)
    s = hot_walls_system(52)
    steps = 60
    ref = oracle_for(s, types).simulate(steps, 0.002, s.r, s.v, s.q, s.L, s.cid)
    _, _, out, per = gpu_run(s, types, steps)
    d = out["r"] - ref["r"]
    for k in range(3):
        if types[k] == P:
            d[:, k] -= np.round(d[:, k] / s.box[k]) * s.box[k]
    assert np.abs(d).max() < 1e-11 * s.box.max()
    assert np.abs(out["v"] - ref["v"]).max() < 1e-11 * np.abs(ref["v"]).max()
    rp = ref["per_step"]
    for k in range(3):
        assert np.abs(per[:, k] - rp[:, k]).max() < 1e-11 * np.abs(rp[:, k]).max(), k
    for k in range(3):
        if types[k] == R:
            assert np.all(out["r"][:, k] >= 0) and np.all(out["r"][:, k] < s.box[k])


def test_reflecting_replicas_equal_single_runs():
    say("case boundaries replicas")
    reps = [hot_walls_system(60 + k) for k in range(2)]
    types, steps = (P, P, R), 15
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP64)
        c.set_boundaries(types)
        c.update_container()
        c.traverse()
        c.step(steps, 0.002, use_graph=True)
        out = c.download()
    off = np.cumsum([0] + [s.n for s in reps])
    for k, s in enumerate(reps):
        _, _, o1, _ = gpu_run(s, types, steps)
        for key in ("r", "v", "F"):
            assert np.array_equal(out[key][off[k]:off[k + 1]], o1[key]), (k, key)


def test_boundary_argument_checks():
    s = scenarios.ljts_box(500, 0.5, 1.0, seed=53)
    with Context(0) as c:
        c.setup_system(s)
        with pytest.raises(MdxError):
            c.set_boundaries((BOUNDARY_OUTFLOW, P, P))
        c.set_boundaries((P, P, R))
        with pytest.raises(MdxError):
            c.set_rmm(0.002)
