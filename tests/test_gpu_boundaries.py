"""Per-dimension boundaries on the GPU (mdx_set_boundaries): forces with no
images across reflecting dimensions, and trajectories with reflecting and
outflow walls, against the oracle (FP64 form, 1e-12 / 1e-11; outflow: the
same survivors, in upload order); replicas bit-identical to single runs;
unsupported combinations rejected."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import (BOUNDARY_OUTFLOW, BOUNDARY_PERIODIC, BOUNDARY_REFLECTING, PREC_FP64, Context,
                    MdxError)
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
P, R, O = BOUNDARY_PERIODIC, BOUNDARY_REFLECTING, BOUNDARY_OUTFLOW


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
        c.set_boundaries((P, P, R))
        with pytest.raises(MdxError):
            c.set_rmm(0.002)
        # outflow walls change the molecule count: no thermostat, long-range
        # correction, reduced-memory mode or adjoint with them
        c.set_boundaries((O, P, P))
        with pytest.raises(MdxError):
            c.set_thermostat(1.0)
        with pytest.raises(MdxError):
            c.set_thermostat(None, lrc=True)
        with pytest.raises(MdxError):
            c.set_rmm(0.002)
        c.update_container()
        c.traverse()
        with pytest.raises(MdxError):
            c.adjoint(None)
    with Context(0) as c:
        c.setup_system(s)
        c.set_thermostat(1.0)
        with pytest.raises(MdxError):
            c.set_boundaries((P, O, P))
    reps = [scenarios.ljts_box(500, 0.5, 1.0, seed=54 + k) for k in range(2)]
    with Context(0) as c:
        c.setup_replicas(reps)
        with pytest.raises(MdxError):
            c.set_boundaries((P, P, O))


def isolated_streams():
    """Five molecules with no partner within rc moving towards the x walls;
    every position is exact (drifts of whole quarters)."""
    box = np.array([12.0, 12.0, 12.0])
    x = np.array([11.5, 11.0, 9.75, 6.0, 0.5])
    vx = np.array([128.0, 32.0, 64.0, 0.0, -32.0])
    n = len(x)
    r = np.stack([x, 3.0 * np.arange(n) % 12.0 + 0.5, (np.arange(n) * 2.9) % 12.0], axis=1)
    v = np.stack([vx, np.zeros(n), np.zeros(n)], axis=1)
    desc = np.array(scenarios.comp_1clj(), np.float64)
    return scenarios.System("streams", 1, desc, np.zeros((0, 2)), 1.0, 2.5, 2.5, box, r, v,
                            np.tile([1.0, 0, 0, 0], (n, 1)), np.zeros((n, 3)),
                            np.zeros(n, np.int32), 1.0, dt=0.0078125)


@pytest.mark.parametrize("use_graph", [False, True])
def test_outflow_deletes_leavers(use_graph):
    """The hand-built case of the oracle's outflow pin: molecule 0 is deleted in
    step 1, 4 in step 2, 1 in step 4, 2 in step 5; molecule 3 stays."""
    s = isolated_streams()
    say("case outflow hand-built graph", use_graph)
    counts = []
    with Context(0) as c:
        c.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, True)
        c.set_boundaries((O, P, P))
        c.upload(s.r, s.v, s.q, s.L, s.cid)
        c.update_container()
        c.traverse()
        for _ in range(10):
            c.step(1, s.dt, use_graph=use_graph)
            counts.append(c.molecule_count())
        out = c.download()
        cells = c.cell_ids()
    assert counts == [4, 3, 3, 2, 1, 1, 1, 1, 1, 1]
    np.testing.assert_array_equal(out["ids"], [4])  # upload ids are 1-based
    np.testing.assert_array_equal(out["r"], s.r[3:4])
    assert len(cells) == 1


def streaming_fluid(seed):
    """LJ fluid flowing towards +x (drift 3 on top of the thermal velocities)."""
    s = scenarios.lj_fcc(ncell=5, T=1.0, seed=seed)
    s.v[:, 0] += 3.0
    return s


@pytest.mark.parametrize("types", [(O, P, P), (O, R, O)])
def test_outflow_trajectory_vs_oracle(types):
    s = streaming_fluid(71)
    steps = 60
    say("case outflow trajectory", types, "N", s.n)
    o = oracle_for(s, types)
    ref = o.simulate_open(steps, 0.002, s.r, s.v, s.q, s.L, s.cid)
    assert len(ref["orig"]) < s.n  # molecules did leave
    with Context(0) as c:
        c.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, True)
        c.set_boundaries(types)
        c.upload(s.r, s.v, s.q, s.L, s.cid)
        c.update_container()
        c.traverse()
        per = c.step(steps, 0.002, use_graph=True)
        out = c.download()
    np.testing.assert_array_equal(out["ids"], ref["orig"] + 1)
    d = out["r"] - ref["r"]
    for k in range(3):
        if types[k] == P:
            d[:, k] -= np.round(d[:, k] / s.box[k]) * s.box[k]
    assert np.abs(d).max() < 1e-11 * s.box.max()
    assert np.abs(out["v"] - ref["v"]).max() < 1e-11 * np.abs(ref["v"]).max()
    assert np.abs(out["F"] - ref["F"]).max() < 1e-11 * np.abs(ref["F"]).max()
    rp = ref["per_step"]
    for k in range(3):
        assert np.abs(per[:, k] - rp[:, k]).max() < 1e-11 * np.abs(rp[:, k]).max(), k
    # the kinetic sum covers exactly the survivors of each step
    assert np.all(np.diff(rp[:, 6]) <= 0)
