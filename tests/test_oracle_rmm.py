"""The oracle's reduced-memory mode (VCP1CLJRMM + LeapfrogRMM) against the
standard Leapfrog it rearranges: in exact arithmetic the two produce the same
positions, and the half-step velocities are the full-step ones plus
(dt/2m) F; in FP64 they agree to rounding. CPU only."""
import pathlib
import sys

import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "ls1-mardyn_amd"))

import scenarios  # noqa: E402
from oracle_lib import Oracle  # noqa: E402


def test_rmm_equals_leapfrog():
    s = scenarios.small_lj(n=300, seed=8, shifted=True)
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box)
    steps, dt = 50, 0.002
    std = o.simulate(steps, dt, s.r, s.v, s.q, s.L, s.cid)
    rmm = o.simulate_rmm(steps, dt, s.r, s.v, s.q, s.cid)
    d = rmm["r"] - std["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < 1e-11
    m = o.component_props()[0][0]
    v_half = std["v"] + 0.5 * dt / m * std["F"]
    assert np.abs(rmm["v"] - v_half).max() < 1e-11 * np.abs(v_half).max()
    # same forces each step: Upot and virial records agree
    np.testing.assert_allclose(rmm["per_step"][:, :2], std["per_step"][:, :2], rtol=1e-11)
    # the record's sum is over the half-step velocities, without mass
    assert abs(rmm["per_step"][-1, 2] - (rmm["v"] ** 2).sum()) < 1e-12 * rmm["per_step"][-1, 2]


def test_rmm_rejects_multisite():
    s = scenarios.small_mixture(n=50)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    try:
        o.simulate_rmm(2, 0.001, s.r, s.v, s.q, s.cid)
    except RuntimeError:
        return
    raise AssertionError("multi-site system accepted")
