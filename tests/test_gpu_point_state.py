"""Orientation state of one-centre LJ systems on the GPU. Their q and L never
enter the dynamics (no site offsets, invI = 0, no torque), so the device keeps
them in upload order instead of permuting them with every re-bin, and applies
ls1's per-update normalisation of q (FullMolecule::setupSoACache,
FullMolecule.cpp:719) when q is read. Checked here: after a mix of container
updates, graph-captured and eager steps (with and without the thermostat) the
downloaded q equals the upload's q normalised once per container update (the
reference normalises at every update), L is bit-identical to the upload, and
positions and velocities equal those of a run from unit quaternions."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def normalise_k(q, k):
    """Quaternion::normalize applied k times, as the reference's k updates do."""
    q = q.copy()
    for _ in range(k):
        qs = 1.0 / np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2]
                           + q[:, 3] * q[:, 3])
        q *= qs[:, None]
    return q


def run(s, q, L, precision, thermostat):
    with Context(0) as c:
        c.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, True, 0, precision)
        if thermostat:
            c.set_thermostat(0.9)
        c.upload(s.r, s.v, q, L, s.cid)
        c.update_container()            # 1 update
        c.traverse()
        c.step(5, 0.002, use_graph=True)   # 6
        c.step(3, 0.002, use_graph=False)  # 9
        mid = c.download()
        c.step(40, 0.002, use_graph=True)  # 49
        return mid, c.download()


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_FP64ACC])
@pytest.mark.parametrize("thermostat", [False, True])
def test_point_lj_orientation_state(precision, thermostat):
    say("point orientation", precision, thermostat)
    s = scenarios.small_lj(n=500, seed=71)
    rng = np.random.default_rng(72)
    q = rng.normal(size=(s.n, 4)) * rng.uniform(0.5, 2.0, size=(s.n, 1))
    L = rng.normal(size=(s.n, 3))
    mid, end = run(s, q, L, precision, thermostat)
    np.testing.assert_allclose(mid["q"], normalise_k(q, 9), rtol=0, atol=1e-15)
    np.testing.assert_allclose(end["q"], normalise_k(q, 49), rtol=0, atol=1e-15)
    assert np.array_equal(mid["L"], L) and np.array_equal(end["L"], L)
    # the orientation has no effect on the trajectory
    unit_q = np.tile([1.0, 0.0, 0.0, 0.0], (s.n, 1))
    mid1, end1 = run(s, unit_q, np.zeros((s.n, 3)), precision, thermostat)
    for key in ("r", "v", "F", "Vi"):
        assert np.array_equal(end[key], end1[key]), key
    assert np.array_equal(end1["q"], unit_q) and not end1["L"].any()
