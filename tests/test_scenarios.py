"""The seeded workloads match their stated definitions (geometry, masses,
inertia, densities, composition, state-point ranges). CPU only."""
import math

import numpy as np
import pytest

import scenarios


def test_water_geometry_and_frame():
    (c,) = scenarios.parse_components(scenarios.comp_water(3.0))
    o = np.array(c["lj"][0][:3])
    h1, h2, msite = (np.array(s[:3]) for s in c["ch"])
    assert math.isclose(np.linalg.norm(h1 - o), 0.3, rel_tol=1e-12)
    assert math.isclose(np.linalg.norm(h2 - o), 0.3, rel_tol=1e-12)
    cosang = np.dot(h1 - o, h2 - o) / 0.09
    assert math.isclose(math.degrees(math.acos(cosang)), 104.5, rel_tol=1e-12)
    # M site on the bisector, 0.05 from the oxygen, massless, charge -1.04
    assert math.isclose(np.linalg.norm(msite - o), 0.05, rel_tol=1e-12)
    assert msite[0] == 0 and msite[1] == 0 and c["ch"][2][3] == 0.0
    assert [s[4] for s in c["ch"]] == [0.52, 0.52, -1.04]
    # masses 16:1:1 normalised to 1, centre of mass at the origin
    assert math.isclose(c["mass"], 1.0, rel_tol=1e-14)
    com = sum(np.array(s[:3]) * s[3] for s in c["lj"] + c["ch"])
    assert np.abs(com).max() < 1e-15


def test_2cljq_inertia_and_quadrupole():
    (c,) = scenarios.parse_components(scenarios.comp_2cljq(3.0))
    assert [s[2] for s in c["lj"]] == [0.25, -0.25]
    np.testing.assert_allclose(c["I"], [0.0625, 0.0625, 0.0], atol=1e-15)
    assert c["qp"][0][6] == 2.0 and c["qp"][0][5] == 1.0  # Q* = 2 along the axis


def test_mixture_composition_small():
    s = scenarios.mixture(n=4000)
    assert s.n == 4000 and (s.cid == 1).sum() == 2000
    assert math.isclose(s.n / np.prod(s.box), 0.5, rel_tol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(s.q, axis=1), 1.0, rtol=1e-14)
    assert s.rc == 3.0 and s.rc_lj == 3.0


def test_binary_lj_definition_small():
    s = scenarios.binary_lj(n=5000)
    comps = scenarios.parse_components(s.comp_desc)
    assert comps[0]["lj"][0][4:6] == [1.0, 1.0] and comps[1]["lj"][0][4:6] == [0.85, 1.1]
    assert (s.cid == 1).sum() == 2000  # 60/40
    assert math.isclose(s.n / np.prod(s.box), 0.7, rel_tol=1e-12)
    xi, eta = s.mix[0]
    assert 0.95 <= xi <= 1.05 and 0.95 <= eta <= 1.05


def test_replica_sweep_ranges():
    reps = scenarios.replicas(nrep=6, nmin=2000, nmax=4000)
    assert len(reps) == 6
    for s in reps:
        assert 2000 <= s.n <= 4000
        rho = s.n / np.prod(s.box)
        assert 0.05 <= rho <= 0.85
        assert 0.7 <= s.temperature <= 1.3
        assert np.all(np.floor(s.box / s.rc) >= 2)  # at least 2 cells per dimension
        comps = scenarios.parse_components(s.comp_desc)
        assert comps[0]["lj"][0][7] == 1.0  # truncated-shifted


@pytest.mark.parametrize("scale", [0.25])
def test_ljts_slab_densities(scale):
    s = scenarios.ljts_slab(scale=scale)
    w = 200.0 * scale
    x0 = 0.5 * (s.box[0] - w)
    inside = (s.r[:, 0] >= x0) & (s.r[:, 0] < x0 + w)
    area = s.box[1] * s.box[2]
    assert math.isclose(inside.sum() / (w * area), 0.73, rel_tol=1e-3)
    assert math.isclose((~inside).sum() / ((s.box[0] - w) * area), 0.02, rel_tol=2e-2)
