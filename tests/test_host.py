"""The C++ host layer (HIPLinkedCells / HIPCellProcessor / HIPLeapfrog over
the C ABI) and its driver mdx_run."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

import scenarios
from ls1inp import read_inp, write_inp
from oracle_lib import Oracle

ROOT = pathlib.Path(__file__).resolve().parents[1]
MDX_RUN = ROOT / "ls1-mardyn_amd" / "mdx_run"
GOLD = ROOT / "tests" / "golden"


def test_driver_usage():
    r = subprocess.run([str(MDX_RUN), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "usage" in r.stderr


@pytest.mark.parametrize("name", ["VectorizationMultiComponentMultiPotentials_50_molecules",
                                  "VectorizationWater", "ForceCalculationTestU0"])
def test_phase_space_round_trip(name, tmp_path):
    src = GOLD / "inp" / f"{name}.inp"
    out = tmp_path / "rt.inp"
    subprocess.run([str(MDX_RUN), str(src), "--read-write-only", str(out)], check=True)
    a, b = read_inp(src), read_inp(out)
    for k in ("box", "comp_desc", "mix", "ids", "cid", "r", "v", "q", "L"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert a.eps_rf == b.eps_rf and a.ncomp == b.ncomp


def pressure(per_step, n, volume, ndof):
    """Domain::getGlobalPressure (Domain.cpp:142-146): p = rho T + rho W / (3 N),
    with T = (sum m v^2 + sum I w^2) / (3 N + rotDOF) (Domain.cpp:240-242)
    and the global virial W (local + long-range correction)."""
    rho = n / volume
    T = (per_step[:, 2] + per_step[:, 3]) / ndof
    return rho * T + rho * (per_step[:, 1] / n) / 3.0


def parsed(stdout, key):
    return np.array([float(m) for m in re.findall(r"\b" + key + r" = ([-0-9.eE+]+)", stdout)])


@pytest.mark.gpu
def test_driver_trajectory_vs_oracle(tmp_path):
    s = scenarios.small_mixture(seed=71)
    f = tmp_path / "mix.inp"
    write_inp(f, s.box, s.ncomp, s.comp_desc, s.mix, s.eps_rf, np.arange(1, s.n + 1), s.cid, s.r,
              s.v, s.q, s.L)
    r = subprocess.run([str(MDX_RUN), str(f), "--cutoff", str(s.rc), "--dt", "0.001", "--steps",
                        "20", "--checkpoint", str(tmp_path / "out.inp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    upot = [float(m) for m in re.findall(r"U_pot = ([-0-9.eE+]+)", r.stdout)]
    assert len(upot) == 21
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc, s.box, periodic=True)
    ref = o.simulate(20, 0.001, s.r, s.v, s.q, s.L, s.cid)
    ref_u = np.concatenate([[sum(ref["initial"]["macro"][[0]]) / 6 + ref["initial"]["macro"][1]
                             + ref["initial"]["macro"][3]], ref["per_step"][:, 0]]) / s.n
    np.testing.assert_allclose(upot, ref_u, rtol=1e-9)
    # the pressure of every step (NVE: no correction terms)
    _, _, rot = o.component_props()
    ndof = 3 * s.n + int(rot[s.cid].sum())
    p_ref = pressure(ref["per_step"], s.n, float(np.prod(s.box)), ndof)
    np.testing.assert_allclose(parsed(r.stdout, "p"), p_ref, rtol=1e-9)
    fin = read_inp(tmp_path / "out.inp")
    d = fin.r - ref["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < 1e-9


@pytest.mark.gpu
def test_driver_rebuild_keeps_precision(tmp_path):
    """mdx_run rebuilds its container (as Simulation::readXML does) after
    setting the FP32 form: the run stays in the FP32 form (its U_pot differs
    from the FP64 run's, within the form's error bound) and on the device it
    was given; the FP64 run matches the oracle."""
    s = scenarios.small_lj(seed=74)
    f = tmp_path / "lj.inp"
    write_inp(f, s.box, s.ncomp, s.comp_desc, s.mix, s.eps_rf, np.arange(1, s.n + 1), s.cid, s.r,
              s.v, s.q, s.L)
    out = {}
    for prec in ("fp64", "fp32"):
        r = subprocess.run([str(MDX_RUN), str(f), "--cutoff", str(s.rc), "--dt", "0.002",
                            "--steps", "5", "--precision", prec, "--device", "0"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out[prec] = parsed(r.stdout, "U_pot")
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box, periodic=True, cell_mode=1)
    ref = o.simulate(5, 0.002, s.r, s.v, s.q, s.L, s.cid)
    np.testing.assert_allclose(out["fp64"][1:], ref["per_step"][:, 0] / s.n, rtol=1e-9)
    rel = np.abs(out["fp32"] - out["fp64"]) / np.abs(out["fp64"])
    assert rel.max() < 1e-5 and rel.max() > 1e-12


@pytest.mark.gpu
def test_driver_nvt_lrc_vs_oracle(tmp_path):
    """mdx_run --nvt T --lrc: the host layer's step loop with the device's NVT
    tail against the oracle's NVT step order (global Upot per molecule incl.
    the correction, final positions)."""
    s = scenarios.lj_fcc(ncell=4, T=1.0, seed=72)
    f = tmp_path / "lj.inp"
    write_inp(f, s.box, s.ncomp, s.comp_desc, s.mix, s.eps_rf, np.arange(1, s.n + 1), s.cid, s.r,
              s.v, s.q, s.L)
    r = subprocess.run([str(MDX_RUN), str(f), "--cutoff", str(s.rc), "--dt", "0.002", "--steps",
                        "20", "--nvt", "0.9", "--lrc", "--checkpoint", str(tmp_path / "out.inp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    upot = [float(m) for m in re.findall(r"U_pot = ([-0-9.eE+]+)", r.stdout)][1:]
    temp = [float(m) for m in re.findall(r"T = ([-0-9.eE+]+)", r.stdout)]
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box, periodic=True)
    ref = o.simulate_nvt(20, 0.002, s.r, s.v, s.q, s.L, s.cid, 0.9, lrc=True)
    np.testing.assert_allclose(upot, ref["per_step"][:, 0] / s.n, rtol=1e-9)
    np.testing.assert_allclose(temp, ref["per_step"][:, 6], rtol=1e-9)
    # p from the step's T (before the velocity scaling) and the global virial
    # including the long-range correction
    rho = s.n / float(np.prod(s.box))
    p_ref = rho * ref["per_step"][:, 6] + rho * (ref["per_step"][:, 1] / s.n) / 3.0
    np.testing.assert_allclose(parsed(r.stdout, "p"), p_ref, rtol=1e-9)
    fin = read_inp(tmp_path / "out.inp")
    d = fin.r - ref["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < 1e-9


@pytest.mark.gpu
def test_driver_rmm_vs_oracle(tmp_path):
    """mdx_run --rmm: the reduced-memory processor through the host layer."""
    s = scenarios.lj_fcc(ncell=4, T=0.8, seed=73, shifted=True)
    f = tmp_path / "lj.inp"
    write_inp(f, s.box, s.ncomp, s.comp_desc, s.mix, s.eps_rf, np.arange(1, s.n + 1), s.cid, s.r,
              s.v, s.q, s.L)
    r = subprocess.run([str(MDX_RUN), str(f), "--cutoff", str(s.rc), "--dt", "0.002", "--steps",
                        "20", "--rmm", "--checkpoint", str(tmp_path / "out.inp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box, periodic=True)
    ref = o.simulate_rmm(20, 0.002, s.r, s.v, s.q, s.cid)
    fin = read_inp(tmp_path / "out.inp")
    d = fin.r - ref["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < 1e-9
    np.testing.assert_allclose(fin.v, ref["v"], rtol=0, atol=1e-9)
