"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's own outputs (tests/golden/ref_golden.npz).

Bars: FP64 form -- forces, torques and virials per molecule within 1e-12
relative to the largest reference magnitude of the quantity, Upot and virial
within 1e-12 relative; cell ids and the traversal's cell-pair set bit-exact.
FP32-pair / FP64-accumulate form -- relative RMS force error <= 1e-5 against
the FP64 oracle (the specified bound), pair set identical.
"""
import pathlib
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context, upot_virial
from ls1inp import read_inp
from oracle_lib import Oracle, macro_to_upot_virial

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
G = np.load(GOLD / "ref_golden.npz")
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rms_rel(a, b):
    return np.sqrt(((a - b) ** 2).sum(1).mean()) / np.sqrt((b ** 2).sum(1).mean())


def gpu_forces(s_or_ps, rc, rc_lj=None, periodic=True, precision=PREC_FP64, cell_mode=0):
    s = s_or_ps
    with Context(0) as ctx:
        ctx.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, rc, rc if rc_lj is None else rc_lj,
                  periodic, cell_mode, precision)
        ctx.upload(s.r, s.v, s.q, s.L, s.cid)
        ctx.set_counting(True)
        out = ctx.forces()
        out["counters"] = ctx.counters()
        out["cell_ids"] = ctx.cell_ids()
        return out


def check_against(out, F, M, Vi, upot, virial, tol=TOL):
    for got, ref, name in ((out["F"], F, "F"), (out["M"], M, "M"), (out["Vi"], Vi, "Vi")):
        if np.abs(ref).max() == 0:
            assert np.abs(got).max() == 0, name
        else:
            assert rel_err(got, ref) < tol, (name, rel_err(got, ref))
    u, w = upot_virial(out["macro"])
    assert abs(u - upot) <= tol * max(abs(upot), 1e-300), (u, upot)
    assert abs(w - virial) <= tol * max(abs(virial), 1e-300), (w, virial)


def test_kat_u0_f0():
    for name, rc, fexp, uexp, wexp in (
            ("ForceCalculationTestU0", 1.1,
             np.array([[-24, -24, 0], [24, -24, 0], [-24, 24, 0], [24, 24, 0]], float), 0.0, 96.0),
            ("ForceCalculationTestF0", 1.3, np.zeros((4, 3)), -4.0, 0.0)):
        say("KAT", name)
        ps = read_inp(GOLD / "inp" / f"{name}.inp")
        out = gpu_forces(ps, rc, periodic=False)
        np.testing.assert_allclose(out["F"], fexp, atol=1e-7)
        u, w = upot_virial(out["macro"])
        assert abs(u - uexp) < 1e-8 and abs(w - wexp) < 1e-6


FIXTURES = [
    ("VectorizationLennardJones", 35.0),
    ("VectorizationLennardJones1CLJ", 35.0),
    ("VectorizationCharge", 35.0),
    ("VectorizationChargeDipole", 35.0),
    ("VectorizationChargeQuadrupole", 35.0),
    ("VectorizationDipole", 35.0),
    ("VectorizationDipoleQuadrupole", 35.0),
    ("VectorizationQuadrupole", 35.0),
    ("VectorizationWater", 6.16),
    ("VectorizationMultiComponentMultiPotentials", 35.0),
    ("VectorizationMultiComponentMultiPotentials_50_molecules", 35.0),
]


@pytest.mark.parametrize("name,rc", FIXTURES)
def test_reference_fixture_fp64(name, rc):
    say("fixture", name, "rc", rc)
    ps = read_inp(GOLD / "inp" / f"{name}.inp")
    out = gpu_forces(ps, rc, periodic=False)
    check_against(out, G[f"{name}/F"], G[f"{name}/M"], G[f"{name}/Vi"], *G[f"{name}/upot_virial"])
    o = Oracle.from_phase_space(ps, rc, periodic=False)
    ref = o.forces(ps.r, ps.q, ps.cid)
    # molecule-centre tests are counted once per unordered site pair on the
    # GPU; in-cutoff site pairs must equal the oracle's exactly
    np.testing.assert_array_equal(out["counters"][1:], ref["counters"][1:])


@pytest.mark.parametrize("sysname,seed", [("small_lj", 41), ("small_mixture", 42)])
def test_periodic_fp64_vs_oracle(sysname, seed):
    say("periodic", sysname, seed)
    s = getattr(scenarios, sysname)(seed=seed)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    ref = o.forces(s.r, s.q, s.cid)
    out = gpu_forces(s, s.rc, s.rc_lj)
    u, w = macro_to_upot_virial(ref["macro"])
    check_against(out, ref["F"], ref["M"], ref["Vi"], u, w)
    np.testing.assert_array_equal(out["counters"][1:], ref["counters"][1:])
    np.testing.assert_allclose(out["q"], ref["q"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("tag", ["periodic_lj", "periodic_mix"])
def test_periodic_fp64_vs_reference_golden(tag):
    say("golden", tag)
    desc = G[f"{tag}/comp_desc"]
    rc, eps_rf = G[f"{tag}/rc_epsrf"]
    ncomp = len(scenarios.parse_components(desc))
    mix = np.array([[1.0, 1.0]] * (ncomp * (ncomp - 1) // 2))
    if ncomp == 3:
        mix[1] = (0.97, 1.02)
    n = len(G[f"{tag}/r"])
    with Context(0) as ctx:
        ctx.setup(ncomp, desc, mix, eps_rf, G[f"{tag}/box"], rc, rc, True)
        ctx.upload(G[f"{tag}/r"], np.zeros((n, 3)), G[f"{tag}/q"], np.zeros((n, 3)), G[f"{tag}/cid"])
        out = ctx.forces()
    check_against(out, G[f"{tag}/F"], G[f"{tag}/M"], G[f"{tag}/Vi"], *G[f"{tag}/upot_virial"])


def test_lj_fcc_32000_fp64_vs_oracle():
    """The 32,000-molecule FCC LJ fluid (20^3 unit cells, rho 0.6, rc 2.5)."""
    say("lj_fcc 32000")
    s = scenarios.lj_fcc()
    o = Oracle(1, s.comp_desc, None, 1.0, s.rc, s.rc, s.box, periodic=True)
    # on the perfect lattice every net force is cancellation noise (~1e-15):
    # compare on thermally displaced positions, as after the first steps
    rng = np.random.default_rng(3)
    s.r = np.mod(s.r + rng.normal(scale=0.05, size=s.r.shape), s.box)
    ref = o.forces(s.r, s.q, s.cid)
    out = gpu_forces(s, s.rc)
    u, w = macro_to_upot_virial(ref["macro"])
    check_against(out, ref["F"], ref["M"], ref["Vi"], u, w)
    np.testing.assert_array_equal(out["cell_ids"], o.cell_index(s.r))


def test_cell_ids_bit_exact():
    say("cell ids")
    rng = np.random.default_rng(7)
    for box, rc, mode in (((10.0, 7.3, 12.1), 2.5, 0), ((37.6386,) * 3, 2.5, 1),
                          ((9.0, 9.5, 11.0), 3.0, 0)):
        box = np.array(box)
        o = Oracle(1, scenarios.comp_1clj(), None, 1.0, rc, rc, box, periodic=True, cell_mode=mode)
        cpd, cl, _, _ = o.cell_geometry()
        pts = rng.uniform(0, box, size=(4000, 3))
        k = rng.integers(0, cpd - 2, size=(1000, 3))
        edges = np.minimum(k * cl, np.nextafter(box, 0))
        pts = np.vstack([pts, edges, np.nextafter(edges, np.inf), np.nextafter(edges, -np.inf).clip(0),
                         np.zeros((1, 3)), np.nextafter(box, 0)[None]])
        pts = np.minimum(pts, np.nextafter(box, 0))
        n = len(pts)
        with Context(0) as ctx:
            ctx.setup(1, scenarios.comp_1clj(), None, 1.0, box, rc, rc, True, mode)
            ctx.upload(pts, np.zeros((n, 3)), np.tile([1.0, 0, 0, 0], (n, 1)), np.zeros((n, 3)),
                       np.zeros(n, np.int32))
            ctx.update_container()
            got = ctx.cell_ids()
            gcpd, gcl = ctx.cell_geometry()
        np.testing.assert_array_equal(gcpd, cpd)
        np.testing.assert_array_equal(gcl, cl)
        np.testing.assert_array_equal(got, o.cell_index(pts))


@pytest.mark.parametrize("box,rc", [((10.0, 12.5, 7.5), 2.5), ((11.0, 7.0, 9.0), 3.0)])
def test_cell_pair_enumeration_equals_c08(box, rc):
    say("cell pairs", box, rc)
    box = np.array(box)
    o = Oracle(1, scenarios.comp_1clj(), None, 1.0, rc, rc, box, periodic=True)
    ref = {tuple(sorted(p)) for p in o.c08_pairs()}
    with Context(0) as ctx:
        ctx.setup(1, scenarios.comp_1clj(), None, 1.0, box, rc, rc, True)
        got = ctx.cell_pairs()
    assert {tuple(sorted(p)) for p in got} == ref
    assert len(got) == len(ref)


def test_leapfrog_vs_reference():
    say("leapfrog")
    desc = G["leapfrog/comp_desc"]
    ncomp = len(scenarios.parse_components(desc))
    mix = np.array([[1.0, 1.0]] * (ncomp * (ncomp - 1) // 2))
    cid = G["leapfrog/cid"]
    n = len(cid)
    o = Oracle(ncomp, desc, mix, 1.0, 2.5, 2.5, np.full(3, 100.0), periodic=True)
    # drive the device Leapfrog with prescribed forces: upload the state,
    # compute forces on a huge-box copy (no neighbours -> F = M = 0) is not
    # useful, so compare the half steps through a zero-force reference
    r1, v1, q1, L1 = o.leapfrog_pre(0.002, G["leapfrog/r0"], G["leapfrog/v0"], G["leapfrog/q0"],
                                    G["leapfrog/L0"], np.zeros((n, 3)), np.zeros((n, 3)), cid)
    with Context(0) as ctx:
        ctx.setup(ncomp, desc, mix, 1.0, np.full(3, 100.0), 2.5, 2.5, True)
        ctx.upload(G["leapfrog/r0"], G["leapfrog/v0"], G["leapfrog/q0"], G["leapfrog/L0"], cid)
        ctx.leapfrog_pre(0.002)
        out = ctx.download()
    np.testing.assert_allclose(out["r"], r1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(out["v"], v1, rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["q"], q1, rtol=0, atol=1e-14)
    np.testing.assert_allclose(out["L"], L1, rtol=0, atol=1e-15)


def test_trajectory_100_steps_fp64():
    """LJTS from an FCC start, 100 Leapfrog steps: per-step Upot/virial and
    final state against the oracle running the same step order."""
    say("trajectory 100")
    s = scenarios.lj_fcc(ncell=6)
    desc = np.array(scenarios.comp_1clj(rc_shift=s.rc, shifted=True))
    o = Oracle(1, desc, None, 1.0, s.rc, s.rc, s.box, periodic=True)
    ref = o.simulate(100, 0.002, s.r, s.v, s.q, s.L, s.cid)
    with Context(0) as ctx:
        ctx.setup(1, desc, None, 1.0, s.box, s.rc, s.rc, True)
        ctx.upload(s.r, s.v, s.q, s.L, s.cid)
        ctx.update_container()
        ctx.traverse()
        per = ctx.step(100, 0.002)
        out = ctx.download()
    rp = ref["per_step"]
    assert np.abs(per[:, 0] - rp[:, 0]).max() < 1e-9 * np.abs(rp[:, 0]).max()
    assert np.abs(per[:, 1] - rp[:, 1]).max() < 1e-9 * np.abs(rp[:, 1]).max()
    assert np.abs(per[:, 2] - rp[:, 2]).max() < 1e-9 * np.abs(rp[:, 2]).max()
    d = out["r"] - ref["r"]
    d -= np.round(d / s.box) * s.box
    assert np.abs(d).max() < 1e-9
    assert rel_err(out["F"], ref["F"]) < 1e-8


def test_graph_replay_matches_eager():
    say("graph replay")
    s = scenarios.small_mixture(seed=55)
    outs = []
    for use_graph in (False, True):
        with Context(0) as ctx:
            ctx.setup_system(s)
            ctx.update_container()
            ctx.traverse()
            per = [ctx.step(3, 0.001, use_graph=use_graph) for _ in range(3)]
            outs.append((np.vstack(per), ctx.download()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    for k in ("r", "v", "q", "L", "F", "M"):
        np.testing.assert_array_equal(outs[0][1][k], outs[1][1][k])


@pytest.mark.parametrize("sysname", ["small_lj", "small_mixture"])
def test_fused_steps_equal_single_steps(sysname):
    """A multi-step call does step n's second and step n+1's first half kick in
    one pass; the result equals one-step calls bit for bit."""
    say("fused steps", sysname)
    s = getattr(scenarios, sysname)(seed=57)
    outs = []
    for split in (False, True):
        with Context(0) as ctx:
            ctx.setup_system(s)
            ctx.update_container()
            ctx.traverse()
            if split:
                per = np.vstack([ctx.step(1, 0.001, use_graph=True) for _ in range(7)])
            else:
                per = ctx.step(7, 0.001, use_graph=True)
            outs.append((per, ctx.download()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    for k in ("r", "v", "q", "L", "F", "M", "Vi"):
        np.testing.assert_array_equal(outs[0][1][k], outs[1][1][k])


@pytest.mark.parametrize("sysname,seed", [("small_lj", 61), ("small_mixture", 62)])
def test_fp32_form_error_bound(sysname, seed):
    say("fp32 form", sysname)
    s = getattr(scenarios, sysname)(seed=seed)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    ref = o.forces(s.r, s.q, s.cid)
    out = gpu_forces(s, s.rc, s.rc_lj, precision=PREC_FP32_FP64ACC)
    ef = rms_rel(out["F"], ref["F"])
    say("  rel RMS force error", ef)
    assert ef < 1e-5
    if np.abs(ref["M"]).max() > 0:
        assert rms_rel(out["M"], ref["M"]) < 1e-5
    u, w = macro_to_upot_virial(ref["macro"])
    gu, gw = upot_virial(out["macro"])
    assert abs(gu - u) < 1e-5 * abs(u) and abs(gw - w) < 1e-5 * abs(w)
    np.testing.assert_array_equal(out["counters"][1:], ref["counters"][1:])
