"""The benchmark workloads against the FP64 oracle, at sizes the oracle
finishes in seconds, plus the split LJ / electrostatic cutoffs and the
rebuild (float cutoff) cell geometry.

Bars, as in test_gpu_parity.py: FP64 form -- forces, torques and virials per
molecule within 1e-12 relative to the largest oracle magnitude, Upot and
virial within 1e-12, in-cutoff site-pair counters equal; FP32-pair /
FP64-accumulate form -- relative RMS force (and torque) error <= 1e-5.
Trajectories (FP64, a few tens of steps from the same start) within 1e-10.

Workloads: the LJTS vapour-liquid slab at 0.2 of its linear size (same
densities, temperature, cutoff and aspect ratio); the 2CLJQ / water mixture
at its real cutoff 3 and density 0.5 with 8,000 molecules; a batch of 12
uneven replicas of the state-point sweep, each replica against its own
oracle run; the full 512-replica sweep once, with property checks.
"""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context, upot_virial
from ls1inp import read_inp
from oracle_lib import Oracle, macro_to_upot_virial
from test_gpu_parity import GOLD, check_against, rel_err, rms_rel

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_for(s, rc_lj=None, cell_mode=0, periodic=True):
    return Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj if rc_lj is None else rc_lj,
                  s.box, periodic=periodic, cell_mode=cell_mode)


def gpu_eval(s, precision, rc_lj=None, cell_mode=0, steps=0):
    with Context(0) as ctx:
        ctx.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc,
                  s.rc_lj if rc_lj is None else rc_lj, True, cell_mode, precision)
        ctx.upload(s.r, s.v, s.q, s.L, s.cid)
        ctx.set_counting(True)
        out = ctx.forces()
        out["counters"] = ctx.counters()
        # per-kernel visit counts: each kernel its own side, halved in total
        kc = ctx.kernel_stats()[1]
        np.testing.assert_array_equal(kc.sum(0) // 2, out["counters"])
        ctx.set_counting(False)
        if steps:
            out["per_step"] = ctx.step(steps, s.dt, use_graph=True)
            out["after"] = ctx.download()
        return out


def check_fp64(out, ref):
    u, w = macro_to_upot_virial(ref["macro"])
    check_against(out, ref["F"], ref["M"], ref["Vi"], u, w)
    np.testing.assert_array_equal(out["counters"][1:], ref["counters"][1:])


def check_fp32(out, ref):
    ef = rms_rel(out["F"], ref["F"])
    say("  fp32 rel RMS force error", ef)
    assert ef < 1e-5
    if np.abs(ref["M"]).max() > 0:
        assert rms_rel(out["M"], ref["M"]) < 1e-5
    u, w = macro_to_upot_virial(ref["macro"])
    gu, gw = upot_virial(out["macro"])
    assert abs(gu - u) < 1e-5 * abs(u) and abs(gw - w) < 1e-5 * abs(w)
    np.testing.assert_array_equal(out["counters"][1:], ref["counters"][1:])


def check_trajectory(out, ref, box):
    rp, per = ref["per_step"], out["per_step"]
    for col in range(4):
        scale = max(np.abs(rp[:, col]).max(), 1e-300)
        assert np.abs(per[:, col] - rp[:, col]).max() <= TRAJ_TOL * scale, col
    a = out["after"]
    d = a["r"] - ref["r"]
    d -= np.round(d / box) * box
    assert np.abs(d).max() < TRAJ_TOL * box.max()
    assert rel_err(a["v"], ref["v"]) < TRAJ_TOL
    assert rel_err(a["F"], ref["F"]) < TRAJ_TOL
    if np.abs(ref["L"]).max() > 0:
        assert rel_err(a["L"], ref["L"]) < TRAJ_TOL
        assert np.abs(a["q"] - ref["q"]).max() < TRAJ_TOL


def test_ljts_slab_scaled():
    s = scenarios.ljts_slab(scale=0.2)
    say("case ljts_slab scale 0.2 N", s.n, "box", s.box.tolist())
    o = oracle_for(s)
    ref = o.forces(s.r, s.q, s.cid)
    check_fp64(gpu_eval(s, PREC_FP64), ref)
    check_fp32(gpu_eval(s, PREC_FP32_FP64ACC), ref)
    steps = 20
    say("  trajectory", steps)
    traj = o.simulate(steps, s.dt, s.r, s.v, s.q, s.L, s.cid)
    check_trajectory(gpu_eval(s, PREC_FP64, steps=steps), traj, s.box)


def test_mixture_rc3():
    s = scenarios.mixture(n=8000)
    say("case mixture N", s.n, "rc", s.rc, "box", s.box[0])
    o = oracle_for(s)
    ref = o.forces(s.r, s.q, s.cid)
    assert np.abs(ref["M"]).max() > 0
    check_fp64(gpu_eval(s, PREC_FP64), ref)
    check_fp32(gpu_eval(s, PREC_FP32_FP64ACC), ref)
    steps = 10
    say("  trajectory", steps)
    traj = o.simulate(steps, s.dt, s.r, s.v, s.q, s.L, s.cid)
    check_trajectory(gpu_eval(s, PREC_FP64, steps=steps), traj, s.box)


def test_mixture_dense_cells():
    """The mixture on 3 x 3 x 3 cells (rc 4.9, about 59 molecules per cell):
    each cell's charges span two 64-lane chunks, and every charge and
    quadrupole pass walks the whole box in many LDS chunks, split at molecule
    boundaries for charge partners."""
    s = scenarios.mixture(n=1600, rc=4.9)
    say("case mixture dense cells N", s.n, "rc", s.rc, "box", s.box[0])
    assert int(s.box[0] / s.rc) == 3
    o = oracle_for(s)
    ref = o.forces(s.r, s.q, s.cid)
    check_fp64(gpu_eval(s, PREC_FP64), ref)
    check_fp32(gpu_eval(s, PREC_FP32_FP64ACC), ref)


def test_mixture_cell_edge_equal_to_cutoff():
    """A box of exactly 6 rc: the cell edge equals the cutoff, so the
    quadrupole kernel runs cell by cell instead of two cells per tile (the
    common tile holds cells a cell edge away from one of the two, which are
    only out of range when the edge exceeds rc); both forms against the
    oracle as in the mixture test. The other mixture tests cover the two-cell
    tiles (cell edges above rc) and a row of three cells."""
    s = scenarios.mixture(n=3000)
    scale = 6.0 * s.rc / s.box[0]
    s.r = s.r * scale
    s.box = np.full(3, 6.0 * s.rc)
    say("case mixture box 6 rc N", s.n, "box", s.box[0])
    o = oracle_for(s)
    ref = o.forces(s.r, s.q, s.cid)
    assert np.abs(ref["M"]).max() > 0
    check_fp64(gpu_eval(s, PREC_FP64), ref)
    check_fp32(gpu_eval(s, PREC_FP32_FP64ACC), ref)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32_FP64ACC])
def test_replica_batch_each_vs_oracle(prec):
    reps = scenarios.replicas(nrep=12, nmin=2000, nmax=4000)
    say("case replicas x12 prec", prec, [s.n for s in reps])
    steps = 10
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=prec)
        c.set_counting(True)
        c.update_container()
        mac = c.traverse()
        cnt = c.counters()
        c.set_counting(False)
        f0 = c.download()
        per = c.step(steps, reps[0].dt, use_graph=True) if prec == PREC_FP64 else None
        f1 = c.download()
    off = np.cumsum([0] + [s.n for s in reps])
    tot = np.zeros(8, np.int64)
    for k, s in enumerate(reps):
        sl = slice(off[k], off[k + 1])
        o = oracle_for(s)
        ref = o.forces(s.r, s.q, s.cid)
        tot += ref["counters"]
        got = {key: f0[key][sl] for key in ("F", "M", "Vi")}
        got["macro"] = mac[k]
        if prec == PREC_FP64:
            u, w = macro_to_upot_virial(ref["macro"])
            check_against(got, ref["F"], ref["M"], ref["Vi"], u, w)
            traj = o.simulate(steps, s.dt, s.r, s.v, s.q, s.L, s.cid)
            out = {"per_step": per[:, k, :], "after": {key: f1[key][sl] for key in f1}}
            check_trajectory(out, traj, s.box)
        else:
            assert rms_rel(got["F"], ref["F"]) < 1e-5, k
            u, _ = macro_to_upot_virial(ref["macro"])
            assert abs(upot_virial(mac[k])[0] - u) < 1e-5 * abs(u), k
    # the batch's pair counters are the sums of the replicas' own
    np.testing.assert_array_equal(cnt[1:], tot[1:])


def test_full_replica_sweep_properties():
    """All 512 replicas of the sweep (6.29 M molecules) in one context:
    finite forces and sums for every replica, one step, and three sampled
    replicas bit-identical to running them alone; the smallest replica
    against the oracle."""
    reps = scenarios.replicas()
    say("case replicas x512 N", sum(s.n for s in reps))
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP32_FP64ACC)
        c.update_container()
        mac = c.traverse()
        f0 = c.download()
        per = c.step(2, reps[0].dt, use_graph=True)
    assert mac.shape == (512, 4) and np.all(np.isfinite(mac))
    assert np.all(np.isfinite(per)) and np.all(np.isfinite(f0["F"])) and np.all(np.isfinite(f0["Vi"]))
    assert np.all(per[:, :, 2] > 0)  # every replica has kinetic energy
    off = np.cumsum([0] + [s.n for s in reps])
    ns = np.array([s.n for s in reps])
    for k in (int(np.argmin(ns)), int(np.argmax(ns)), 257):
        s = reps[k]
        sl = slice(off[k], off[k + 1])
        say("  sampled replica", k, "N", s.n)
        with Context(0) as c:
            c.setup_system(s, periodic=True, precision=PREC_FP32_FP64ACC)
            c.update_container()
            m1 = c.traverse()
            d1 = c.download()
        assert np.array_equal(f0["F"][sl], d1["F"]) and np.array_equal(f0["Vi"][sl], d1["Vi"])
        np.testing.assert_allclose(mac[k], m1, rtol=1e-12, atol=1e-12 * np.abs(m1).max())
    k = int(np.argmin(ns))
    s = reps[k]
    ref = oracle_for(s).forces(s.r, s.q, s.cid)
    assert rms_rel(f0["F"][off[k]:off[k + 1]], ref["F"]) < 1e-5


# --------------------------------------------------- split cutoffs rcLJ < rc

SPLIT_FIXTURES = [("VectorizationWater", 12.0), ("VectorizationMultiComponentMultiPotentials", 35.0)]


@pytest.mark.parametrize("name,rc", SPLIT_FIXTURES)
def test_split_cutoff_fixture(name, rc):
    ps = read_inp(GOLD / "inp" / f"{name}.inp")
    rc_lj = 0.8 * rc
    say("case split cutoff", name, rc, rc_lj)
    o = Oracle.from_phase_space(ps, rc, rc_lj, periodic=False)
    ref = o.forces(ps.r, ps.q, ps.cid)
    full = Oracle.from_phase_space(ps, rc, rc, periodic=False).forces(ps.r, ps.q, ps.cid)
    assert ref["counters"][1] < full["counters"][1]  # the split removes LJ pairs
    for prec in (PREC_FP64, PREC_FP32_FP64ACC):
        with Context(0) as ctx:
            ctx.setup(ps.ncomp, ps.comp_desc, ps.mix, ps.eps_rf, ps.box, rc, rc_lj, False, 0, prec)
            ctx.upload(ps.r, ps.v, ps.q, ps.L, ps.cid)
            ctx.set_counting(True)
            out = ctx.forces()
            out["counters"] = ctx.counters()
        (check_fp64 if prec == PREC_FP64 else check_fp32)(out, ref)


@pytest.mark.parametrize("sysname", ["small_mixture", "mixture"])
def test_split_cutoff_periodic(sysname):
    s = scenarios.small_mixture(seed=71) if sysname == "small_mixture" else scenarios.mixture(n=4000)
    rc_lj = 0.8 * s.rc if sysname == "small_mixture" else 2.5
    say("case split cutoff", sysname, "rc", s.rc, "rcLJ", rc_lj)
    o = oracle_for(s, rc_lj=rc_lj)
    ref = o.forces(s.r, s.q, s.cid)
    check_fp64(gpu_eval(s, PREC_FP64, rc_lj=rc_lj), ref)
    check_fp32(gpu_eval(s, PREC_FP32_FP64ACC, rc_lj=rc_lj), ref)


def test_split_cutoff_nvt_with_long_range_correction():
    """rcLJ = 2.5 < rc = 3 with the homogeneous long-range correction and the
    thermostat: the correction integrates the LJ tail from rcLJ."""
    s = scenarios.mixture(n=4000)
    rc_lj = 2.5
    say("case split cutoff NVT + LRC")
    o = oracle_for(s, rc_lj=rc_lj)
    steps = 5
    ref = o.simulate_nvt(steps, s.dt, s.r, s.v, s.q, s.L, s.cid, 1.0, lrc=True)
    with Context(0) as ctx:
        ctx.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, rc_lj, True, 0, PREC_FP64)
        ctx.upload(s.r, s.v, s.q, s.L, s.cid)
        ctx.set_thermostat(1.0, lrc=True)
        ctx.update_container()
        ctx.traverse()
        per = ctx.step(steps, s.dt, use_graph=True)
        st = ctx.thermostat_state()
        after = ctx.download()
    rp = ref["per_step"]
    for col in range(4):
        assert np.abs(per[:, col] - rp[:, col]).max() <= 1e-10 * np.abs(rp[:, col]).max(), col
    np.testing.assert_allclose(st[0, 1:4], ref["lrc"], rtol=1e-13)
    assert rel_err(after["v"], ref["v"]) < 1e-10


# ------------------------------------- rebuild geometry (float cutoff) cases

F32 = lambda x: float(np.float32(x))  # noqa: E731

# x: 9 cells with the double cutoff, 10 with the rebuild's float one; then a
# box with a single rebuild cell in x (rc <= L < 2 rc)
GEOM_BOXES = [(np.array([25.9999995, 13.0, 26.0]), 2.6), (np.array([10 * F32(2.6), 7.8, 9.1]), 2.6),
              (np.array([4.0, 8.0, 11.0]), 2.6)]


def lj_system(box, rc, n, seed):
    rng = np.random.default_rng(seed)
    r = scenarios.rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.9)
    n = len(r)
    desc = np.array(scenarios.comp_1clj(rc_shift=rc, shifted=True), np.float64)
    return scenarios.System("geom", 1, desc, np.zeros((0, 2)), 1.0, rc, rc, box, r,
                            rng.normal(scale=0.5, size=(n, 3)), np.tile([1.0, 0, 0, 0], (n, 1)),
                            np.zeros((n, 3)), np.zeros(n, np.int32), 1.0)


def brute_images(s):
    """All pairs over the 27 image shifts {-L, 0, L}^3 within rc (LJ,
    truncated-shifted): with a single cell in a dimension every image the
    linked cells see is one of these, and none further is within rc."""
    eps24, sig2, shift6 = 24.0, 1.0, 0.0
    o = oracle_for(s)
    shift6 = o.lj_tables()[2][0, 0]
    F = np.zeros_like(s.r)
    u6 = 0.0
    shifts = np.array([[a, b, c] for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)]) * s.box
    for i in range(s.n):
        d = s.r[i][None, None, :] - (s.r[None, :, :] + shifts[:, None, :])
        r2 = (d ** 2).sum(-1)
        m = (r2 < s.rc ** 2) & (r2 > 0)
        r2i = np.where(m, 1.0 / np.where(m, r2, 1.0), 0.0)
        l6 = (sig2 * r2i) ** 3
        sc = eps24 * r2i * (2 * l6 * l6 - l6)
        F[i] = (d * sc[..., None]).sum((0, 1))
        u6 += 0.5 * np.where(m, eps24 * (l6 * l6 - l6) + shift6, 0.0).sum()
    return F, u6 / 6.0


@pytest.mark.parametrize("case", range(len(GEOM_BOXES)))
def test_rebuild_geometry(case):
    box, rc = GEOM_BOXES[case]
    s = lj_system(box, rc, int(0.5 * np.prod(box)), seed=90 + case)
    say("case rebuild geometry", box.tolist(), "rc", rc, "N", s.n)
    o = oracle_for(s, cell_mode=1)
    cpd, cl, _, _ = o.cell_geometry()
    ref = o.forces(s.r, s.q, s.cid)
    with Context(0) as ctx:
        ctx.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, rc, rc, True, 1, PREC_FP64)
        ctx.upload(s.r, s.v, s.q, s.L, s.cid)
        ctx.set_counting(True)
        out = ctx.forces()
        out["counters"] = ctx.counters()
        gcpd, gcl = ctx.cell_geometry()
        ids = ctx.cell_ids()
        out["per_step"] = ctx.step(5, s.dt, use_graph=True)
        out["after"] = ctx.download()
    np.testing.assert_array_equal(gcpd, cpd)
    np.testing.assert_array_equal(gcl, cl)
    np.testing.assert_array_equal(ids, o.cell_index(s.r))
    check_fp64(out, ref)
    check_trajectory(out, o.simulate(5, s.dt, s.r, s.v, s.q, s.L, s.cid), s.box)
    if cpd.min() == 3:
        F, u = brute_images(s)
        assert rel_err(ref["F"], F) < 1e-12
        assert abs(macro_to_upot_virial(ref["macro"])[0] - u) < 1e-12 * abs(u)


def test_ctor_geometry_refuses_one_cell():
    """The ctor (cell_mode 0) needs two cells per dimension (LinkedCells.cpp:
    79-96); the rebuild accepts one."""
    box, rc = GEOM_BOXES[2]
    with Context(0) as ctx:
        with pytest.raises(Exception):
            ctx.setup(1, scenarios.comp_1clj(), None, 1.0, box, rc, rc, True, 0, PREC_FP64)
