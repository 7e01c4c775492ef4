#!/usr/bin/env python3
"""Benchmark: molecule-steps/s of the short-range force path on one MI355X.

Workload (default): the multi-site mixture the metric's target is quoted on --
1,000,000 rigid molecules in a periodic cube at rho 0.5, T 1.0, rc 3: 50 %
2CLJQ (two LJ sites 0.5 apart, a point quadrupole Q = 2 on the axis), 50 %
4-site water-like (LJ oxygen, charges +0.52, +0.52, -1.04), seeded random
orientations, Leapfrog dt 0.002. One step = Leapfrog half kick + drift +
periodic wrap, counting-sort re-bin, SoA site build, force traversal (LJ,
charge and quadrupole pair kernels), calcFM, Upot/virial sums, Leapfrog half
kick -- all on the device, one hipGraph per step, inputs resident in HBM
before timing starts.

Both precision forms run: the FP32-pair / FP64-accumulate form gives `value`
(the metric's chosen form); the plain FP64 form, the ratio of the two and the
FP32 form's force error against the FP64 form are reported beside it. Every
pair kernel gets its own roofline entry (FlopCounter flops of its share of
the pairs / its HIP-event duration); `roofline` is the longest one of the
chosen form. The LJTS vapour-liquid slab (3.96 M one-centre molecules) runs
too and is reported under `ljts_slab`.

The CPU baseline is the C oracle (a scalar restatement of the reference
algorithm, C08 traversal parallelised with OpenMP as the reference's is)
timed on this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "ls1-mardyn_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import scenarios  # noqa: E402
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context  # noqa: E402

# MI355X published peaks
PEAK_FP32_VECTOR_TF = 157.3
PEAK_FP64_VECTOR_TF = 78.6
# FlopCounter convention (src/particleContainer/adapter/FlopCounter.cpp:19-76):
# 8 flops per molecule-centre distance test, per site pair inside the cutoff
# LJ 41, C-C 34, C-D 62, D-D 135, C-Q 79, D-Q 151, Q-Q 161.
FLOPS_PER_PAIR = np.array([8, 41, 34, 62, 135, 79, 151, 161], dtype=np.float64)


PAIR_SHAPE = None  # (cells per block, blocks per XCD) or None: autotune


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNEL_NAMES = ("k_pairs<LJ>", "k_pairs<charge>", "k_pairs<dipole>", "k_pairs<quadrupole>")


def measured_traffic(workload, scale, tag, shape, kernel):
    """HBM bytes per launch of one pair kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (copy-calibrated), or None when no
    summary covers this workload, form, launch shape and kernel."""
    if scale != 1.0:
        return None
    f = ROOT / "profiles" / f"k_pairs_{workload}_traffic.json"
    try:
        d = json.loads(f.read_text())
        return d[tag]["%d,%d" % tuple(shape)][kernel]["traffic_bytes"]
    except (OSError, ValueError, KeyError, TypeError):
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_workload(name, scale):
    if name == "ljts_slab":
        return scenarios.ljts_slab(scale=scale)
    if name == "lj_fcc":
        return scenarios.lj_fcc()
    if name == "mixture":
        return scenarios.mixture(n=int(1_000_000 * scale ** 3))
    if name == "binary_lj":
        return scenarios.binary_lj(n=int(500_000 * scale ** 3))
    raise SystemExit(f"unknown workload {name}")


def run_gpu(s, precision, steps, warmup, kernel_reps, rmm=False):
    t_start = time.perf_counter()

    def phase(msg):
        log(f"    [{time.perf_counter() - t_start:7.2f} s] {msg}")

    with Context(0) as ctx:
        ctx.setup_system(s, periodic=True, precision=precision)
        phase("uploaded")
        ctx.update_container()
        ctx.traverse()
        init = ctx.download()
        phase("first traversal")
        # work counters of one traversal (separate, uncounted runs are timed)
        ctx.set_counting(True)
        ctx.traverse()
        counters = ctx.counters()
        kcounters = ctx.kernel_stats()[1]
        ctx.set_counting(False)
        phase("counted traversal")
        if PAIR_SHAPE is None:
            shape, _ = ctx.autotune(reps=5)
            phase(f"autotuned launch shape (cells per block, blocks per XCD) = {shape}")
        else:
            shape = PAIR_SHAPE
            ctx.set_pair_config(*shape)
            phase(f"fixed launch shape (cells per block, blocks per XCD) = {shape}")
        # pair-kernel duration: HIP events on the context stream around the
        # pair kernels, averaged over kernel_reps traversals
        ctx.set_timing(True)
        pair_ms, kernel_ms = [], []
        for _ in range(kernel_reps):
            ctx.traverse()
            pair_ms.append(ctx.last_kernel_ms()[0])
            kernel_ms.append(ctx.kernel_stats()[0])
        ctx.set_timing(False)
        phase(f"timed traversals: pair kernels {np.mean(pair_ms):.3f} ms")
        if warmup:
            ctx.step(warmup, s.dt, use_graph=True)
        ctx.synchronize()
        phase("warm-up steps")
        try:
            import torch

            if torch.cuda.is_available():
                torch.cuda.synchronize()
        except Exception:
            pass
        t0 = time.perf_counter()
        per = ctx.step(steps, s.dt, use_graph=True)
        ctx.synchronize()
        t1 = time.perf_counter()
        rmm_s = None
        if rmm:
            # the reduced-memory processor on the same state (half kick, then
            # r-update / re-bin / force-with-kick / v^2 sums per step)
            ctx.set_rmm(s.dt)
            if warmup:
                ctx.step(warmup, s.dt, use_graph=True)
            ctx.synchronize()
            t2 = time.perf_counter()
            ctx.step(steps, s.dt, use_graph=True)
            ctx.synchronize()
            rmm_s = time.perf_counter() - t2
            phase(f"reduced-memory mode: {rmm_s * 1e3 / steps:.3f} ms/step")
    return dict(seconds=t1 - t0, per_step=per, counters=counters, pair_ms=float(np.mean(pair_ms)),
                pair_ms_all=pair_ms, init=init, shape=shape, rmm_seconds=rmm_s,
                kernel_counters=kcounters, kernel_ms=np.mean(np.array(kernel_ms), axis=0))


def kernel_rooflines(r, tag, workload, scale, ref_checks):
    """One roofline entry per pair kernel that ran, on the FlopCounter basis
    (FlopCounter.cpp:19-76): the site pairs the kernel evaluates (its share:
    every site evaluates its own side, so a pair of one type counts half in
    its kernel's visits twice and a mixed pair half in each of the two
    kernels) at the per-type flops, plus 8 flops per molecule-centre check,
    with ls1's count of checks (each unordered molecule pair once, from the
    oracle's C08 walk) shared among the kernels in proportion to the checks
    each makes; the kernels' flops then add up to the FlopCounter total of a
    traversal. flops_issued_basis counts each kernel's own checks instead (a
    multi-site kernel checks every site pair's molecule centres)."""
    peak = PEAK_FP32_VECTOR_TF if tag == "fp32" else PEAK_FP64_VECTOR_TF
    kc = np.asarray(r["kernel_counters"], np.float64) / 2.0
    own_checks = kc[:, 0].sum()
    out = []
    for t, name in enumerate(KERNEL_NAMES):
        ms = float(r["kernel_ms"][t])
        if ms <= 0:
            continue
        cnt = kc[t]
        site_flops = float((FLOPS_PER_PAIR[1:] * cnt[1:]).sum())
        share = cnt[0] / own_checks if own_checks > 0 else 0.0
        flops = site_flops + 8.0 * ref_checks * share
        achieved = flops / (ms * 1e-3) / 1e12
        out.append({"kernel": name, "bound": "valu", "achieved": achieved, "peak": peak,
                    "unit": "TFLOP/s", "frac": achieved / peak, "flops_per_launch": flops,
                    "flops_issued_basis": site_flops + 8.0 * float(cnt[0]),
                    "kernel_ms": ms, "molecule_checks_made": float(cnt[0]),
                    "site_pairs": {k: float(v) for k, v in zip(
                        ("LJ", "CC", "CD", "DD", "CQ", "DQ", "QQ"), cnt[1:]) if v},
                    "traffic": measured_traffic(workload, scale, tag, r["shape"], name)})
    return out


def cpu_baseline(workload, budget_s):
    """The oracle (scalar C restatement of the reference path; its C08
    traversal and molecule loops run on OpenMP threads as the reference's do)
    on a bounded sample of the workload: same physics and densities, fewer
    molecules. Threads = OMP_NUM_THREADS (the host share this process gets),
    else every core."""
    from oracle_lib import Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    if workload == "ljts_slab":
        s = scenarios.ljts_slab(scale=0.4)
        sample = f"LJTS slab scaled to 256x64x64 ({s.n:,} molecules), same densities/T/rc"
    elif workload == "mixture":
        s = scenarios.mixture(n=200_000)
        sample = f"multi-site mixture, {s.n:,} molecules (same rho, T, rc, 50/50 2CLJQ/water)"
    else:
        s = scenarios.lj_fcc(ncell=20)
        sample = f"LJ FCC 20^3 cells ({s.n:,} molecules)"
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    o.set_threads(threads)
    t0 = time.perf_counter()
    o.forces(s.r, s.q, s.cid)
    one = time.perf_counter() - t0
    nsteps = max(1, int(budget_s / max(one, 1e-3)))
    t0 = time.perf_counter()
    o.simulate(nsteps, s.dt, s.r, s.v, s.q, s.L, s.cid)
    dt = time.perf_counter() - t0
    return dict(value=s.n * nsteps / dt, unit="molecule-steps/s", cores=threads, kind="port",
                sample=f"{sample}, {nsteps} steps, {dt:.1f} s, {threads} OpenMP threads",
                cpu_model=cpu_model(),
                reference_probe={
                    "value": 4.02e6, "unit": "molecule-steps/s", "cores": 8,
                    "note": "the reference itself (DPDP, AVX2, C08, 8 OpenMP threads) on a 32,448-"
                            "molecule one-centre LJ fluid in the survey container (Xeon @ 2.7 GHz); "
                            "orientation only, not this host"})


def run_step_variants(args):
    """The step around the pair kernels in its ls1 variants, on the 1 M
    mixture in the FP32 form (same state, same launch shape, one hipGraph per
    step): plain NVE steps; NVT steps with the velocity-scaling thermostat and
    the homogeneous long-range correction on the device; reflecting walls in
    z; outflow walls in z (leaving molecules deleted, survivors compacted at
    the next re-bin). Milliseconds per step for each, and the NVE pair-kernel
    time for reference."""
    from ls1amd import BOUNDARY_OUTFLOW, BOUNDARY_PERIODIC, BOUNDARY_REFLECTING

    s = scenarios.mixture(n=int(1_000_000 * args.scale ** 3))
    log(f"step variants: mixture N = {s.n}")
    out = {"workload": "mixture step variants", "molecules": int(s.n), "steps": args.steps,
           "precision": "FP32 pair / FP64 accumulate"}
    variants = (
        ("nve", None, None),
        ("nvt_thermostat_lrc", s.temperature, None),
        ("walls_reflecting_z", None, BOUNDARY_REFLECTING),
        ("walls_outflow_z", None, BOUNDARY_OUTFLOW),
    )
    for name, T, wall in variants:
        with Context(0) as ctx:
            ctx.setup_system(s, periodic=True, precision=PREC_FP32_FP64ACC)
            ctx.set_pair_config(*(PAIR_SHAPE or (1, 0)))
            if T is not None:
                ctx.set_thermostat(np.full(1, T), lrc=True)
            if wall is not None:
                ctx.set_boundaries([BOUNDARY_PERIODIC, BOUNDARY_PERIODIC, wall])
            ctx.update_container()
            ctx.traverse()
            ctx.step(args.warmup, s.dt, use_graph=True)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.step(args.steps, s.dt, use_graph=True)
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            rec = {"ms_per_step": ms, "molecule_steps_per_s": s.n * 1e3 / ms}
            if wall == BOUNDARY_OUTFLOW:
                rec["molecules_left"] = int(ctx.molecule_count())
            out[name] = rec
            log(f"  {name}: {ms:.3f} ms/step")
    return out


def run_replicas(args):
    """Batched replicas: all replicas in one context, one launch per kernel
    per step over the flattened cell table, one hipGraph per step, autotuned
    pair-kernel launch shape; compared with the default launch shape and with
    per-replica plain launches (one context and graph per replica, stepped
    one after another)."""
    reps = scenarios.replicas(nrep=args.replicas)
    ntot = int(sum(s.n for s in reps))
    dt = reps[0].dt
    log(f"replicas: {len(reps)} systems, {ntot} molecules")
    out = {}
    with Context(0) as ctx:
        ctx.setup_replicas(reps, periodic=True, precision=PREC_FP32_FP64ACC)
        ctx.update_container()
        ctx.traverse()
        ctx.set_counting(True)
        ctx.traverse()
        counters = ctx.counters()
        ctx.set_counting(False)
        timed = {}
        for tag in ("default", "autotuned"):
            if tag == "autotuned":
                (cpb, bpx), best = ctx.autotune(reps=3)
                out["autotuned_shape"] = {"cells_per_block": cpb, "blocks_per_xcd": bpx}
            ctx.set_timing(True)
            pm = []
            for _ in range(args.kernel_reps):
                ctx.traverse()
                pm.append(ctx.last_kernel_ms()[0])
            ctx.set_timing(False)
            ctx.step(args.warmup, dt, use_graph=True)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.step(args.steps, dt, use_graph=True)
            ctx.synchronize()
            sec = time.perf_counter() - t0
            timed[tag] = {"value": ntot * args.steps / sec, "ms_per_step": sec * 1e3 / args.steps,
                          "pair_kernel_ms": float(np.mean(pm))}
            log(f"  {tag}: {timed[tag]['ms_per_step']:.3f} ms/step, pair {timed[tag]['pair_kernel_ms']:.3f} ms")
    out.update(timed)
    # per-replica plain launches
    ctxs = []
    try:
        for s in reps:
            c = Context(0)
            ctxs.append(c)
            c.setup_system(s, periodic=True, precision=PREC_FP32_FP64ACC)
            c.update_container()
            c.traverse()
            c.step(1, dt, use_graph=True)
        for c in ctxs:
            c.synchronize()
        t0 = time.perf_counter()
        for c in ctxs:
            c.step(args.plain_steps, dt, use_graph=True)
        for c in ctxs:
            c.synchronize()
        sec = time.perf_counter() - t0
        out["per_replica_launches"] = {"value": ntot * args.plain_steps / sec,
                                       "ms_per_step": sec * 1e3 / args.plain_steps,
                                       "steps": args.plain_steps}
        log(f"  per-replica launches: {out['per_replica_launches']['ms_per_step']:.3f} ms/step")
    finally:
        for c in ctxs:
            c.close()
    flops = float((FLOPS_PER_PAIR * counters).sum())
    best = out["autotuned"]
    achieved = flops / (best["pair_kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": "molecule-steps/s on 1 MI355X, batched replicas (FP32 pair / FP64 accumulate)",
        "value": best["value"], "unit": "molecule-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 accumulation/sums/Leapfrog",
        "data": "synthetic (seeded random sequential addition, Maxwell-Boltzmann velocities)",
        "config": {"workload": "replicas", "replicas": len(reps), "molecules": ntot,
                   "n_range": [2000, 40000], "rho_range": [0.05, 0.85], "T_range": [0.7, 1.3],
                   "rc": 2.5, "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VECTOR_TF,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VECTOR_TF, "traffic": None,
                     "kernel": "k_pairs<LJ>", "flops_per_launch": flops,
                     "kernel_ms": best["pair_kernel_ms"]},
        "replicas_detail": out,
        "speedup_vs_per_replica_launches": best["value"] / out["per_replica_launches"]["value"],
    }
    if not args.no_cpu_baseline:
        from oracle_lib import Oracle

        t0, done, k = time.perf_counter(), 0, 0
        while time.perf_counter() - t0 < args.cpu_budget and k < len(reps):
            s = reps[k]
            o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
            o.simulate(2, s.dt, s.r, s.v, s.q, s.L, s.cid)
            done += 2 * s.n
            k += 1
        sec = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": done / sec, "unit": "molecule-steps/s", "cores": 1,
                                "kind": "port",
                                "sample": f"first {k} replicas run one after another, 2 steps each, {sec:.1f} s"}
    return line


def run_adjoint(args):
    """Differentiation workload: binary LJ mixture, 500,000 molecules, FP64:
    forward traversal, then the adjoint (6 eps/sigma pair derivatives of
    Upot and virial + the force VJP for a seeded Gaussian cotangent)."""
    s = scenarios.binary_lj(n=int(500_000 * args.scale ** 3))
    cot = np.random.default_rng(9).normal(size=(s.n, 3))
    with Context(0) as ctx:
        ctx.setup_system(s, periodic=True, precision=PREC_FP64)
        ctx.update_container()
        ctx.traverse()
        for _ in range(args.warmup):
            ctx.adjoint(cot)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.traverse()
        fwd = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.adjoint(cot)
        adj = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.adjoint(None)
        adj_p = (time.perf_counter() - t0) / args.steps
        # kernel times (HIP events on the context stream, transfers excluded)
        ctx.set_timing(True)
        fk, ak, vk = [], [], []
        for _ in range(args.steps):
            ctx.traverse()
            fk.append(ctx.last_kernel_ms()[0])
            ctx.adjoint(cot)
            a_ms, v_ms = ctx.last_kernel_ms()
            ak.append(a_ms)
            vk.append(v_ms)
        ctx.set_timing(False)
    fwd_k = float(np.mean(fk))
    return {"workload": "binary_lj adjoint", "molecules": int(s.n),
            "forward_pair_kernel_ms": fwd_k, "adjoint_kernels_ms": float(np.mean(ak)),
            "adjoint_kernels_with_vjp_ms": float(np.mean(vk)),
            "adjoint_over_forward": float(np.mean(vk)) / fwd_k,
            "host_call_ms": {"forward_traverse": fwd * 1e3, "adjoint_with_vjp": adj * 1e3,
                             "adjoint_params_only": adj_p * 1e3},
            "note": "kernel times from HIP events; host-call times include the cotangent "
                    "upload and the result download (pageable host memory)"}


def rel_rms(a, b):
    den = float((b ** 2).sum(1).mean())
    return float(np.sqrt(((a - b) ** 2).sum(1).mean() / den)) if den > 0 else None


def precision_report(scale, drift_steps):
    """Error of the FP32-pair / FP64-accumulate form against the plain FP64
    form (which matches the FP64 oracle to 1e-12, tests/test_gpu_parity.py):
    one traversal of the multi-site mixture and of the LJTS slab (force,
    torque, Upot, virial), and total-energy drift over drift_steps Leapfrog
    steps of the 32,000-molecule FCC LJ fluid (full LJ as specified, and the
    truncated-shifted variant whose energy is continuous at rc)."""
    from ls1amd import upot_virial

    out = {}
    cases = (("mixture", lambda: scenarios.mixture(n=int(1_000_000 * scale ** 3))),
             ("ljts_slab", lambda: scenarios.ljts_slab(scale=scale)))
    for name, make in cases:
        s = make()
        log(f"precision: {name} N = {s.n}")
        res = {}
        for tag, p in (("fp32", PREC_FP32_FP64ACC), ("fp64", PREC_FP64)):
            with Context(0) as ctx:
                ctx.setup_system(s, periodic=True, precision=p)
                ctx.update_container()
                mac = ctx.traverse()
                d = ctx.download()
                res[tag] = (d, upot_virial(mac))
        (d32, (u32, w32)), (d64, (u64, w64)) = res["fp32"], res["fp64"]
        out[name] = {"molecules": int(s.n),
                     "force_rel_rms": rel_rms(d32["F"], d64["F"]),
                     "force_max_abs_over_rms": float(np.abs(d32["F"] - d64["F"]).max()
                                                     / np.sqrt((d64["F"] ** 2).sum(1).mean())),
                     "torque_rel_rms": rel_rms(d32["M"], d64["M"]),
                     "virial_tensor_rel_rms": rel_rms(d32["Vi"], d64["Vi"]),
                     "upot_rel": abs(u32 - u64) / abs(u64), "virial_rel": abs(w32 - w64) / abs(w64)}
    for shifted in (False, True):
        s = scenarios.lj_fcc(shifted=shifted)
        key = f"energy_drift_lj_fcc_{'shifted' if shifted else 'full'}_{drift_steps}_steps"
        log(f"precision: {key}")
        rec = {}
        for tag, p in (("fp32", PREC_FP32_FP64ACC), ("fp64", PREC_FP64)):
            with Context(0) as ctx:
                ctx.setup_system(s, periodic=True, precision=p)
                ctx.update_container()
                u0, _ = upot_virial(ctx.traverse())
                e0 = u0 + 0.5 * float((s.v ** 2).sum())  # m = 1
                per = np.concatenate([ctx.step(min(1000, drift_steps - k), s.dt, use_graph=True)
                                      for k in range(0, drift_steps, 1000)])
            e = per[:, 0] + 0.5 * (per[:, 2] + per[:, 3])
            rec[tag] = {"E0": e0, "E_end": float(e[-1]), "drift_rel": float((e[-1] - e0) / abs(e0)),
                        "max_dev_rel": float(np.abs(e - e0).max() / abs(e0)),
                        "upot_end": float(per[-1, 0])}
        rec["energy_diff_end_rel"] = abs(rec["fp32"]["E_end"] - rec["fp64"]["E_end"]) / abs(
            rec["fp64"]["E_end"])
        out[key] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixture")
    ap.add_argument("--no-slab", action="store_true",
                    help="skip the LJTS slab reported beside a multi-site headline")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--precision-report", action="store_true",
                    help="only report the FP32 form's errors against FP64 (one JSON line)")
    ap.add_argument("--drift-steps", type=int, default=10_000)
    ap.add_argument("--replicas", type=int, default=512,
                    help="replica count of the batched-replicas workload")
    ap.add_argument("--plain-steps", type=int, default=10,
                    help="steps of the per-replica plain-launch comparison")
    ap.add_argument("--pair-shape", default=None, metavar="CPB,BPX",
                    help="fixed pair-kernel launch shape instead of the autotuner (for profiling)")
    args = ap.parse_args()
    if args.pair_shape:
        global PAIR_SHAPE
        PAIR_SHAPE = tuple(int(x) for x in args.pair_shape.split(","))
    if args.precision_report:
        print(json.dumps({"precision_report": precision_report(args.scale, args.drift_steps)}),
              flush=True)
        return
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "copies": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path; run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    if args.workload == "replicas":
        print(json.dumps(run_replicas(args)), flush=True)
        return
    if args.workload == "step-variants":
        print(json.dumps(run_step_variants(args)), flush=True)
        return

    line = measure(args.workload, args.scale, args)
    if args.workload != "ljts_slab" and not args.no_slab:
        # the one-centre slab (the LJ kernel's workload) beside the headline
        slab = measure("ljts_slab", 1.0, args, quiet=True)
        line["ljts_slab"] = {k: slab[k] for k in (
            "value", "ms_per_step", "roofline", "kernels", "fp64_plain", "speedup_fp32_vs_fp64",
            "force_rel_rms_err_fp32_vs_fp64", "config", "launch_shape", "reduced_memory_mode",
            "molecule_checks") if k in slab}
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, OpenMP)")
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_budget)
    print(json.dumps(line), flush=True)


def reference_checks(s):
    """Molecule-centre distance checks of the reference's C08 traversal on the
    full workload (halo pairs of periodic boundaries counted from both sides,
    as ls1's FlopCounter does) and of the pairs with macroscopic sums (each
    unordered pair once), from the oracle's counting walk."""
    from oracle_lib import Oracle

    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    return o.count_checks(s.r, s.q, s.cid)


def measure(workload, scale, args, quiet=False):
    log(f"generating workload {workload} scale {scale}")
    s = make_workload(workload, scale)
    log(f"  N = {s.n}, box = {s.box.tolist()}")

    res = {}
    forms = {"fp32": PREC_FP32_FP64ACC, "fp64": PREC_FP64}
    for tag in (["fp32", "fp64"] if args.only == "both" else [args.only]):
        log(f"GPU run {tag}: warmup {args.warmup}, timed {args.steps} steps")
        res[tag] = run_gpu(s, forms[tag], args.steps, args.warmup, args.kernel_reps,
                           rmm=workload in ("ljts_slab", "lj_fcc"))
        r = res[tag]
        log(f"  {r['seconds'] * 1e3 / args.steps:.3f} ms/step, pair kernels {r['pair_ms']:.3f} ms "
            f"(per kernel {np.round(r['kernel_ms'], 3).tolist()})")

    main_tag = "fp32" if "fp32" in res else "fp64"
    r = res[main_tag]
    ms = r["seconds"] * 1e3 / args.steps
    value = s.n * args.steps / r["seconds"]
    log("reference-equivalent molecule-centre checks (oracle counting walk)")
    ref_all, ref_macro = reference_checks(s)
    kern = {tag: kernel_rooflines(res[tag], tag, workload, scale, ref_macro) for tag in res}
    top = max(kern[main_tag], key=lambda k: k["kernel_ms"])
    all_flops = float(8.0 * ref_macro + (FLOPS_PER_PAIR[1:] * r["counters"][1:]).sum())
    potential = {"ljts_slab": "LJTS 12-6", "mixture": "2CLJQ + 4-site water (LJ, charge, "
                 "quadrupole)", "lj_fcc": "LJ 12-6", "binary_lj": "binary LJ 12-6"}.get(workload,
                                                                                      workload)
    out = {
        "metric": "molecule-steps/s on 1 MI355X: FP32/FP64-acc vs FP64, at stated force error",
        "value": value,
        "unit": "molecule-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 accumulation/sums/Leapfrog" if main_tag == "fp32" else "f64",
        "data": "synthetic (seeded lattice / random sequential addition, random orientations, "
                "Maxwell-Boltzmann velocities)",
        "config": {"workload": f"{workload}" + ("" if scale == 1.0 else f"@scale{scale}"),
                   "molecules": int(s.n), "box": s.box.tolist(), "rc": s.rc, "dt": s.dt,
                   "potential": potential, "parallelism": "single GPU"},
        "roofline": top,
        "kernels": kern,
        "pair_kernels_ms": r["pair_ms"],
        "pair_kernels_flops_per_launch": all_flops,
        "pair_kernels_achieved_tflops": all_flops / (r["pair_ms"] * 1e-3) / 1e12,
        "launch_shape": {"cells_per_block": r["shape"][0], "blocks_per_xcd": r["shape"][1]},
    }
    if "fp64" in res and "fp32" in res:
        r64 = res["fp64"]
        v64 = s.n * args.steps / r64["seconds"]
        F32, F64 = res["fp32"]["init"]["F"], r64["init"]["F"]
        err = float(np.sqrt(((F32 - F64) ** 2).sum(1).mean() / (F64 ** 2).sum(1).mean()))
        out["fp64_plain"] = {
            "value": v64, "ms_per_step": r64["seconds"] * 1e3 / args.steps,
            "pair_kernels_ms": r64["pair_ms"],
            "roofline": max(kern["fp64"], key=lambda k: k["kernel_ms"])}
        out["speedup_fp32_vs_fp64"] = value / v64
        out["force_rel_rms_err_fp32_vs_fp64"] = err
        M32, M64 = res["fp32"]["init"]["M"], r64["init"]["M"]
        if np.abs(M64).max() > 0:
            out["torque_rel_rms_err_fp32_vs_fp64"] = float(np.sqrt(
                ((M32 - M64) ** 2).sum(1).mean() / (M64 ** 2).sum(1).mean()))
        u32, u64 = res["fp32"]["per_step"][:, 0], r64["per_step"][:, 0]
        out["upot_rel_err_first_step"] = float(abs(u32[0] - u64[0]) / abs(u64[0]))
    for tag in res:
        if res[tag].get("rmm_seconds"):
            rs = res[tag]["rmm_seconds"]
            out.setdefault("reduced_memory_mode", {})[tag] = {
                "value": s.n * args.steps / rs, "ms_per_step": rs * 1e3 / args.steps,
                "note": "VCP1CLJRMM + LeapfrogRMM equivalent: kick inside the pair kernel"}
    out["molecule_checks"] = {
        "this_build": float(r["counters"][0]),
        "reference_c08_all": float(ref_all),
        "reference_c08_with_macroscopic_sums": float(ref_macro),
        "note": "flops use ls1's checks with macroscopic sums (each unordered molecule pair "
                "once); this build checks molecule centres once per pair of sites of the "
                "kernel's types (charge partners once per molecule), ls1's FlopCounter also "
                "counts the second visit of boundary-halo pairs (reference_c08_all)"}
    return out


if __name__ == "__main__":
    main()
