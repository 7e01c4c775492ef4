#!/usr/bin/env python3
"""Benchmark: molecule-steps/s of the short-range force path on one MI355X.

Workload (default): truncated-shifted LJ vapour-liquid slab, 640 x 160 x 160
box with a 200-wide liquid slab at rho 0.73 and vapour at rho 0.02, T 0.8,
rc 2.5, Leapfrog dt 0.002 -- 3.96 M molecules (the 4 M one-GPU subdomain
share; the literal 320 x 80 x 80 box cannot hold 4 M molecules at these
densities). One step = Leapfrog half kick + drift + periodic wrap,
counting-sort re-bin, SoA site build, force traversal, calcFM, Upot/virial
sums, Leapfrog half kick -- all on the device, one hipGraph per step, inputs
resident in HBM before timing starts.

Both precision forms run: the FP32-pair / FP64-accumulate form gives `value`
(the metric's chosen form); the plain FP64 form and the FP32 form's force
error against it are reported beside it.

The CPU baseline is the C oracle (a scalar port of the reference algorithm)
timed on this host on a bounded sample of the same workload.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, scale, tag):
    """HBM bytes per pair-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (copy-calibrated), or
    None when no summary covers it."""
    if workload != "ljts_slab" or scale != 1.0 or tag != "fp32":
        return None
    f = ROOT / "profiles" / "k_pairs_lj_ljts_slab_3.96M_traffic.json"
    try:
        return json.loads(f.read_text())["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


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


def run_gpu(s, precision, steps, warmup, kernel_reps):
    with Context(0) as ctx:
        ctx.setup_system(s, periodic=True, precision=precision)
        ctx.update_container()
        ctx.traverse()
        init = ctx.download()
        # work counters of one traversal (separate, uncounted runs are timed)
        ctx.set_counting(True)
        ctx.traverse()
        counters = ctx.counters()
        ctx.set_counting(False)
        # pair-kernel duration: HIP events on the context stream around the
        # pair kernels, averaged over kernel_reps traversals
        ctx.set_timing(True)
        pair_ms = []
        for _ in range(kernel_reps):
            ctx.traverse()
            pair_ms.append(ctx.last_kernel_ms()[0])
        ctx.set_timing(False)
        if warmup:
            ctx.step(warmup, s.dt, use_graph=True)
        ctx.synchronize()
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
    return dict(seconds=t1 - t0, per_step=per, counters=counters, pair_ms=float(np.mean(pair_ms)),
                pair_ms_all=pair_ms, init=init)


def cpu_baseline(workload, budget_s):
    """Oracle (scalar C port, 1 thread) on a small sample of the workload:
    same physics and densities, box scaled down."""
    from oracle_lib import Oracle

    if workload == "ljts_slab":
        s = scenarios.ljts_slab(scale=0.25)
        sample = "LJTS slab scaled to 160x40x40 (61,920 molecules), same densities/T/rc"
    elif workload == "mixture":
        s = scenarios.mixture(n=20_000)
        sample = "multi-site mixture, 20,000 molecules"
    else:
        s = scenarios.lj_fcc(ncell=10)
        sample = "LJ FCC 10^3 cells (4,000 molecules)"
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    t0 = time.perf_counter()
    o.forces(s.r, s.q, s.cid)
    one = time.perf_counter() - t0
    nsteps = max(1, int(budget_s / max(one, 1e-3)))
    t0 = time.perf_counter()
    o.simulate(nsteps, s.dt, s.r, s.v, s.q, s.L, s.cid)
    dt = time.perf_counter() - t0
    return dict(value=s.n * nsteps / dt, unit="molecule-steps/s", cores=1, kind="port",
                sample=f"{sample}, {nsteps} steps, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="ljts_slab")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
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
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    log(f"generating workload {args.workload} scale {args.scale}")
    s = make_workload(args.workload, args.scale)
    log(f"  N = {s.n}, box = {s.box.tolist()}")

    res = {}
    forms = {"fp32": PREC_FP32_FP64ACC, "fp64": PREC_FP64}
    for tag in (["fp32", "fp64"] if args.only == "both" else [args.only]):
        log(f"GPU run {tag}: warmup {args.warmup}, timed {args.steps} steps")
        res[tag] = run_gpu(s, forms[tag], args.steps, args.warmup, args.kernel_reps)
        r = res[tag]
        log(f"  {r['seconds'] * 1e3 / args.steps:.3f} ms/step, pair kernel {r['pair_ms']:.3f} ms")

    main_tag = "fp32" if "fp32" in res else "fp64"
    r = res[main_tag]
    ms = r["seconds"] * 1e3 / args.steps
    value = s.n * args.steps / r["seconds"]

    # algorithmic flops of one traversal (each unordered pair once)
    flops = float((FLOPS_PER_PAIR * r["counters"]).sum())
    achieved = flops / (r["pair_ms"] * 1e-3) / 1e12
    peak = PEAK_FP32_VECTOR_TF if main_tag == "fp32" else PEAK_FP64_VECTOR_TF
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
        "data": "synthetic (seeded random sequential addition, Maxwell-Boltzmann velocities)",
        "config": {"workload": f"{args.workload}" + ("" if args.scale == 1.0 else f"@scale{args.scale}"),
                   "molecules": int(s.n), "box": s.box.tolist(), "rc": s.rc, "dt": s.dt,
                   "potential": "LJTS 12-6" if args.workload == "ljts_slab" else args.workload,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": measured_traffic(args.workload, args.scale, main_tag),
                     "kernel": "k_pairs<LJ>", "flops_per_launch": flops,
                     "kernel_ms": r["pair_ms"]},
    }
    if "fp64" in res and "fp32" in res:
        r64 = res["fp64"]
        v64 = s.n * args.steps / r64["seconds"]
        F32, F64 = res["fp32"]["init"]["F"], r64["init"]["F"]
        err = float(np.sqrt(((F32 - F64) ** 2).sum(1).mean() / (F64 ** 2).sum(1).mean()))
        fl64 = float((FLOPS_PER_PAIR * r64["counters"]).sum())
        out["fp64_plain"] = {
            "value": v64, "ms_per_step": r64["seconds"] * 1e3 / args.steps,
            "pair_kernel_ms": r64["pair_ms"],
            "roofline": {"bound": "valu", "achieved": fl64 / (r64["pair_ms"] * 1e-3) / 1e12,
                         "peak": PEAK_FP64_VECTOR_TF, "unit": "TFLOP/s",
                         "frac": fl64 / (r64["pair_ms"] * 1e-3) / 1e12 / PEAK_FP64_VECTOR_TF}}
        out["speedup_fp32_vs_fp64"] = value / v64
        out["force_rel_rms_err_fp32_vs_fp64"] = err
        u32, u64 = res["fp32"]["per_step"][:, 0], r64["per_step"][:, 0]
        out["upot_rel_err_first_step"] = float(abs(u32[0] - u64[0]) / abs(u64[0]))
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, 1 thread)")
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
