"""Writes tests/golden/ref_golden.npz from the reference's own code.

Runs oracle/_ref/libls1ref.so -- ls1-MarDyn's FullMolecule / Component /
Comp2Param / Quaternion sources and the potforce.h Legacy pair physics,
compiled from the reference tree by ``make -C oracle ref`` -- on:

* every force fixture the reference's VectorizedCellProcessorTest uses
  (tests/golden/inp/*.inp, copied data files), at the cutoffs that test uses
  (VectorizedCellProcessorTest.cpp:45-134, 136-388), open boundaries;
* two small seeded periodic systems (one-centre LJ, and the 2CLJQ + 4-site
  water mixture), minimum image;
* one Leapfrog half-step pair (upd_preF, upd_postF) on the mixture;
* component masses, moments of inertia and the LJ parameter tables.

Usage: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1] / "ls1-mardyn_amd"))

from ls1inp import read_inp  # noqa: E402
from oracle_lib import RefLegacy  # noqa: E402
import scenarios  # noqa: E402

FIXTURES = [
    ("ForceCalculationTestU0", 1.1),
    ("ForceCalculationTestF0", 1.3),
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


def nlj_total(desc, ncomp):
    p, n = 0, 0
    for _ in range(ncomp):
        nlj, nc, nd, nq = (int(x) for x in desc[p:p + 4])
        n += nlj
        p += 4 + 8 * nlj + 5 * nc + 7 * nd + 7 * nq + 3
    return n


def main():
    out = {}
    for name, rc in FIXTURES:
        ps = read_inp(HERE / "inp" / f"{name}.inp")
        ref = RefLegacy.from_phase_space(ps, rc)
        res = ref.forces(ps.r, ps.q, ps.cid, ps.box, periodic=False)
        for k in ("F", "M", "Vi"):
            out[f"{name}/{k}"] = res[k]
        out[f"{name}/upot_virial"] = np.array([res["upot"], res["virial"]])
        m, I, rot = ref.component_props()
        out[f"{name}/mass"], out[f"{name}/inertia"], out[f"{name}/rotdof"] = m, I, rot
        n = nlj_total(ps.comp_desc, ps.ncomp)
        e, s, sh = ref.lj_tables(n)
        out[f"{name}/eps24"], out[f"{name}/sig2"], out[f"{name}/shift6"] = e, s, sh

    for tag, sysd in (("periodic_lj", scenarios.small_lj(seed=11)),
                      ("periodic_mix", scenarios.small_mixture(seed=12))):
        ref = RefLegacy(sysd.ncomp, sysd.comp_desc, sysd.mix, sysd.eps_rf, sysd.rc, sysd.rc)
        res = ref.forces(sysd.r, sysd.q, sysd.cid, sysd.box, periodic=True)
        out[f"{tag}/r"], out[f"{tag}/q"], out[f"{tag}/cid"] = sysd.r, sysd.q, sysd.cid
        out[f"{tag}/box"], out[f"{tag}/comp_desc"] = sysd.box, sysd.comp_desc
        out[f"{tag}/rc_epsrf"] = np.array([sysd.rc, sysd.eps_rf])
        for k in ("F", "M", "Vi"):
            out[f"{tag}/{k}"] = res[k]
        out[f"{tag}/upot_virial"] = np.array([res["upot"], res["virial"]])

    sysd = scenarios.small_mixture(seed=13)
    rng = np.random.default_rng(14)
    n = len(sysd.r)
    F = rng.normal(size=(n, 3)) * 5.0
    M = rng.normal(size=(n, 3)) * 0.5
    L = rng.normal(size=(n, 3)) * 0.05
    ref = RefLegacy(sysd.ncomp, sysd.comp_desc, sysd.mix, sysd.eps_rf, sysd.rc, sysd.rc)
    dt = 0.002
    r1, v1, q1, L1 = ref.upd_preF(dt, sysd.r, sysd.v, sysd.q, L, F, M, sysd.cid)
    v2, L2, sums = ref.upd_postF(dt, v1, L1, q1, F, M, sysd.cid)
    out.update({"leapfrog/r0": sysd.r, "leapfrog/v0": sysd.v, "leapfrog/q0": sysd.q,
                "leapfrog/L0": L, "leapfrog/F": F, "leapfrog/M": M, "leapfrog/cid": sysd.cid,
                "leapfrog/comp_desc": sysd.comp_desc, "leapfrog/r1": r1, "leapfrog/v1": v1,
                "leapfrog/q1": q1, "leapfrog/L1": L1, "leapfrog/v2": v2, "leapfrog/L2": L2,
                "leapfrog/sums": sums})
    np.savez_compressed(HERE / "ref_golden.npz", **out)
    print("wrote", HERE / "ref_golden.npz", len(out), "arrays")


if __name__ == "__main__":
    main()
