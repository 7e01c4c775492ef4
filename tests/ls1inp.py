"""Reader for ls1-MarDyn ASCII phase-space files (``.inp``).

Mirrors ``ASCIIReader::readPhaseSpaceHeader`` / ``readPhaseSpace``
(src/io/ASCIIReader.cpp:52-400): whitespace-separated tokens, ``#`` comment
lines, the component block of ``NumberOfComponents`` (per component
``nlj nc nd nq ntersoff``, then LJ ``x y z m eps sigma rc_shift do_shift``,
charges ``x y z m q``, dipoles ``x y z ex ey ez absMy``, quadrupoles
``x y z ex ey ez absQ``, then three moments of inertia), the mixing pairs
``xi eta`` and epsilon_RF, then ``NumberOfMolecules`` and rows in the
``ICRVQD`` / ``ICRV`` / ``IRV`` formats (component ids are 1-based in the
file, 0-based here).

The component block is returned flattened to the record layout the C ABI
and the oracle take (see include/ls1amd.h).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class PhaseSpace:
    box: np.ndarray            # (3,)
    temperature: float
    ncomp: int
    comp_desc: np.ndarray      # flat float64 component records
    mix: np.ndarray            # (npairs, 2) xi, eta
    eps_rf: float
    ids: np.ndarray            # (N,) uint64
    cid: np.ndarray            # (N,) int32, 0-based
    r: np.ndarray              # (N, 3)
    v: np.ndarray              # (N, 3)
    q: np.ndarray              # (N, 4) qw qx qy qz
    L: np.ndarray              # (N, 3) angular momentum D


def _tokens(text: str):
    for line in text.splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        for t in s.split():
            yield t


def read_inp(path) -> PhaseSpace:
    with open(path, "r") as fh:
        text = fh.read()
    it = _tokens(text)
    head = [next(it), next(it), next(it)]
    if head[0] != "mardyn" or head[1] != "trunk":
        raise ValueError(f"{path}: not an ls1 phase-space file")
    box = np.zeros(3)
    temperature = 0.0
    ncomp = 0
    desc: list[float] = []
    mix = []
    eps_rf = 1.0
    nmol = None
    fmt = "ICRVQD"
    while True:
        tok = next(it)
        if tok in ("currentTime", "t"):
            next(it)
        elif tok in ("Temperature", "T"):
            temperature = float(next(it))
        elif tok in ("Length", "L"):
            box = np.array([float(next(it)) for _ in range(3)])
        elif tok in ("NumberOfComponents", "C"):
            ncomp = int(next(it))
            for _ in range(ncomp):
                nlj, nc, nd, nq, nt = (int(next(it)) for _ in range(5))
                if nt != 0:
                    raise ValueError("tersoff sites are not supported")
                desc += [nlj, nc, nd, nq]
                for _ in range(nlj):
                    desc += [float(next(it)) for _ in range(8)]
                for _ in range(nc):
                    desc += [float(next(it)) for _ in range(5)]
                for _ in range(nd):
                    desc += [float(next(it)) for _ in range(7)]
                for _ in range(nq):
                    desc += [float(next(it)) for _ in range(7)]
                desc += [float(next(it)) for _ in range(3)]
            for _ in range(ncomp * (ncomp - 1) // 2):
                mix.append((float(next(it)), float(next(it))))
            eps_rf = float(next(it))
        elif tok in ("NumberOfMolecules", "N"):
            nmol = int(next(it))
        elif tok in ("MoleculeFormat", "M"):
            fmt = next(it)
            break
        else:
            raise ValueError(f"{path}: unexpected token {tok!r}")
    if nmol is None:
        raise ValueError(f"{path}: NumberOfMolecules missing")
    ids = np.zeros(nmol, np.uint64)
    cid = np.zeros(nmol, np.int32)
    r = np.zeros((nmol, 3))
    v = np.zeros((nmol, 3))
    q = np.zeros((nmol, 4))
    q[:, 0] = 1.0
    L = np.zeros((nmol, 3))
    for i in range(nmol):
        ids[i] = int(next(it))
        if fmt != "IRV":
            cid[i] = int(next(it)) - 1
        r[i] = [float(next(it)) for _ in range(3)]
        v[i] = [float(next(it)) for _ in range(3)]
        if fmt in ("ICRVQD", "ICRVQDV"):
            q[i] = [float(next(it)) for _ in range(4)]
            L[i] = [float(next(it)) for _ in range(3)]
        if fmt == "ICRVQDV":
            for _ in range(3):
                next(it)
    return PhaseSpace(box, temperature, ncomp, np.asarray(desc, np.float64),
                      np.asarray(mix, np.float64).reshape(-1, 2), eps_rf,
                      ids, cid, r, v, q, L)


def write_inp(path, box, ncomp, comp_desc, mix, eps_rf, ids, cid, r, v, q, L, temperature=0.0):
    """Writes the ICRVQD format ASCIIReader reads (1-based component ids)."""
    desc = [float(x) for x in comp_desc]
    lines = ["mardyn trunk 20100525", "currentTime\t0",
             " Length\t%.17g %.17g %.17g" % tuple(box), f" Temperature\t{temperature!r}",
             f" NumberOfComponents\t{ncomp}"]
    p = 0
    for _ in range(ncomp):
        n = [int(x) for x in desc[p:p + 4]]
        lines.append("\t".join(str(k) for k in n) + "\t0")
        p += 4
        for width, count in zip((8, 5, 7, 7), n):
            for _ in range(count):
                lines.append(" ".join("%.17g" % x for x in desc[p:p + width]))
                p += width
        lines.append(" ".join("%.17g" % x for x in desc[p:p + 3]))
        p += 3
    for xi, eta in np.asarray(mix, np.float64).reshape(-1, 2):
        lines.append("%.17g %.17g" % (xi, eta))
    lines.append("%.17g" % eps_rf)
    lines.append(f" NumberOfMolecules\t{len(r)}")
    lines.append(" MoleculeFormat\tICRVQD")
    for i in range(len(r)):
        vals = list(r[i]) + list(v[i]) + list(q[i]) + list(L[i])
        lines.append(f"{int(ids[i])}\t{int(cid[i]) + 1}\t" + "\t".join("%.17g" % x for x in vals))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
