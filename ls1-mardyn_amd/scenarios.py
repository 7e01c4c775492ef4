"""Seeded workloads for the benchmark and the tests.

Component records use the ls1 ASCII-header layout (src/io/ASCIIReader.cpp:
152-232) that the C ABI takes: per component ``nlj nc nd nq``, LJ sites
``x y z m eps sigma rc_shift do_shift``, charges ``x y z m q``, dipoles and
quadrupoles ``x y z ex ey ez abs``, then ``I11 I22 I33`` (0 = computed from
the sites). Sites are given in the principal-axes / centre-of-mass frame, as
ls1 assumes (src/molecules/Component.cpp:151-179).

Workloads (reduced units, sigma = eps = 1):

* ``lj_fcc``       one-centre LJ 12-6, FCC lattice, rho 0.6, T 0.7, rc 2.5
                   (32,000 molecules at the default 20^3 unit cells).
* ``ljts_slab``    truncated-shifted LJ vapour-liquid slab: liquid slab at
                   rho 0.73 between vapour at rho 0.02, T 0.8, rc 2.5, random
                   sequential addition with minimum separation 0.8. The
                   default size is a 640 x 160 x 160 box with a 200-wide slab
                   (3.96 M molecules); ``scale`` shrinks all lengths.
* ``mixture``      rigid 2CLJQ (two LJ sites 0.5 apart, quadrupole Q = 2 on
                   the axis) and 4-site water-like molecules (LJ oxygen, two
                   +0.52 hydrogens, a -1.04 M site), 50/50, rho 0.5, T 1.0,
                   rc 3, random orientations, FCC lattice start.
* ``binary_lj``    60/40 one-centre LJ mixture, eps22 = 0.85, sigma22 = 1.1,
                   seeded Lorentz-Berthelot factors xi, eta in [0.95, 1.05],
                   rho 0.7, T 0.9, rc 2.5.
* ``replicas``     a state-point sweep of independent periodic LJTS systems
                   (rc 2.5): molecule counts log-uniform in [2,000, 40,000],
                   rho uniform in [0.05, 0.85], T uniform in [0.7, 1.3], each
                   drawn from the seed; 512 replicas by default.
* ``small_lj`` / ``small_mixture``: seconds-sized periodic systems for parity.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
GEN_SO = HERE / "host" / "libls1gen.so"


@dataclasses.dataclass
class System:
    name: str
    ncomp: int
    comp_desc: np.ndarray
    mix: np.ndarray
    eps_rf: float
    rc: float
    rc_lj: float
    box: np.ndarray
    r: np.ndarray
    v: np.ndarray
    q: np.ndarray
    L: np.ndarray
    cid: np.ndarray
    temperature: float
    dt: float = 0.002

    @property
    def n(self):
        return len(self.r)


# ------------------------------------------------------------- components

def comp_1clj(eps=1.0, sigma=1.0, m=1.0, rc_shift=2.5, shifted=False):
    return [1, 0, 0, 0, 0.0, 0.0, 0.0, m, eps, sigma, rc_shift, 1.0 if shifted else 0.0,
            0.0, 0.0, 0.0]


def comp_2cljq(rc_shift=3.0, Q=2.0, half_len=0.25):
    return [2, 0, 0, 1,
            0.0, 0.0, half_len, 0.5, 1.0, 1.0, rc_shift, 0.0,
            0.0, 0.0, -half_len, 0.5, 1.0, 1.0, rc_shift, 0.0,
            0.0, 0.0, 0.0, 0.0, 0.0, 1.0, Q,
            0.0, 0.0, 0.0]


def comp_water(rc_shift=3.0, d_oh=0.3, angle_deg=104.5, d_om=0.05, qh=0.52):
    """Planar C2v molecule in the x-z plane, symmetry axis z, origin at the
    centre of mass. Masses O:H:H = 16:1:1, normalised to 1; M site massless."""
    th = math.radians(angle_deg) / 2.0
    mo, mh = 16.0 / 18.0, 1.0 / 18.0
    zh_rel = d_oh * math.cos(th)
    zo = -(2.0 * mh * zh_rel) / (mo + 2.0 * mh)
    zh = zo + zh_rel
    xh = d_oh * math.sin(th)
    return [1, 3, 0, 0,
            0.0, 0.0, zo, mo, 1.0, 1.0, rc_shift, 0.0,
            xh, 0.0, zh, mh, qh,
            -xh, 0.0, zh, mh, qh,
            0.0, 0.0, zo + d_om, 0.0, -2.0 * qh,
            0.0, 0.0, 0.0]


def parse_components(desc):
    """Per component: dict(nlj, nc, nd, nq, mass, I, sites...)."""
    out, p = [], 0
    desc = list(desc)
    while p < len(desc):
        nlj, nc, nd, nq = (int(x) for x in desc[p:p + 4])
        p += 4
        lj = [desc[p + 8 * i:p + 8 * i + 8] for i in range(nlj)]
        p += 8 * nlj
        ch = [desc[p + 5 * i:p + 5 * i + 5] for i in range(nc)]
        p += 5 * nc
        dp = [desc[p + 7 * i:p + 7 * i + 7] for i in range(nd)]
        p += 7 * nd
        qp = [desc[p + 7 * i:p + 7 * i + 7] for i in range(nq)]
        p += 7 * nq
        Iov = desc[p:p + 3]
        p += 3
        mass, I = 0.0, np.zeros(3)
        for s in lj + ch:
            x, y, z, m = s[:4]
            mass += m
            I += m * np.array([y * y + z * z, x * x + z * z, x * x + y * y])
        I = np.array([Iov[d] if Iov[d] > 0 else I[d] for d in range(3)])
        out.append(dict(nlj=nlj, nc=nc, nd=nd, nq=nq, mass=mass, I=I, lj=lj, ch=ch, dp=dp, qp=qp))
    return out


# -------------------------------------------------------------- helpers

def _gen_lib():
    if not GEN_SO.exists():
        raise RuntimeError(f"{GEN_SO} missing: run __graft_entry__.build()")
    L = C.CDLL(str(GEN_SO))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    L.ls1gen_rsa.restype = C.c_long
    L.ls1gen_rsa.argtypes = [C.c_uint64, C.c_int, dp, dp, lp, dp, C.c_double, C.c_long, dp]
    return L


def rsa_positions(seed, regions, box, min_sep, max_trials=2000):
    """regions: list of (lo[3], hi[3], count). Random sequential addition."""
    lo = np.ascontiguousarray([r[0] for r in regions], np.float64).ravel()
    hi = np.ascontiguousarray([r[1] for r in regions], np.float64).ravel()
    cnt = np.ascontiguousarray([r[2] for r in regions], np.int64)
    out = np.zeros((int(cnt.sum()), 3))
    got = _gen_lib().ls1gen_rsa(int(seed), len(regions), lo, hi, cnt,
                                np.ascontiguousarray(box, np.float64), float(min_sep),
                                int(max_trials), out)
    if got != out.shape[0]:
        raise RuntimeError(f"random sequential addition placed {got} of {out.shape[0]}")
    return out


def random_quaternions(rng, n):
    """Uniform random unit quaternions (Shoemake's subgroup algorithm)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2),
                  a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3)], axis=1)
    return q


def quat_rotate(q, d):
    """Quaternion::rotate (src/molecules/Quaternion.cpp:43-64), vectorised."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    ww, xx, yy, zz = w * w, x * x, y * y, z * z
    wx, wy, wz, xy, xz, yz = w * x, w * y, w * z, x * y, x * z, y * z
    return np.stack([
        (ww + xx - yy - zz) * d[:, 0] + 2 * (xy - wz) * d[:, 1] + 2 * (wy + xz) * d[:, 2],
        2 * (wz + xy) * d[:, 0] + (ww - xx + yy - zz) * d[:, 1] + 2 * (yz - wx) * d[:, 2],
        2 * (xz - wy) * d[:, 0] + 2 * (wx + yz) * d[:, 1] + (ww - xx - yy + zz) * d[:, 2]], axis=1)


def maxwell_boltzmann(rng, cid, comps, T, q=None):
    """Translational velocities and lab-frame angular momenta at temperature
    T; net momentum removed."""
    n = len(cid)
    mass = np.array([c["mass"] for c in comps])[cid]
    v = rng.normal(size=(n, 3)) * np.sqrt(T / mass)[:, None]
    v -= (mass[:, None] * v).sum(0) / mass.sum()
    Ib = np.array([c["I"] for c in comps])[cid]
    w = rng.normal(size=(n, 3)) * np.sqrt(T / np.where(Ib > 0, Ib, 1.0))
    Lb = np.where(Ib > 0, Ib * w, 0.0)
    L = quat_rotate(q, Lb) if q is not None else Lb
    return v, L


def fcc_lattice(ncell, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(ncell)] * 3, indexing="ij"), -1).reshape(-1, 3)
    r = (g[:, None, :] + base[None, :, :]).reshape(-1, 3) * a + 0.25 * a
    return r


# ------------------------------------------------------------- workloads

def lj_fcc(ncell=20, rho=0.6, T=0.7, rc=2.5, seed=1, shifted=False):
    a = (4.0 / rho) ** (1.0 / 3.0)
    r = fcc_lattice(ncell, a)
    n = len(r)
    desc = np.array(comp_1clj(rc_shift=rc, shifted=shifted), np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed)
    cid = np.zeros(n, np.int32)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System("lj_fcc", 1, desc, np.zeros((0, 2)), 1.0, rc, rc, np.full(3, ncell * a), r, v, q,
                  L * 0.0, cid, T)


def ljts_slab(scale=1.0, rho_l=0.73, rho_v=0.02, T=0.8, rc=2.5, seed=2, min_sep=0.8):
    box = np.array([640.0, 160.0, 160.0]) * scale
    w = 200.0 * scale
    x0, x1 = 0.5 * (box[0] - w), 0.5 * (box[0] + w)
    n_l = int(round(rho_l * w * box[1] * box[2]))
    n_v1 = int(round(rho_v * x0 * box[1] * box[2]))
    n_v2 = int(round(rho_v * (box[0] - x1) * box[1] * box[2]))
    r = rsa_positions(seed, [((x0, 0, 0), (x1, box[1], box[2]), n_l),
                             ((0, 0, 0), (x0, box[1], box[2]), n_v1),
                             ((x1, 0, 0), (box[0], box[1], box[2]), n_v2)], box, min_sep)
    n = len(r)
    desc = np.array(comp_1clj(rc_shift=rc, shifted=True), np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed + 1000)
    cid = np.zeros(n, np.int32)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System("ljts_slab", 1, desc, np.zeros((0, 2)), 1.0, rc, rc, box, r, v, q, L * 0.0, cid,
                  T)


def mixture(n=1_000_000, rho=0.5, T=1.0, rc=3.0, seed=3):
    """Molecules start on an FCC lattice at density rho (k^3 unit cells, the
    smallest k with 4 k^3 >= n, random vacancies removed down to n): random
    placement cannot avoid site overlaps of extended molecules at rho = 0.5,
    and the overlaps blow the system up within a few steps."""
    rng = np.random.default_rng(seed + 1000)
    k = int(math.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-9))
    Lb = (n / rho) ** (1.0 / 3.0)
    box = np.full(3, Lb)
    r = fcc_lattice(k, Lb / k)
    r = r[np.sort(rng.permutation(len(r))[:n])]
    desc = np.array(comp_2cljq(rc) + comp_water(rc), np.float64)
    comps = parse_components(desc)
    cid = np.zeros(n, np.int32)
    cid[rng.permutation(n)[: n // 2]] = 1
    q = random_quaternions(rng, n)
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System("mixture", 2, desc, np.array([[1.0, 1.0]]), 1.0e10, rc, rc, box, r, v, q, L, cid,
                  T)


def binary_lj(n=500_000, rho=0.7, T=0.9, rc=2.5, seed=4, min_sep=0.85):
    Lb = (n / rho) ** (1.0 / 3.0)
    box = np.full(3, Lb)
    r = rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, min_sep)
    desc = np.array(comp_1clj(1.0, 1.0, 1.0, rc) + comp_1clj(0.85, 1.1, 1.0, rc), np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed + 1000)
    xi, eta = rng.uniform(0.95, 1.05, size=2)
    cid = np.zeros(n, np.int32)
    cid[rng.permutation(n)[: int(round(0.4 * n))]] = 1
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System("binary_lj", 2, desc, np.array([[xi, eta]]), 1.0, rc, rc, box, r, v, q, L * 0.0,
                  cid, T)


def ljts_box(n, rho, T, rc=2.5, seed=0, min_sep=0.8, name="ljts_box"):
    """One periodic truncated-shifted LJ system at density rho, temperature T."""
    box = np.full(3, (n / rho) ** (1.0 / 3.0))
    r = rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, min_sep)
    desc = np.array(comp_1clj(rc_shift=rc, shifted=True), np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed + 1000)
    cid = np.zeros(n, np.int32)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System(name, 1, desc, np.zeros((0, 2)), 1.0, rc, rc, box, r, v, q, L * 0.0, cid, T)


def replicas(nrep=512, nmin=2000, nmax=40000, seed=5, rc=2.5):
    """Independent LJTS systems of a state-point sweep (see the module notes)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nrep)).astype(np.int64)
    rhos = rng.uniform(0.05, 0.85, nrep)
    Ts = rng.uniform(0.7, 1.3, nrep)
    return [ljts_box(int(ns[k]), float(rhos[k]), float(Ts[k]), rc, seed=10_000 + 7 * k + seed,
                     name=f"replica{k}") for k in range(nrep)]


def small_lj(n=400, rho=0.6, T=0.7, rc=2.5, seed=11, shifted=False):
    Lb = (n / rho) ** (1.0 / 3.0)
    box = np.full(3, Lb)
    r = rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.85)
    desc = np.array(comp_1clj(rc_shift=rc, shifted=shifted), np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed + 1000)
    cid = np.zeros(n, np.int32)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    return System("small_lj", 1, desc, np.zeros((0, 2)), 1.0, rc, rc, box, r, v, q, L * 0.0, cid, T)


def small_mixture(n=150, rho=0.3, T=1.0, rc=2.5, seed=12, with_dipoles=True):
    """2CLJQ + water (+ a Stockmayer-like dipolar component when
    with_dipoles), so every site-pair kernel type occurs."""
    Lb = (n / rho) ** (1.0 / 3.0)
    box = np.full(3, Lb)
    r = rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.9)
    parts = comp_2cljq(rc) + comp_water(rc)
    ncomp = 2
    if with_dipoles:
        # one LJ site, a point dipole 0.1 off-centre and a charge; moments set
        parts += [1, 1, 1, 0,
                  0.0, 0.0, 0.0, 1.0, 1.0, 1.0, rc, 1.0,
                  0.0, 0.0, -0.1, 0.0, 0.3,
                  0.0, 0.0, 0.1, 0.0, 0.0, 1.0, 1.2,
                  0.02, 0.02, 0.01]
        ncomp = 3
    desc = np.array(parts, np.float64)
    comps = parse_components(desc)
    rng = np.random.default_rng(seed + 1000)
    cid = rng.integers(0, ncomp, size=n).astype(np.int32)
    q = random_quaternions(rng, n)
    v, L = maxwell_boltzmann(rng, cid, comps, T, q)
    mix = np.array([[1.0, 1.0]] * (ncomp * (ncomp - 1) // 2))
    if ncomp == 3:
        mix[1] = (0.97, 1.02)
    return System("small_mixture", ncomp, desc, mix, 5.0, rc, rc, box, r, v, q, L, cid, T)
