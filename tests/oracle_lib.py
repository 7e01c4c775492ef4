"""ctypes bindings for the CPU checkers (test infrastructure only).

* ``Oracle``   -- oracle/liboracle.so, the C restatement of the reference path.
* ``RefLegacy``-- oracle/_ref/libls1ref.so, the reference's own molecule code
  and potforce.h Legacy pair physics (present only where the reference
  sources were available to build it).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libls1ref.so"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

MACRO_NAMES = ("upot6lj", "upotXpoles", "virial", "myRF")
COUNTER_NAMES = ("molpairs", "LJ", "CC", "CD", "DD", "CQ", "DQ", "QQ")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def macro_to_upot_virial(mac):
    """VectorizedCellProcessor::endTraversal, VectorizedCellProcessor.cpp:152-154."""
    return mac[0] / 6.0 + mac[1] + mac[3], mac[2] + 3.0 * mac[3]


class Oracle:
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not ORACLE_SO.exists():
                raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
            L = C.CDLL(str(ORACLE_SO))
            L.ox_create.restype = C.c_void_p
            L.ox_create.argtypes = [C.c_int, _dp, C.c_int, C.c_void_p, C.c_double, C.c_double,
                                    C.c_double, _dp, C.c_int, C.c_int]
            L.ox_destroy.argtypes = [C.c_void_p]
            L.ox_nlj_total.argtypes = [C.c_void_p]
            L.ox_lj_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
            L.ox_component_props.argtypes = [C.c_void_p, _dp, _dp, _ip]
            L.ox_cell_geometry.argtypes = [C.c_void_p, _ip, _dp, _dp, _dp]
            L.ox_cell_index_of_point.restype = C.c_long
            L.ox_cell_index_of_point.argtypes = [C.c_void_p, _dp]
            L.ox_cell_index_batch.argtypes = [C.c_void_p, C.c_int, _dp, _lp]
            L.ox_c08_pairs.argtypes = [C.c_void_p, _ip, C.c_int]
            L.ox_wrap.argtypes = [C.c_void_p, C.c_int, _dp]
            L.ox_forces.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _lp]
            L.ox_forces_brute.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _dp]
            L.ox_leapfrog_pre.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                          _dp, _dp, _ip]
            L.ox_leapfrog_post.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                           _dp, _ip, _dp]
            L.ox_simulate.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                      _dp, _ip, _dp, _dp, _dp, _dp]
            L.ox_set_lj_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
            L.ox_lrc_homogeneous.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
            L.ox_thermostat_params.argtypes = [_dp, C.c_double, C.c_int, _dp]
            L.ox_velocity_scaling.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip,
                                              _dp, C.c_double]
            L.ox_set_boundaries.argtypes = [C.c_void_p, _ip]
            L.ox_reflect.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _ip,
                                     C.c_void_p]
            L.ox_simulate_open.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                           _dp, _dp, _ip, _dp, _dp, _dp, _ip, _dp]
            L.ox_simulate_rmm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                          _ip, _dp]
            L.ox_simulate_nvt.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp,
                                          _dp, _ip, _dp, _dp, _dp, C.c_void_p, C.c_double,
                                          C.c_int, _dp]
            L.ox_lj_shift_rc.argtypes = [C.c_void_p, _dp]
            L.ox_adjoint.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp]
            L.ox_set_threads.argtypes = [C.c_void_p, C.c_int]
            L.ox_count_checks.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _lp]
            L.ox_site_pair.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, _dp]
            cls._lib = L
        return cls._lib

    def __init__(self, ncomp, comp_desc, mix, eps_rf, rc, rc_lj, box, periodic=True,
                 cell_mode=0):
        L = self.lib()
        desc = _f64(comp_desc)
        self._mix = _f64(mix).ravel() if mix is not None and len(mix) else None
        mixp = self._mix.ctypes.data_as(C.c_void_p) if self._mix is not None else None
        self.box = _f64(box)
        self.h = L.ox_create(int(ncomp), desc, desc.size, mixp, float(eps_rf), float(rc),
                             float(rc_lj), self.box, int(bool(periodic)), int(cell_mode))
        if not self.h:
            raise ValueError("ox_create failed (bad component records)")
        self.ncomp = int(ncomp)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib().ox_destroy(h)
            self.h = None

    @classmethod
    def from_phase_space(cls, ps, rc, rc_lj=None, periodic=True, cell_mode=0):
        return cls(ps.ncomp, ps.comp_desc, ps.mix, ps.eps_rf, rc, rc if rc_lj is None else rc_lj,
                   ps.box, periodic, cell_mode)

    def set_threads(self, n):
        """OpenMP threads of the traversal and molecule loops (1: the serial
        path every parity test uses)."""
        self.lib().ox_set_threads(self.h, int(n))

    def count_checks(self, r, q, cid):
        """Molecule-centre distance checks of the reference's C08 traversal:
        (all, those with macroscopic sums)."""
        out = np.zeros(2, np.int64)
        if self.lib().ox_count_checks(self.h, len(r), _f64(r), _f64(q),
                                      np.ascontiguousarray(cid, np.int32), out):
            raise RuntimeError("oracle count failed")
        return int(out[0]), int(out[1])

    def lj_tables(self):
        n = self.lib().ox_nlj_total(self.h)
        e, s, sh = (np.zeros(n * n) for _ in range(3))
        self.lib().ox_lj_tables(self.h, e, s, sh)
        return e.reshape(n, n), s.reshape(n, n), sh.reshape(n, n)

    def component_props(self):
        m = np.zeros(self.ncomp)
        I = np.zeros(3 * self.ncomp)
        rot = np.zeros(self.ncomp, np.int32)
        self.lib().ox_component_props(self.h, m, I, rot)
        return m, I.reshape(-1, 3), rot

    def set_lj_tables(self, eps24, sig2, shift6):
        self.lib().ox_set_lj_tables(self.h, _f64(eps24).ravel(), _f64(sig2).ravel(),
                                    _f64(shift6).ravel())

    def lj_shift_rc(self):
        n = self.lib().ox_nlj_total(self.h)
        out = np.zeros(n * n)
        self.lib().ox_lj_shift_rc(self.h, out)
        return out.reshape(n, n)

    def adjoint(self, r, q, cid, cot):
        n = len(r)
        nl = self.lib().ox_nlj_total(self.h)
        dUe, dUs, dWe, dWs = (np.zeros((nl, nl)) for _ in range(4))
        dUq, dWq = np.zeros((self.ncomp, 8)), np.zeros((self.ncomp, 8))
        vjp = np.zeros((n, 3))
        self.lib().ox_adjoint(self.h, n, _f64(r), _f64(q).copy(), np.ascontiguousarray(cid, np.int32),
                              _f64(cot), dUe, dUs, dWe, dWs, dUq, dWq, vjp)
        return dict(dU_deps=dUe, dU_dsig=dUs, dW_deps=dWe, dW_dsig=dWs, dU_dq=dUq, dW_dq=dWq,
                    vjp=vjp)

    def cell_geometry(self):
        cpd = np.zeros(3, np.int32)
        cl, hmin, hmax = np.zeros(3), np.zeros(3), np.zeros(3)
        self.lib().ox_cell_geometry(self.h, cpd, cl, hmin, hmax)
        return cpd, cl, hmin, hmax

    def cell_index(self, r):
        r = _f64(r).reshape(-1, 3)
        out = np.zeros(len(r), np.int64)
        self.lib().ox_cell_index_batch(self.h, len(r), r, out)
        return out

    def c08_pairs(self):
        n = self.lib().ox_c08_pairs(self.h, np.zeros(2, np.int32), 0)
        out = np.zeros(2 * n, np.int32)
        self.lib().ox_c08_pairs(self.h, out, n)
        return out.reshape(n, 2)

    def wrap(self, r):
        r = _f64(r).copy()
        self.lib().ox_wrap(self.h, len(r), r)
        return r

    def forces(self, r, q, cid, brute=False):
        r = _f64(r)
        q = _f64(q).copy()
        cid = np.ascontiguousarray(cid, np.int32)
        n = len(r)
        F, M, Vi = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        mac = np.zeros(4)
        cnt = np.zeros(8, np.int64)
        if brute:
            rc = self.lib().ox_forces_brute(self.h, n, r, q, cid, F, M, Vi, mac)
        else:
            rc = self.lib().ox_forces(self.h, n, r, q, cid, F, M, Vi, mac, cnt)
        if rc:
            raise RuntimeError("oracle force evaluation failed")
        return dict(F=F, M=M, Vi=Vi, macro=mac, q=q, counters=cnt)

    def leapfrog_pre(self, dt, r, v, q, L, F, M, cid):
        r, v, q, L = (_f64(a).copy() for a in (r, v, q, L))
        self.lib().ox_leapfrog_pre(self.h, len(r), float(dt), r, v, q, L, _f64(F), _f64(M),
                                   np.ascontiguousarray(cid, np.int32))
        return r, v, q, L

    def leapfrog_post(self, dt, v, L, q, F, M, cid):
        v, L = _f64(v).copy(), _f64(L).copy()
        sums = np.zeros(4)
        self.lib().ox_leapfrog_post(self.h, len(v), float(dt), v, L, _f64(q), _f64(F), _f64(M),
                                    np.ascontiguousarray(cid, np.int32), sums)
        return v, L, sums

    def simulate(self, nsteps, dt, r, v, q, L, cid):
        """Initial forces, then nsteps of the reference step order."""
        r, v, q, L = (_f64(a).copy() for a in (r, v, q, L))
        cid = np.ascontiguousarray(cid, np.int32)
        f0 = self.forces(r, q, cid)
        q = f0["q"]
        F, M, Vi = f0["F"].copy(), f0["M"].copy(), f0["Vi"].copy()
        per = np.zeros(6 * max(nsteps, 1))
        rc = self.lib().ox_simulate(self.h, len(r), int(nsteps), float(dt), r, v, q, L, cid, F, M,
                                    Vi, per)
        if rc:
            raise RuntimeError("oracle simulate failed")
        return dict(r=r, v=v, q=q, L=L, F=F, M=M, Vi=Vi, per_step=per.reshape(-1, 6)[:nsteps],
                    initial=f0)

    def lrc_homogeneous(self, nmol, volume):
        """UpotCorr, VirialCorr, VirialCorr/(3N) of Homogeneous (ox_lrc_homogeneous)."""
        out = np.zeros(3)
        if self.lib().ox_lrc_homogeneous(self.h, _f64(nmol), float(volume), out):
            raise RuntimeError("long-range correction: rcLJ too small")
        return out

    @classmethod
    def thermostat_params(cls, sums, T_target, heuristics=True):
        """T, beta_trans, beta_rot, explode from the upd_postF sums."""
        out = np.zeros(4)
        cls.lib().ox_thermostat_params(_f64(sums), float(T_target), int(bool(heuristics)), out)
        return out

    def velocity_scaling(self, v, L, F, M, q, cid, tp, T_target):
        v, L, F, M = (_f64(a).copy() for a in (v, L, F, M))
        self.lib().ox_velocity_scaling(self.h, len(v), v, L, F, M, _f64(q),
                                       np.ascontiguousarray(cid, np.int32), _f64(tp),
                                       float(T_target))
        return v, L, F, M

    def simulate_nvt(self, nsteps, dt, r, v, q, L, cid, T_target, lrc=False, heuristics=True):
        """ox_simulate with the NVT tail: initial forces (with the long-range
        virial addition when lrc), then nsteps; per_step rows are Upot,
        virial, summv2, sumIw2, Upot_LJ, Upot_rest, T, beta_trans, beta_rot."""
        r, v, q, L = (_f64(a).copy() for a in (r, v, q, L))
        cid = np.ascontiguousarray(cid, np.int32)
        terms = None
        if lrc:
            nmol = np.bincount(cid, minlength=self.ncomp).astype(np.float64)
            terms = self.lrc_homogeneous(nmol, float(np.prod(self.box)))
        f0 = self.forces(r, q, cid)
        q = f0["q"]
        F, M, Vi = f0["F"].copy(), f0["M"].copy(), f0["Vi"].copy()
        if terms is not None:
            Vi = terms[2] + Vi
        per = np.zeros(9 * max(nsteps, 1))
        tp = terms.ctypes.data_as(C.c_void_p) if terms is not None else None
        rc = self.lib().ox_simulate_nvt(self.h, len(r), int(nsteps), float(dt), r, v, q, L, cid,
                                        F, M, Vi, tp, float(T_target), int(bool(heuristics)), per)
        if rc:
            raise RuntimeError("oracle simulate failed")
        return dict(r=r, v=v, q=q, L=L, F=F, M=M, Vi=Vi, per_step=per.reshape(-1, 9)[:nsteps],
                    lrc=terms)

    def set_boundaries(self, types):
        """Per dimension: 0 periodic, 1 outflow, 2 reflecting."""
        if self.lib().ox_set_boundaries(self.h, np.ascontiguousarray(types, np.int32)):
            raise ValueError("unsupported boundary type")

    def reflect(self, dt, r, v, F, cid, with_dead=False):
        """The wall rule: new velocities (reflecting walls) and, with
        with_dead, the deletion flags of outflow walls."""
        v = _f64(v).copy()
        dead = np.zeros(len(v), np.int32)
        self.lib().ox_reflect(self.h, len(v), float(dt), _f64(r), v, _f64(F),
                              np.ascontiguousarray(cid, np.int32),
                              dead.ctypes.data_as(C.c_void_p) if with_dead else None)
        return (v, dead.astype(bool)) if with_dead else v

    def simulate_open(self, nsteps, dt, r, v, q, L, cid):
        """simulate with outflow walls: survivors (compacted, orig = their
        upload index) and per-step rows with a 7th column N."""
        r, v, q, L = (_f64(a).copy() for a in (r, v, q, L))
        cid = np.ascontiguousarray(cid, np.int32).copy()
        f0 = self.forces(r, q, cid)
        q = f0["q"]
        F, M, Vi = f0["F"].copy(), f0["M"].copy(), f0["Vi"].copy()
        orig = np.arange(len(r), dtype=np.int32)
        per = np.zeros(7 * max(nsteps, 1))
        k = self.lib().ox_simulate_open(self.h, len(r), int(nsteps), float(dt), r, v, q, L, cid,
                                        F, M, Vi, orig, per)
        if k < 0:
            raise RuntimeError("oracle simulate_open failed")
        return dict(r=r[:k], v=v[:k], q=q[:k], L=L[:k], F=F[:k], M=M[:k], Vi=Vi[:k], cid=cid[:k],
                    orig=orig[:k], per_step=per.reshape(-1, 7)[:nsteps])

    def simulate_rmm(self, nsteps, dt, r, v, q, cid):
        """Reduced-memory mode: initial half kick, then nsteps; returns the
        half-step velocities and the per-step records (sum v^2 in column 2)."""
        r, v, q = (_f64(a).copy() for a in (r, v, q))
        cid = np.ascontiguousarray(cid, np.int32)
        per = np.zeros(6 * max(nsteps, 1))
        rc = self.lib().ox_simulate_rmm(self.h, len(r), int(nsteps), float(dt), r, v, q, cid, per)
        if rc:
            raise RuntimeError("oracle simulate_rmm failed")
        return dict(r=r, v=v, per_step=per.reshape(-1, 6)[:nsteps])

    @classmethod
    def site_pair(cls, kind, ri, ei, pi, rj, ej, pj, eps24=0.0, sig2=0.0, shift6=0.0,
                  eps_rf_invrc3=0.0):
        out = np.zeros(11)
        cls.lib().ox_site_pair(int(kind), _f64(ri), _f64(ei), float(pi), _f64(rj), _f64(ej),
                               float(pj), float(eps24), float(sig2), float(shift6),
                               float(eps_rf_invrc3), out)
        return out


class RefLegacy:
    _lib = None

    @classmethod
    def available(cls):
        return REF_SO.exists()

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(str(REF_SO))
            L.ref_create.restype = C.c_void_p
            L.ref_create.argtypes = [C.c_int, _dp, C.c_int, C.c_void_p, C.c_double, C.c_double,
                                     C.c_double]
            L.ref_destroy.argtypes = [C.c_void_p]
            L.ref_component_props.argtypes = [C.c_void_p, _dp, _dp, _ip]
            L.ref_lj_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
            L.ref_forces.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _dp, C.c_int, _dp, _dp,
                                     _dp, _dp]
            L.ref_upd_preF.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _ip]
            L.ref_upd_postF.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                        _ip, _dp]
            cls._lib = L
        return cls._lib

    def __init__(self, ncomp, comp_desc, mix, eps_rf, rc, rc_lj):
        L = self.lib()
        desc = _f64(comp_desc)
        self._mix = _f64(mix).ravel() if mix is not None and len(mix) else None
        mixp = self._mix.ctypes.data_as(C.c_void_p) if self._mix is not None else None
        self.h = L.ref_create(int(ncomp), desc, desc.size, mixp, float(eps_rf), float(rc),
                              float(rc_lj))
        if not self.h:
            raise ValueError("ref_create failed")
        self.ncomp = int(ncomp)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib().ref_destroy(h)
            self.h = None

    @classmethod
    def from_phase_space(cls, ps, rc, rc_lj=None):
        return cls(ps.ncomp, ps.comp_desc, ps.mix, ps.eps_rf, rc, rc if rc_lj is None else rc_lj)

    def component_props(self):
        m = np.zeros(self.ncomp)
        I = np.zeros(3 * self.ncomp)
        rot = np.zeros(self.ncomp, np.int32)
        self.lib().ref_component_props(self.h, m, I, rot)
        return m, I.reshape(-1, 3), rot

    def lj_tables(self, nlj):
        e, s, sh = (np.zeros(nlj * nlj) for _ in range(3))
        self.lib().ref_lj_tables(self.h, e, s, sh)
        return e.reshape(nlj, nlj), s.reshape(nlj, nlj), sh.reshape(nlj, nlj)

    def forces(self, r, q, cid, box, periodic=False):
        n = len(r)
        F, M, Vi = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        mac = np.zeros(2)
        self.lib().ref_forces(self.h, n, _f64(r), _f64(q), np.ascontiguousarray(cid, np.int32),
                              _f64(box), int(bool(periodic)), F, M, Vi, mac)
        return dict(F=F, M=M, Vi=Vi, upot=mac[0], virial=mac[1])

    def upd_preF(self, dt, r, v, q, L, F, M, cid):
        r, v, q, L = (_f64(a).copy() for a in (r, v, q, L))
        self.lib().ref_upd_preF(self.h, len(r), float(dt), r, v, q, L, _f64(F), _f64(M),
                                np.ascontiguousarray(cid, np.int32))
        return r, v, q, L

    def upd_postF(self, dt, v, L, q, F, M, cid):
        v, L = _f64(v).copy(), _f64(L).copy()
        sums = np.zeros(2)
        self.lib().ref_upd_postF(self.h, len(v), float(dt), v, L, _f64(q), _f64(F), _f64(M),
                                 np.ascontiguousarray(cid, np.int32), sums)
        return v, L, sums
