"""ctypes binding of the C ABI in include/ls1amd.h (libls1amd.so).

Plumbing for the tests, the smoke check and the benchmark. There is no
fallback: if the HIP library is missing or no device is visible, the calls
raise.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libls1amd.so"

PREC_FP64 = 0
PREC_FP32_FP64ACC = 1
BOUNDARY_PERIODIC, BOUNDARY_OUTFLOW, BOUNDARY_REFLECTING = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

# every symbol include/ls1amd.h declares
EXPORTS = (
    "mdx_last_error", "mdx_create", "mdx_destroy", "mdx_set_components", "mdx_set_lj_tables",
    "mdx_set_domain", "mdx_set_precision", "mdx_upload", "mdx_download", "mdx_update_container",
    "mdx_traverse", "mdx_leapfrog_pre", "mdx_leapfrog_post", "mdx_step", "mdx_synchronize",
    "mdx_cell_ids", "mdx_cell_geometry", "mdx_cell_pairs", "mdx_set_counting", "mdx_counters",
    "mdx_upload_device", "mdx_set_timing", "mdx_last_kernel_ms", "mdx_set_charges",
    "mdx_adjoint", "mdx_lj_tables", "mdx_set_replicas", "mdx_set_pair_config", "mdx_autotune",
    "mdx_set_thermostat", "mdx_thermostat_state", "mdx_set_rmm", "mdx_set_boundaries",
    "mdx_adjoint_noalloc", "mdx_molecule_count", "mdx_kernel_stats",
)

_lib = None


class MdxError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MdxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.mdx_last_error.restype = C.c_char_p
    L.mdx_last_error.argtypes = [vp]
    L.mdx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.mdx_destroy.argtypes = [vp]
    L.mdx_set_components.argtypes = [vp, C.c_int, _dp, C.c_int, vp, C.c_double]
    L.mdx_set_lj_tables.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.mdx_set_domain.argtypes = [vp, _dp, C.c_double, C.c_double, C.c_int, C.c_int]
    L.mdx_set_precision.argtypes = [vp, C.c_int]
    L.mdx_upload.argtypes = [vp, C.c_int64, _dp, _dp, _dp, _dp, _ip, vp]
    L.mdx_download.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.mdx_update_container.argtypes = [vp]
    L.mdx_traverse.argtypes = [vp, vp]
    L.mdx_leapfrog_pre.argtypes = [vp, C.c_double]
    L.mdx_leapfrog_post.argtypes = [vp, C.c_double, vp]
    L.mdx_step.argtypes = [vp, C.c_int, C.c_double, vp, C.c_int]
    L.mdx_synchronize.argtypes = [vp]
    L.mdx_cell_ids.argtypes = [vp, C.c_int64, _lp]
    L.mdx_cell_geometry.argtypes = [vp, _ip, _dp]
    L.mdx_cell_pairs.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.mdx_set_counting.argtypes = [vp, C.c_int]
    L.mdx_counters.argtypes = [vp, _lp]
    L.mdx_upload_device.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
    L.mdx_set_timing.argtypes = [vp, C.c_int]
    L.mdx_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mdx_set_charges.argtypes = [vp, C.c_int, _dp]
    L.mdx_lj_tables.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.mdx_set_replicas.argtypes = [vp, C.c_int, _dp, _lp, C.c_double, C.c_double, C.c_int,
                                   C.c_int]
    L.mdx_set_pair_config.argtypes = [vp, C.c_int, C.c_int]
    L.mdx_autotune.argtypes = [vp, C.c_int, _ip, C.POINTER(C.c_double)]
    L.mdx_set_thermostat.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int]
    L.mdx_thermostat_state.argtypes = [vp, _dp]
    L.mdx_set_rmm.argtypes = [vp, C.c_double]
    L.mdx_set_boundaries.argtypes = [vp, _ip]
    L.mdx_adjoint.argtypes = [vp, C.c_int64] + [vp] * 8
    L.mdx_adjoint_noalloc.argtypes = [vp, C.c_int64] + [vp] * 8
    L.mdx_molecule_count.argtypes = [vp, vp]
    L.mdx_kernel_stats.argtypes = [vp, vp, vp]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context: model, domain, state (mirrors ls1's
    VectorizedCellProcessor + LinkedCells + Leapfrog for one subdomain)."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.mdx_create(int(device), C.byref(h))
        if rc != 0:
            raise MdxError(f"mdx_create failed ({rc}): {self.L.mdx_last_error(None).decode()}")
        self.h = h
        self.n = 0
        self.nrep = 1
        self.open = False

    def _chk(self, rc, what):
        if rc != 0:
            raise MdxError(f"{what} failed ({rc}): {self.L.mdx_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def setup(self, ncomp, comp_desc, mix, eps_rf, box, rc, rc_lj, periodic=True, cell_mode=0,
              precision=PREC_FP64):
        desc = np.ascontiguousarray(comp_desc, np.float64)
        self.ncomp, self.nlj, pos = int(ncomp), 0, 0
        for _ in range(self.ncomp):  # flat component records: counts, sites, moments
            nlj, nc, nd, nq = (int(x) for x in desc[pos:pos + 4])
            self.nlj += nlj
            pos += 4 + 8 * nlj + 5 * nc + 7 * nd + 7 * nq + 3
        self._mix = np.ascontiguousarray(mix, np.float64).ravel() if mix is not None and len(mix) else None
        self._chk(self.L.mdx_set_components(self.h, int(ncomp), desc, desc.size, _ptr(self._mix),
                                            float(eps_rf)), "mdx_set_components")
        self._chk(self.L.mdx_set_domain(self.h, np.ascontiguousarray(box, np.float64), float(rc),
                                        float(rc_lj), int(bool(periodic)), int(cell_mode)),
                  "mdx_set_domain")
        self.nrep = 1
        self.open = False  # a new domain has periodic boundaries
        self.set_precision(precision)
        return self

    def setup_system(self, s, periodic=True, precision=PREC_FP64, cell_mode=0):
        self.setup(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.box, s.rc, s.rc_lj, periodic,
                   cell_mode, precision)
        self.upload(s.r, s.v, s.q, s.L, s.cid)
        return self

    def set_replicas(self, boxes, counts, rc, rc_lj, periodic=True, cell_mode=0):
        boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 3)
        counts = np.ascontiguousarray(counts, np.int64)
        self._chk(self.L.mdx_set_replicas(self.h, len(boxes), boxes.ravel(), counts, float(rc),
                                          float(rc_lj), int(bool(periodic)), int(cell_mode)),
                  "mdx_set_replicas")
        self.nrep = len(boxes)

    def set_pair_config(self, cells_per_block=1, blocks_per_xcd=0):
        self._chk(self.L.mdx_set_pair_config(self.h, int(cells_per_block), int(blocks_per_xcd)),
                  "mdx_set_pair_config")

    def autotune(self, reps=3):
        ch = np.zeros(2, np.int32)
        ms = C.c_double(0)
        self._chk(self.L.mdx_autotune(self.h, int(reps), ch, C.byref(ms)), "mdx_autotune")
        return (int(ch[0]), int(ch[1])), ms.value

    def set_thermostat(self, T_target=None, lrc=False, explosion_heuristics=True):
        """NVT tail of every step (global velocity-scaling thermostat per
        replica, homogeneous long-range correction); T_target: scalar or one
        value per replica, None or <= 0 for no thermostat."""
        nrep = self.nrep
        if T_target is None:
            tp = None
        else:
            t = np.ascontiguousarray(np.broadcast_to(np.asarray(T_target, np.float64), (nrep,)))
            tp = t.ctypes.data_as(C.c_void_p)
        self._chk(self.L.mdx_set_thermostat(self.h, nrep, tp, int(bool(lrc)),
                                            int(bool(explosion_heuristics))),
                  "mdx_set_thermostat")

    def set_boundaries(self, types):
        """Per-dimension boundary types: BOUNDARY_PERIODIC, BOUNDARY_OUTFLOW or
        BOUNDARY_REFLECTING."""
        t = np.ascontiguousarray(types, np.int32)
        self._chk(self.L.mdx_set_boundaries(self.h, t), "mdx_set_boundaries")
        self.open = bool(np.any(t == BOUNDARY_OUTFLOW))

    def set_rmm(self, dt):
        """Reduced-memory one-centre mode (half-step velocities, kick in the
        force kernel); runs the initial half kick. dt = 0 switches it off."""
        self._chk(self.L.mdx_set_rmm(self.h, float(dt)), "mdx_set_rmm")

    def thermostat_state(self):
        """[nrep, 8]: T_target, UpotCorr, VirialCorr, VirialCorr/(3N), T,
        beta_trans, beta_rot, explosion flag of the last step."""
        out = np.zeros((self.nrep, 8))
        self._chk(self.L.mdx_thermostat_state(self.h, out), "mdx_thermostat_state")
        return out

    def setup_replicas(self, systems, periodic=True, precision=PREC_FP64, cell_mode=0):
        """Batched replicas from a list of systems that share components and
        cutoffs (each keeps its own box and molecules)."""
        s0 = systems[0]
        desc = np.ascontiguousarray(s0.comp_desc, np.float64)
        self.setup(s0.ncomp, desc, s0.mix, s0.eps_rf, s0.box, s0.rc, s0.rc_lj, periodic,
                   cell_mode, precision)
        self.set_replicas([s.box for s in systems], [s.n for s in systems], s0.rc, s0.rc_lj,
                          periodic, cell_mode)
        cat = lambda k: np.concatenate([np.asarray(getattr(s, k)) for s in systems])
        self.upload(cat("r"), cat("v"), cat("q"), cat("L"), cat("cid"))
        return self

    def set_precision(self, p):
        self._chk(self.L.mdx_set_precision(self.h, int(p)), "mdx_set_precision")

    def set_lj_tables(self, eps24, sig2, shift6):
        n = int(np.sqrt(np.asarray(eps24).size))
        self._chk(self.L.mdx_set_lj_tables(self.h, n, np.ascontiguousarray(eps24, np.float64).ravel(),
                                           np.ascontiguousarray(sig2, np.float64).ravel(),
                                           np.ascontiguousarray(shift6, np.float64).ravel()),
                  "mdx_set_lj_tables")

    def upload(self, r, v, q, L, cid, ids=None):
        n = len(r)
        self.n = n
        f = lambda a, w: np.ascontiguousarray(np.asarray(a, np.float64).reshape(n, w))
        idp = None
        if ids is not None:
            self._ids = np.ascontiguousarray(ids, np.uint64)
            idp = _ptr(self._ids)
        self._chk(self.L.mdx_upload(self.h, n, f(r, 3), f(v, 3), f(q, 4), f(L, 3),
                                    np.ascontiguousarray(cid, np.int32), idp), "mdx_upload")

    def molecule_count(self):
        """Live molecules: the uploaded count less those deleted at outflow walls."""
        k = C.c_int64(0)
        self._chk(self.L.mdx_molecule_count(self.h, C.byref(k)), "mdx_molecule_count")
        return k.value

    def download(self):
        """State in upload order (with outflow walls: the surviving molecules,
        ids tell which)."""
        n = self.molecule_count() if self.open else self.n
        out = dict(r=np.zeros((n, 3)), v=np.zeros((n, 3)), q=np.zeros((n, 4)), L=np.zeros((n, 3)),
                   F=np.zeros((n, 3)), M=np.zeros((n, 3)), Vi=np.zeros((n, 3)),
                   ids=np.zeros(n, np.uint64))
        self._chk(self.L.mdx_download(self.h, n, *(_ptr(out[k]) for k in
                                                   ("r", "v", "q", "L", "F", "M", "Vi", "ids"))),
                  "mdx_download")
        return out

    def update_container(self):
        self._chk(self.L.mdx_update_container(self.h), "mdx_update_container")

    def traverse(self):
        """macro sums [4] ([nrep, 4] for batched replicas)."""
        mac = np.zeros((self.nrep, 4))
        self._chk(self.L.mdx_traverse(self.h, _ptr(mac)), "mdx_traverse")
        return mac[0] if self.nrep == 1 else mac

    def forces(self):
        """update_container + traverse, then the molecule outputs."""
        self.update_container()
        mac = self.traverse()
        out = self.download()
        out["macro"] = mac
        return out

    def leapfrog_pre(self, dt):
        self._chk(self.L.mdx_leapfrog_pre(self.h, float(dt)), "mdx_leapfrog_pre")

    def leapfrog_post(self, dt):
        sums = np.zeros((self.nrep, 4))
        self._chk(self.L.mdx_leapfrog_post(self.h, float(dt), _ptr(sums)), "mdx_leapfrog_post")
        return sums[0] if self.nrep == 1 else sums

    def step(self, nsteps, dt, use_graph=False):
        """Per-step records [nsteps, 6] ([nsteps, nrep, 6] for batched replicas):
        Upot, virial, sum m v^2, sum I w^2, Upot_LJ, Upot_rest."""
        per = np.zeros(6 * self.nrep * max(nsteps, 1))
        self._chk(self.L.mdx_step(self.h, int(nsteps), float(dt), _ptr(per), int(bool(use_graph))),
                  "mdx_step")
        per = per.reshape(-1, self.nrep, 6)[:nsteps]
        return per[:, 0] if self.nrep == 1 else per

    def synchronize(self):
        self._chk(self.L.mdx_synchronize(self.h), "mdx_synchronize")

    def cell_ids(self):
        out = np.zeros(self.molecule_count() if self.open else self.n, np.int64)
        self._chk(self.L.mdx_cell_ids(self.h, len(out), out), "mdx_cell_ids")
        return out

    def cell_geometry(self):
        cpd = np.zeros(3, np.int32)
        cl = np.zeros(3)
        self._chk(self.L.mdx_cell_geometry(self.h, cpd, cl), "mdx_cell_geometry")
        return cpd, cl

    def cell_pairs(self):
        cnt = C.c_int64(0)
        self._chk(self.L.mdx_cell_pairs(self.h, None, 0, C.byref(cnt)), "mdx_cell_pairs")
        out = np.zeros(2 * cnt.value, np.int32)
        self._chk(self.L.mdx_cell_pairs(self.h, _ptr(out), cnt.value, C.byref(cnt)),
                  "mdx_cell_pairs")
        return out.reshape(-1, 2)

    def set_counting(self, on):
        self._chk(self.L.mdx_set_counting(self.h, int(bool(on))), "mdx_set_counting")

    def counters(self):
        out = np.zeros(8, np.int64)
        self._chk(self.L.mdx_counters(self.h, out), "mdx_counters")
        return out

    def set_timing(self, on):
        self._chk(self.L.mdx_set_timing(self.h, int(bool(on))), "mdx_set_timing")

    def kernel_stats(self):
        """Per pair kernel (LJ, charge, dipole, quadrupole i sites): the last
        timed traversal's durations [4] (ms) and the last counted traversal's
        raw visit counts [4, 8] (each kernel counts its own side of a pair)."""
        ms = np.zeros(4)
        cnt = np.zeros((4, 8), np.int64)
        self._chk(self.L.mdx_kernel_stats(self.h, _ptr(ms), _ptr(cnt)), "mdx_kernel_stats")
        return ms, cnt

    def last_kernel_ms(self):
        a, b = C.c_double(0), C.c_double(0)
        self._chk(self.L.mdx_last_kernel_ms(self.h, C.byref(a), C.byref(b)), "mdx_last_kernel_ms")
        return a.value, b.value


    def lj_tables(self):
        """(eps24, sig2, shift6, shift_rc), each [nlj, nlj]."""
        nl = self.nlj
        t = [np.zeros((nl, nl)) for _ in range(4)]
        self._chk(self.L.mdx_lj_tables(self.h, nl, *(_ptr(x) for x in t)), "mdx_lj_tables")
        return tuple(t)

    def set_charges(self, q):
        """q[comp, site] (ncomp x 8) charges; rebuilds the site arrays."""
        q = np.ascontiguousarray(q, np.float64)
        self._chk(self.L.mdx_set_charges(self.h, q.shape[0], q.ravel()), "mdx_set_charges")

    def adjoint(self, cot=None, noalloc=False):
        """FP64 adjoint for the state of the last update_container:
        dU/dW with respect to LJ pair eps/sigma ([nlj, nlj]) and charges
        ([ncomp, 8]); with a cotangent on the molecule forces, the force
        vector-Jacobian product with respect to positions ([n, 3]).
        noalloc: through mdx_adjoint_noalloc (the library's per-device
        scratch array instead of buffers allocated on the first call)."""
        n, nl, ncomp = self.n, self.nlj, self.ncomp
        out = {k: np.zeros((nl, nl)) for k in ("dU_deps", "dU_dsig", "dW_deps", "dW_dsig")}
        out["dU_dq"], out["dW_dq"] = np.zeros((ncomp, 8)), np.zeros((ncomp, 8))
        c = None
        if cot is not None:
            c = np.ascontiguousarray(cot, np.float64).reshape(n, 3)
            out["vjp"] = np.zeros((n, 3))
        fn, name = ((self.L.mdx_adjoint_noalloc, "mdx_adjoint_noalloc") if noalloc
                    else (self.L.mdx_adjoint, "mdx_adjoint"))
        self._chk(fn(self.h, n, _ptr(c), *(_ptr(out[k]) for k in (
            "dU_deps", "dU_dsig", "dW_deps", "dW_dsig", "dU_dq", "dW_dq")), _ptr(out.get("vjp"))),
            name)
        return out


def upot_virial(macro):
    """VectorizedCellProcessor::endTraversal (VectorizedCellProcessor.cpp:152-154)."""
    return macro[0] / 6.0 + macro[1] + macro[3], macro[2] + 3.0 * macro[3]
