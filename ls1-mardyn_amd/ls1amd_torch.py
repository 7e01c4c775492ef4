"""torch.autograd bridge for force-field fitting on the HIP force path.

`ForceFieldEngine` holds one FP64 device context with a molecule state and
evaluates Upot, virial and the molecule forces for given LJ pair parameters,
charges and positions. `ForceField.apply(eps, sigma, charges, r, engine)` is
the differentiable call:
  * eps, sigma: 1-D tensors over the unordered LJ lookup-id pairs (a <= b,
    row-major upper triangle, `engine.pairs`);
  * charges: [ncomp, 8] tensor (entries beyond a component's charge count are
    ignored);
  * r: [n, 3] molecule positions.
It returns (Upot, virial, F). The backward pass calls the device adjoint
(recomputing the pair geometry from the cell lists, no stored tape):
  d/d(eps, sigma, charges) of Upot and virial, and the force vector-Jacobian
  product with respect to r; dUpot/dr = -F.
Not provided (the call raises instead of returning a wrong gradient): the
derivative of the forces with respect to the parameters, and of the virial
with respect to positions.

The truncation shift moves with (eps, sigma) exactly where ls1's Comp2Param
rules build it from the pair's own parameters (a shifted site with itself,
every cross-component pair at rcLJ, Comp2Param.cpp:27-38,79-82,102-105);
elsewhere it keeps its table value.
"""
from __future__ import annotations

import numpy as np
import torch

from ls1amd import PREC_FP64, Context


class ForceFieldEngine:
    def __init__(self, system, device=0, periodic=True):
        self.s = system
        self.ctx = Context(device)
        self.ctx.setup(system.ncomp, system.comp_desc, system.mix, system.eps_rf, system.box,
                       system.rc, system.rc_lj, periodic, precision=PREC_FP64)
        self.nlj = self.ctx.nlj
        self.ncomp = self.ctx.ncomp
        e24, s2, sh6, src = self.ctx.lj_tables()
        self.base_shift6, self.shift_rc = sh6, src
        self.pairs = [(a, b) for a in range(self.nlj) for b in range(a, self.nlj)]
        iu = tuple(np.array(self.pairs).T) if self.pairs else ((), ())
        self.eps0 = e24[iu] / 24.0
        self.sig0 = np.sqrt(s2[iu])
        self.q0 = self._charges_from_desc()
        self._q_dev = self.q0.copy()
        self._loaded = None

    def _charges_from_desc(self):
        q = np.zeros((self.ncomp, 8))
        d, p = np.asarray(self.s.comp_desc, np.float64), 0
        for k in range(self.ncomp):
            nlj, nc, nd, nq = (int(x) for x in d[p:p + 4])
            p += 4 + 8 * nlj
            for s in range(nc):
                q[k, s] = d[p + 5 * s + 4]
            p += 5 * nc + 7 * nd + 7 * nq + 3
        return q

    def tables(self, eps, sig):
        """Full [nlj, nlj] eps24 / sig2 / shift6 tables from pair parameters."""
        nl = self.nlj
        e = np.zeros((nl, nl))
        g = np.zeros((nl, nl))
        for k, (a, b) in enumerate(self.pairs):
            e[a, b] = e[b, a] = eps[k]
            g[a, b] = g[b, a] = sig[k]
        eps24, sig2 = 24.0 * e, g * g
        shift6 = self.base_shift6.copy()
        m = self.shift_rc > 0
        s6 = (sig2[m] / self.shift_rc[m] ** 2) ** 3
        shift6[m] = eps24[m] * (s6 - s6 * s6)
        return eps24, sig2, shift6

    def load(self, eps, sig, q, r):
        """Puts parameters and positions on the device (no-op if unchanged)."""
        key = (eps.tobytes(), sig.tobytes(), q.tobytes(), r.tobytes())
        if key == self._loaded:
            return
        s = self.s
        self.ctx.set_lj_tables(*self.tables(eps, sig))
        if np.any(q != self._q_dev):
            self.ctx.set_charges(q)
            self._q_dev = q.copy()
        self.ctx.upload(r, s.v, s.q, s.L, s.cid)
        self.ctx.update_container()
        self._macro = self.ctx.traverse()
        self._F = self.ctx.download()["F"]
        self._loaded = key

    def evaluate(self, eps, sig, q, r):
        self.load(eps, sig, q, r)
        m = self._macro
        return m[0] / 6.0 + m[1] + m[3], m[2] + 3.0 * m[3], self._F.copy()

    def adjoint(self, eps, sig, q, r, cot=None):
        self.load(eps, sig, q, r)
        out = self.ctx.adjoint(cot)
        iu = tuple(np.array(self.pairs).T) if self.pairs else ((), ())
        return dict(dU_deps=out["dU_deps"][iu], dU_dsig=out["dU_dsig"][iu],
                    dW_deps=out["dW_deps"][iu], dW_dsig=out["dW_dsig"][iu],
                    dU_dq=out["dU_dq"], dW_dq=out["dW_dq"], vjp=out.get("vjp"))

    def close(self):
        self.ctx.close()


def _np(t):
    return np.ascontiguousarray(t.detach().cpu().numpy(), dtype=np.float64)


class ForceField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sig, charges, r, engine):
        a = (_np(eps), _np(sig), _np(charges), _np(r))
        U, W, F = engine.evaluate(*a)
        ctx.engine = engine
        ctx.arrays = a
        ctx.F = F
        dt = dict(dtype=torch.float64, device=r.device)
        return torch.tensor(U, **dt), torch.tensor(W, **dt), torch.as_tensor(F, **dt)

    @staticmethod
    def backward(ctx, gU, gW, gF):
        eng = ctx.engine
        eps, sig, q, r = ctx.arrays
        need_p = any(ctx.needs_input_grad[:3])
        need_r = ctx.needs_input_grad[3]
        gFn = _np(gF) if gF is not None else None
        f_live = gFn is not None and np.any(gFn != 0)
        gUv = float(gU) if gU is not None else 0.0
        gWv = float(gW) if gW is not None else 0.0
        if need_p and f_live:
            raise NotImplementedError("d(forces)/d(parameters) is not provided by the adjoint")
        if need_r and gWv != 0.0:
            raise NotImplementedError("d(virial)/d(positions) is not provided by the adjoint")
        adj = eng.adjoint(eps, sig, q, r, cot=gFn if (need_r and f_live) else None)
        dev = dict(dtype=torch.float64, device=gU.device if gU is not None else "cpu")
        g_eps = g_sig = g_q = g_r = None
        if ctx.needs_input_grad[0]:
            g_eps = torch.as_tensor(gUv * adj["dU_deps"] + gWv * adj["dW_deps"], **dev)
        if ctx.needs_input_grad[1]:
            g_sig = torch.as_tensor(gUv * adj["dU_dsig"] + gWv * adj["dW_dsig"], **dev)
        if ctx.needs_input_grad[2]:
            g_q = torch.as_tensor(gUv * adj["dU_dq"] + gWv * adj["dW_dq"], **dev)
        if need_r:
            gr = -gUv * ctx.F
            if f_live:
                gr = gr + adj["vjp"]
            g_r = torch.as_tensor(gr, **dev)
        return g_eps, g_sig, g_q, g_r, None


def force_field(eps, sig, charges, r, engine):
    """(Upot, virial, F) as differentiable functions of the inputs."""
    return ForceField.apply(eps, sig, charges, r, engine)
