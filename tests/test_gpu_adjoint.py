"""GPU adjoint (mdx_adjoint) and its torch.autograd wrapper.

* against the oracle's FP64 brute-force adjoint (itself pinned by finite
  differences in test_oracle_adjoint.py) on the same seeded systems;
* against FP64 central finite differences (h = 1e-6, relative 1e-6) of the
  GPU's own forward path, up to the full 500,000-molecule binary LJ mixture;
* the torch.autograd.Function wrapper against the same quantities.
"""
import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP64, Context, MdxError, upot_virial
from oracle_lib import Oracle
from test_oracle_adjoint import small_binary, small_charged, tables_from_theta

pytestmark = pytest.mark.gpu

H = 1e-6
KEYS = ("dU_deps", "dU_dsig", "dW_deps", "dW_dsig", "dU_dq", "dW_dq", "vjp")


def gpu_ctx(s):
    c = Context(0)
    c.setup_system(s, periodic=True, precision=PREC_FP64)
    c.update_container()
    c.traverse()
    return c


def uw(c):
    return upot_virial(c.traverse())


@pytest.mark.parametrize("make", [small_binary, small_charged])
def test_adjoint_matches_oracle(make):
    s = make()
    print("case", make.__name__, s.n, flush=True)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    cot = np.random.default_rng(3).normal(size=(s.n, 3))
    ref = o.adjoint(s.r, s.q, s.cid, cot)
    with gpu_ctx(s) as c:
        got = c.adjoint(cot)
        again = c.adjoint(cot)
    for k in KEYS:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        np.testing.assert_allclose(got[k], ref[k], rtol=0, atol=1e-11 * scale, err_msg=k)
        # fixed-order reductions: repeated calls are bit-identical
        assert np.array_equal(got[k], again[k]), k


def small_multi(ncomp, seed=83, n=360):
    """One-centre LJ fluid of ncomp components (ncomp LJ lookup ids)."""
    rho = 0.7
    box = np.full(3, (n / rho) ** (1 / 3))
    r = scenarios.rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.85)
    parts = []
    for k in range(ncomp):
        parts += scenarios.comp_1clj(1.0 - 0.07 * k, 0.95 + 0.03 * k, 1.0, 2.5, k % 2 == 1)
    npair = ncomp * (ncomp - 1) // 2
    mix = np.array([[1.0 + 0.01 * (k % 3), 1.0 - 0.02 * (k % 2)] for k in range(npair)])
    rng = np.random.default_rng(seed)
    cid = rng.integers(0, ncomp, n).astype(np.int32)
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    return scenarios.System("multi", ncomp, np.array(parts), mix, 1.0, 2.5, 2.5, box, r,
                            np.zeros((n, 3)), q, np.zeros((n, 3)), cid, 0.9)


@pytest.mark.parametrize("ncomp", [3, 6])
def test_adjoint_many_lj_ids(ncomp):
    """3 LJ ids (parameter sums in registers, 4-id kernel) and 6 (sums in
    per-lane LDS columns) against the oracle's adjoint."""
    s = small_multi(ncomp)
    print("case multi-component LJ", ncomp, s.n, flush=True)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    cot = np.random.default_rng(4).normal(size=(s.n, 3))
    ref = o.adjoint(s.r, s.q, s.cid, cot)
    with gpu_ctx(s) as c:
        got = c.adjoint(cot)
    for k in ("dU_deps", "dU_dsig", "dW_deps", "dW_dsig", "vjp"):
        scale = max(np.abs(ref[k]).max(), 1e-300)
        np.testing.assert_allclose(got[k], ref[k], rtol=0, atol=1e-11 * scale, err_msg=k)


def test_adjoint_after_downloads():
    """Downloads and the adjoint share the context's staging buffer: repeated
    downloads and an adjoint after them keep working and agree."""
    s = small_binary()
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    cot = np.random.default_rng(3).normal(size=(s.n, 3))
    ref = o.adjoint(s.r, s.q, s.cid, cot)
    with gpu_ctx(s) as c:
        d1 = c.download()
        d2 = c.download()
        got = c.adjoint(cot)
        d3 = c.download()
    for k in ("r", "F", "M"):
        assert np.array_equal(d1[k], d2[k]) and np.array_equal(d1[k], d3[k]), k
    np.testing.assert_allclose(got["vjp"], ref["vjp"], rtol=0, atol=1e-11 * np.abs(ref["vjp"]).max())


def test_adjoint_oracle_binary_5000():
    s = scenarios.binary_lj(n=5000)
    print("case binary_lj 5000", flush=True)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
    cot = np.random.default_rng(4).normal(size=(s.n, 3))
    ref = o.adjoint(s.r, s.q, s.cid, cot)
    with gpu_ctx(s) as c:
        got = c.adjoint(cot)
    for k in ("dU_deps", "dU_dsig", "dW_deps", "dW_dsig", "vjp"):
        scale = np.abs(ref[k]).max()
        np.testing.assert_allclose(got[k], ref[k], rtol=0, atol=1e-11 * scale, err_msg=k)


def fd_check_params(c, adj, tol=1e-6):
    e24, s2, sh6, src = c.lj_tables()
    eps, sig = e24 / 24.0, np.sqrt(s2)
    nl = len(eps)
    for a in range(nl):
        for b in range(a, nl):
            for which, ku, kw in ((0, "dU_deps", "dW_deps"), (1, "dU_dsig", "dW_dsig")):
                vals = []
                for sgn in (1, -1):
                    e, g = eps.copy(), sig.copy()
                    t = e if which == 0 else g
                    t[a, b] += sgn * H
                    if a != b:
                        t[b, a] += sgn * H
                    c.set_lj_tables(*tables_from_theta(e, g, src, sh6))
                    vals.append(uw(c))
                c.set_lj_tables(e24, s2, sh6)
                fu = (vals[0][0] - vals[1][0]) / (2 * H)
                fw = (vals[0][1] - vals[1][1]) / (2 * H)
                for got, fd in ((adj[ku][a, b], fu), (adj[kw][a, b], fw)):
                    assert abs(got - fd) <= tol * max(abs(fd), 1e-3), (a, b, which, got, fd)


def smooth_delta(s, seed=7, eps=1e-5):
    """Gaussian direction with the molecules of every pair within eps of a
    cutoff held still: the forces of truncated potentials jump at the cutoff,
    so a pair crossing it inside the difference stencil would dominate the
    finite difference (at 500,000 molecules about 20 pairs lie within 1e-6
    of rc)."""
    from scipy.spatial import cKDTree

    delta = np.random.default_rng(seed).normal(size=(s.n, 3))
    r = np.mod(s.r, s.box)
    r[r >= s.box] -= s.box[np.nonzero(r >= s.box)[1]]
    tree = cKDTree(r, boxsize=s.box)
    for rc in {float(s.rc), float(s.rc_lj)}:
        pairs = tree.query_pairs(rc + eps, output_type="ndarray")
        d = r[pairs[:, 0]] - r[pairs[:, 1]]
        d -= s.box * np.round(d / s.box)
        near = pairs[np.sqrt((d * d).sum(1)) > rc - eps]
        delta[np.unique(near.ravel())] = 0.0
    return delta


def fd_check_vjp(c, s, cot, vjp, tol=1e-6):
    delta = smooth_delta(s)
    vals = []
    for sgn in (1, -1):
        c.upload(s.r + sgn * H * delta, s.v, s.q, s.L, s.cid)
        c.update_container()
        c.traverse()
        vals.append((cot * c.download()["F"]).sum())
    c.upload(s.r, s.v, s.q, s.L, s.cid)
    c.update_container()
    c.traverse()
    fd = (vals[0] - vals[1]) / (2 * H)
    got = (vjp * delta).sum()
    assert abs(got - fd) <= tol * abs(fd), (got, fd)


@pytest.mark.parametrize("make", [small_binary, small_charged])
def test_adjoint_fd_small(make):
    s = make()
    print("case fd", make.__name__, flush=True)
    cot = np.random.default_rng(5).normal(size=(s.n, 3))
    with gpu_ctx(s) as c:
        adj = c.adjoint(cot)
        fd_check_params(c, adj)
        fd_check_vjp(c, s, cot, adj["vjp"])


def test_adjoint_fd_charges():
    s = small_charged()
    print("case fd charges", flush=True)
    with gpu_ctx(s) as c:
        adj = c.adjoint()
        q0 = np.zeros((s.ncomp, 8))
        comps = scenarios.parse_components(s.comp_desc)
        # charges as the model holds them: take them from the component records
        desc = np.asarray(s.comp_desc)
        p = 0
        for k in range(s.ncomp):
            nlj, nc, nd, nq = (int(x) for x in desc[p:p + 4])
            p += 4 + 8 * nlj
            for j in range(nc):
                q0[k, j] = desc[p + 5 * j + 4]
            p += 5 * nc + 7 * nd + 7 * nq + 3
        for k in range(s.ncomp):
            for j in range(comps[k]["nc"]):
                vals = []
                for sgn in (1, -1):
                    q = q0.copy()
                    q[k, j] += sgn * H
                    c.set_charges(q)
                    vals.append(uw(c))
                c.set_charges(q0)
                fu = (vals[0][0] - vals[1][0]) / (2 * H)
                fw = (vals[0][1] - vals[1][1]) / (2 * H)
                assert abs(adj["dU_dq"][k, j] - fu) <= 1e-6 * max(abs(fu), 1e-3), (k, j)
                assert abs(adj["dW_dq"][k, j] - fw) <= 1e-6 * max(abs(fw), 1e-3), (k, j)


def test_adjoint_fd_binary_500k():
    """The differentiation workload: binary LJ, 500,000 molecules, 60/40,
    seeded cross-term factors, rho 0.7, rc 2.5, FP64; the 6 eps/sigma pair
    parameters and a force VJP with a seeded Gaussian cotangent."""
    s = scenarios.binary_lj()
    print("case binary_lj 500k", s.n, flush=True)
    cot = np.random.default_rng(9).normal(size=(s.n, 3))
    with gpu_ctx(s) as c:
        adj = c.adjoint(cot)
        fd_check_params(c, adj)
        fd_check_vjp(c, s, cot, adj["vjp"])


def test_adjoint_rejects_multipoles():
    s = scenarios.small_mixture()
    with gpu_ctx(s) as c:
        with pytest.raises(MdxError, match="LJ and charge"):
            c.adjoint()


def test_torch_wrapper():
    torch = pytest.importorskip("torch")
    from ls1amd_torch import ForceFieldEngine, force_field

    s = small_charged()
    eng = ForceFieldEngine(s)
    try:
        o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box)
        eps = torch.tensor(eng.eps0, dtype=torch.float64, requires_grad=True)
        sig = torch.tensor(eng.sig0, dtype=torch.float64, requires_grad=True)
        q = torch.tensor(eng.q0, dtype=torch.float64, requires_grad=True)
        r = torch.tensor(s.r, dtype=torch.float64)
        U, W, F = force_field(eps, sig, q, r, eng)
        ref = o.forces(s.r, s.q, s.cid)
        u_ref, w_ref = upot_virial(ref["macro"])
        assert abs(U.item() - u_ref) <= 1e-12 * abs(u_ref)
        assert abs(W.item() - w_ref) <= 1e-12 * abs(w_ref)
        (U + 0.3 * W).backward()
        ad = o.adjoint(s.r, s.q, s.cid, np.zeros((s.n, 3)))
        iu = tuple(np.array(eng.pairs).T)
        for got, key in ((eps.grad, "eps"), (sig.grad, "sig")):
            want = ad[f"dU_d{key}"][iu] + 0.3 * ad[f"dW_d{key}"][iu]
            np.testing.assert_allclose(got.numpy(), want, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(q.grad.numpy(), ad["dU_dq"] + 0.3 * ad["dW_dq"], rtol=1e-10,
                                   atol=1e-12)
        # positions: loss = c . F + 0.5 U  ->  grad = vjp(c) - 0.5 F
        rr = torch.tensor(s.r, dtype=torch.float64, requires_grad=True)
        cot = np.random.default_rng(6).normal(size=(s.n, 3))
        U2, _, F2 = force_field(eps.detach(), sig.detach(), q.detach(), rr, eng)
        ((F2 * torch.tensor(cot)).sum() + 0.5 * U2).backward()
        ad2 = o.adjoint(s.r, s.q, s.cid, cot)
        want = ad2["vjp"] - 0.5 * ref["F"]
        np.testing.assert_allclose(rr.grad.numpy(), want, rtol=0, atol=1e-10 * np.abs(want).max())
        # what the adjoint does not provide raises
        e3 = eps.detach().clone().requires_grad_(True)
        _, _, F3 = force_field(e3, sig.detach(), q.detach(), r, eng)
        with pytest.raises(NotImplementedError):
            F3.sum().backward()
    finally:
        eng.close()


def test_torch_fit_recovers_parameters():
    """A force-field fit: Adam on (eps, sigma) of the binary mixture to match
    Upot and virial of a reference parameter set; the loss falls by orders
    of magnitude."""
    torch = pytest.importorskip("torch")
    from ls1amd_torch import ForceFieldEngine, force_field

    s = small_binary()
    eng = ForceFieldEngine(s)
    try:
        r = torch.tensor(s.r, dtype=torch.float64)
        q = torch.tensor(eng.q0, dtype=torch.float64)
        e_true = torch.tensor(eng.eps0)
        s_true = torch.tensor(eng.sig0)
        with torch.no_grad():
            U0, W0, _ = force_field(e_true, s_true, q, r, eng)
        eps = (e_true * 1.05).clone().requires_grad_(True)
        sig = s_true.clone()
        opt = torch.optim.Adam([eps], lr=2e-3)
        losses = []
        for _ in range(150):
            opt.zero_grad()
            U, W, _ = force_field(eps, sig, q, r, eng)
            loss = ((U - U0) / U0) ** 2 + ((W - W0) / W0) ** 2
            loss.backward()
            opt.step()
            losses.append(loss.item())
        assert losses[-1] < 1e-3 * losses[0], (losses[0], losses[-1])
    finally:
        eng.close()
