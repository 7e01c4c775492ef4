"""Very dense cells (about 740 molecules each on a 3 x 3 x 3 grid): the re-bin
ranks each cell's members against the block's cell range staged in LDS when
that range fits (1024 entries) and against global memory when it does not;
the pair kernels walk neighbourhoods far larger than one LDS chunk. Cell ids
bit-exact and FP64 forces / virials / sums within 1e-12 of the oracle, the
FP32 form within the 1e-5 RMS bound, three steps within 1e-11."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context, upot_virial
from oracle_lib import Oracle, macro_to_upot_virial

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense_system(n=20000, L=21.0, rc=7.0, seed=81):
    box = np.full(3, L)
    r = scenarios.rsa_positions(seed, [((0, 0, 0), tuple(box), n)], box, 0.3)
    desc = np.array(scenarios.comp_1clj(sigma=0.3, rc_shift=rc, shifted=True), np.float64)
    rng = np.random.default_rng(seed)
    v = rng.normal(scale=0.5, size=(n, 3))
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    return desc, box, rc, r, v, q, np.zeros((n, 3)), np.zeros(n, np.int32)


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_dense_cells_vs_oracle():
    say("dense cells")
    desc, box, rc, r, v, q, L, cid = dense_system()
    o = Oracle(1, desc, None, 1.0, rc, rc, box, periodic=True)
    ref = o.forces(r, q, cid)
    u_ref, w_ref = macro_to_upot_virial(ref["macro"])
    with Context(0) as c:
        c.setup(1, desc, None, 1.0, box, rc, rc, True, 0, PREC_FP64)
        c.upload(r, v, q, L, cid)
        out = c.forces()
        ids = c.cell_ids()
    np.testing.assert_array_equal(ids, o.cell_index(r))
    for key in ("F", "Vi"):
        assert rel_err(out[key], ref[key]) < 1e-12, key
    u, w = upot_virial(out["macro"])
    assert abs(u - u_ref) < 1e-12 * abs(u_ref) and abs(w - w_ref) < 1e-12 * abs(w_ref)
    with Context(0) as c:
        c.setup(1, desc, None, 1.0, box, rc, rc, True, 0, PREC_FP32_FP64ACC)
        c.upload(r, v, q, L, cid)
        out32 = c.forces()
    d = out32["F"] - ref["F"]
    assert np.sqrt((d * d).sum() / (ref["F"] ** 2).sum()) < 1e-5


def test_dense_cells_steps_vs_oracle():
    say("dense cells steps")
    desc, box, rc, r, v, q, L, cid = dense_system(n=12000, seed=82)
    o = Oracle(1, desc, None, 1.0, rc, rc, box, periodic=True)
    ref = o.simulate(3, 0.0005, r, v, q, L, cid)
    with Context(0) as c:
        c.setup(1, desc, None, 1.0, box, rc, rc, True, 0, PREC_FP64)
        c.upload(r, v, q, L, cid)
        c.update_container()
        c.traverse()
        c.step(3, 0.0005, use_graph=True)
        out = c.download()
    d = out["r"] - ref["r"]
    d -= np.round(d / box) * box
    assert np.abs(d).max() < 1e-11 * box.max()
    assert rel_err(out["v"], ref["v"]) < 1e-11
