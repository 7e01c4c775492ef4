"""mdx_adjoint_noalloc (the adjoint on the library's per-device scratch array,
no device allocation) against mdx_adjoint through the C ABI.

Both run the same kernels in the same launch shape and reduction order on
buffers of the same layout, so every output is bit-identical. Two calls in a
row, the second with a smaller system (the scratch keeps the first call's
contents and every buffer must be rewritten), then the first system again;
and the refusal of a state whose buffers exceed the scratch array.
"""
import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP64, Context, MdxError
from test_oracle_adjoint import small_charged

pytestmark = pytest.mark.gpu

KEYS = ("dU_deps", "dU_dsig", "dW_deps", "dW_dsig", "dU_dq", "dW_dq", "vjp")


def ctx_for(s):
    c = Context(0)
    c.setup_system(s, periodic=True, precision=PREC_FP64)
    c.update_container()
    c.traverse()
    return c


def check_same(a, b):
    for k in KEYS:
        if k in a or k in b:
            assert np.array_equal(a[k], b[k]), k


def test_noalloc_equals_adjoint_two_calls():
    big = scenarios.binary_lj(n=20000)
    small = small_charged()
    cot_big = np.random.default_rng(5).normal(size=(big.n, 3))
    cot_small = np.random.default_rng(6).normal(size=(small.n, 3))
    print("case noalloc binary_lj", big.n, "then charged", small.n, flush=True)
    with ctx_for(big) as c1, ctx_for(small) as c2:
        first = c1.adjoint(cot_big, noalloc=True)
        second = c2.adjoint(cot_small, noalloc=True)  # smaller, on the same scratch
        ref1 = c1.adjoint(cot_big)
        ref2 = c2.adjoint(cot_small)
        check_same(first, ref1)
        check_same(second, ref2)
        assert np.abs(second["dU_dq"]).max() > 0
        # the first system again after the smaller one, and without a cotangent
        check_same(c1.adjoint(cot_big, noalloc=True), ref1)
        p_only = c1.adjoint(None, noalloc=True)
        ref_p = c1.adjoint(None)
        check_same(p_only, ref_p)


def test_noalloc_refuses_states_beyond_the_scratch():
    """1.49 M one-centre molecules: the VJP buffers (48 B per molecule) do not
    fit into the 64 MiB scratch; the parameter derivatives alone do."""
    s = scenarios.lj_fcc(ncell=72)
    print("case noalloc capacity N", s.n, flush=True)
    cot = np.zeros((s.n, 3))
    with ctx_for(s) as c:
        with pytest.raises(MdxError, match="scratch"):
            c.adjoint(cot, noalloc=True)
        check_same(c.adjoint(None, noalloc=True), c.adjoint(None))
