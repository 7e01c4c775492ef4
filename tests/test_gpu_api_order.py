"""Calls that read the cell lists right after an upload (no
mdx_update_container first) bin the state themselves: traversal, autotuner
and cell ids give what the documented order (upload, update, traverse)
gives. A re-bin builds the site records of the context's precision only;
a traversal in the other precision, or the FP64 adjoint, on the same binned
state adds the missing records and gives what a context of that precision
gives."""
import sys

import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context
from oracle_lib import Oracle
from test_oracle_adjoint import small_charged

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("sysname", ["small_lj", "small_mixture"])
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_FP64ACC])
def test_traverse_right_after_upload(sysname, precision):
    say("api order", sysname, precision)
    s = getattr(scenarios, sysname)(seed=91)
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=precision)
        c.update_container()
        m_ref = c.traverse()
        ref = c.download()
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=precision)
        m = c.traverse()
        out = c.download()
    np.testing.assert_array_equal(m, m_ref)
    for key in ("F", "M", "Vi"):
        assert np.array_equal(out[key], ref[key]), key


def test_autotune_and_cell_ids_right_after_upload():
    say("api order autotune")
    s = scenarios.small_lj(seed=92)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc_lj, s.box, periodic=True)
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=PREC_FP64)
        ids = c.cell_ids()
    np.testing.assert_array_equal(ids, o.cell_index(s.r))
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=PREC_FP64)
        (cpb, bpx), ms = c.autotune(reps=1)
        assert 1 <= cpb <= 8 and ms > 0
        c.traverse()
        out = c.download()
    ref = o.forces(s.r, s.q, s.cid)
    assert np.abs(out["F"] - ref["F"]).max() < 1e-12 * np.abs(ref["F"]).max()


@pytest.mark.parametrize("first", [PREC_FP64, PREC_FP32_FP64ACC])
def test_precision_switch_without_rebin(first):
    say("api order precision switch, first", first)
    s = scenarios.small_mixture(seed=93)
    ref = {}
    for p in (PREC_FP64, PREC_FP32_FP64ACC):
        with Context(0) as c:
            c.setup_system(s, periodic=True, precision=p)
            c.update_container()
            ref[p] = (c.traverse(), c.download())
    other = PREC_FP32_FP64ACC if first == PREC_FP64 else PREC_FP64
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=first)
        c.update_container()
        for p in (first, other, first, other):
            c.set_precision(p)
            m = c.traverse()
            out = c.download()
            np.testing.assert_array_equal(m, ref[p][0])
            for key in ("F", "M", "Vi"):
                assert np.array_equal(out[key], ref[p][1][key]), (p, key)


def test_adjoint_in_fp32_context():
    say("api order adjoint after an FP32 re-bin")
    s = small_charged()
    cot = np.random.default_rng(5).normal(size=(s.n, 3))
    got = {}
    for p in (PREC_FP64, PREC_FP32_FP64ACC):
        with Context(0) as c:
            c.setup_system(s, periodic=True, precision=p)
            c.update_container()
            c.traverse()
            got[p] = c.adjoint(cot)
    for k, v in got[PREC_FP64].items():
        assert np.array_equal(got[PREC_FP32_FP64ACC][k], v), k


def test_adjoint_after_graph_steps_in_fp32_context():
    """Graph-launched steps re-bin on the device; the FP64 records an earlier
    adjoint added belong to the old binning and must be rebuilt for the
    next adjoint."""
    say("api order adjoint after graph-launched FP32 steps")
    s = small_charged()
    rng = np.random.default_rng(6)
    s.v = rng.normal(scale=0.5, size=(s.n, 3))
    cot = rng.normal(size=(s.n, 3))
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=PREC_FP32_FP64ACC)
        c.update_container()
        c.traverse()
        c.step(2, 0.002, use_graph=True)
        c.adjoint(cot)
        c.step(3, 0.002, use_graph=True)
        got = c.adjoint(cot)
        st = c.download()
    s.r, s.v, s.q, s.L = st["r"], st["v"], st["q"], st["L"]
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=PREC_FP64)
        c.update_container()
        c.traverse()
        ref = c.adjoint(cot)
    # the fresh context normalises the downloaded q once more (last-bit
    # changes); stale records of the earlier binning differ by the steps'
    # displacements
    for k, v in ref.items():
        scale = max(np.abs(v).max(), 1e-300)
        np.testing.assert_allclose(got[k], v, rtol=0, atol=1e-9 * scale, err_msg=k)
