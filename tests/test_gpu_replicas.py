"""Batched replicas (mdx_set_replicas) and pair-kernel launch shapes.

Each replica of a batch runs the same kernels on the same cells as a context
holding that system alone, so forces, torques and the Leapfrog trajectory of
every replica are bit-identical to the single-system runs (which the parity
tests pin to the FP64 oracle at 1e-12); the macroscopic sums differ only in
summation order. The pair kernels' launch shape (cells per work item,
persistent blocks) changes which block walks which cells, never what a cell
computes: results are bit-identical across shapes.
"""
import numpy as np
import pytest

import scenarios
from ls1amd import PREC_FP32_FP64ACC, PREC_FP64, Context, MdxError, upot_virial

pytestmark = pytest.mark.gpu


def small_batch():
    # uneven sizes and densities, one replica at the 2-cells-per-dimension edge
    spec = [(2000, 0.85, 0.7), (900, 0.05, 1.3), (3000, 0.4, 1.0), (500, 0.75, 0.9),
            (4000, 0.2, 1.1), (1500, 0.6, 0.8)]
    return [scenarios.ljts_box(n, rho, T, seed=100 + k) for k, (n, rho, T) in enumerate(spec)]


def single(s, prec, steps=0):
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=prec)
        c.update_container()
        mac = c.traverse()
        per = c.step(steps, s.dt, use_graph=True) if steps else None
        return c.download(), mac, per


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32_FP64ACC])
def test_batch_equals_single_systems(prec):
    reps = small_batch()
    print("case batch", prec, [s.n for s in reps], flush=True)
    steps = 10
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=prec)
        c.update_container()
        mac = c.traverse()
        f0 = c.download()
        per = c.step(steps, reps[0].dt, use_graph=True)
        f1 = c.download()
    assert mac.shape == (len(reps), 4) and per.shape == (steps, len(reps), 6)
    off = np.cumsum([0] + [s.n for s in reps])
    for k, s in enumerate(reps):
        d0, m0, _ = single(s, prec)
        d1, _, p1 = single(s, prec, steps)
        sl = slice(off[k], off[k + 1])
        assert np.array_equal(f0["F"][sl], d0["F"]), k
        assert np.array_equal(f0["Vi"][sl], d0["Vi"]), k
        np.testing.assert_allclose(mac[k], m0, rtol=1e-12, atol=1e-12 * np.abs(m0).max())
        for key in ("r", "v", "F"):
            assert np.array_equal(f1[key][sl], d1[key]), (k, key)
        np.testing.assert_allclose(per[:, k, :4], p1[:, :4], rtol=1e-11,
                                   atol=1e-11 * np.abs(p1[:, :4]).max())


@pytest.mark.parametrize("shape", [(2, 0), (4, 16), (1, 8), (8, 64), (16, 0), (16, 32)])
def test_launch_shapes_bit_identical(shape):
    s = scenarios.small_mixture(n=400)
    print("case shape", shape, flush=True)
    with Context(0) as c:
        c.setup_system(s, periodic=True, precision=PREC_FP64)
        c.update_container()
        m_ref = c.traverse()
        d_ref = c.download()
        c.set_pair_config(*shape)
        m = c.traverse()
        d = c.download()
    for key in ("F", "M", "Vi"):
        assert np.array_equal(d[key], d_ref[key]), key
    np.testing.assert_allclose(m, m_ref, rtol=1e-13, atol=1e-13 * np.abs(m_ref).max())


def test_autotune_on_batch():
    reps = small_batch()
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP32_FP64ACC)
        c.update_container()
        c.traverse()
        ref = c.download()["F"]
        (cpb, bpx), ms = c.autotune(reps=2)
        assert cpb in (1, 2, 4, 8) and bpx in (0, 512, 1024) and ms > 0
        c.traverse()
        assert np.array_equal(c.download()["F"], ref)
        c.step(3, reps[0].dt, use_graph=True)


def test_batch_rejects_single_system_calls():
    reps = small_batch()[:2]
    with Context(0) as c:
        c.setup_replicas(reps, periodic=True, precision=PREC_FP64)
        with pytest.raises(MdxError):
            c.cell_ids()
        with pytest.raises(MdxError):
            c.upload(reps[0].r, reps[0].v, reps[0].q, reps[0].L, reps[0].cid)
