"""Pins the oracle's cell geometry and cutoff masks with checks that do not
come from the oracle itself.

* Cell count per dimension: LinkedCells' ctor takes floor(L / rc) cells of
  edge L / n (LinkedCells.cpp:48-57); LinkedCells::rebuild does the same with
  the cutoff rounded to float (`float rc = _cutoffRadius / _cellsInCutoff`,
  LinkedCells.cpp:148-151). The invariant that fixes n: the edge L / n is at
  least the effective cutoff and n is the largest such count. It is
  evaluated here by a search over n, with the effective cutoff rc (ctor) or
  float32(rc) (rebuild), on boxes chosen so that floor(L / rc) and
  floor(L / float(rc)) differ, so a dropped float rounding fails.
* Point location: a containment brute force against the geometry derived
  above (not the oracle's).
* Split cutoffs (CellProcessor(rc, rcLJ), CellProcessor.h:37-39): LJ site
  pairs are masked by rcLJ, the electrostatic ones by rc, both on the
  molecule-centre distance (VectorizedCellProcessor.cpp:965-966). Two
  molecules at a centre distance between rcLJ and rc must see the charge
  pairs only; the expected force is summed here from single-site-pair
  bodies (ox_site_pair, pinned by finite differences in test_oracle.py).
"""
import math

import numpy as np
import pytest

import scenarios
from oracle_lib import Oracle, macro_to_upot_virial

F32 = lambda x: float(np.float32(x))  # noqa: E731

# (box, rc): boxes where the rebuild's float cutoff changes the count, one
# with a single rebuild cell in x, and plain ones
BOXES = [
    ((25.9999995, 13.0, 26.0), 2.6),
    ((10 * F32(2.6), 7.8, 9.1), 2.6),
    ((5.1999999, 7.8, 11.0), 2.6),
    ((4.0, 8.0, 11.0), 2.6),
    ((10.0, 7.3, 12.1), 2.5),
    ((37.6386,) * 3, 2.5),
    ((125.99210498948729,) * 3, 3.0),
]


def expected_cells(L, rc_eff):
    """Largest n >= 1 with L / n >= rc_eff (0 if none)."""
    n = int(L / rc_eff) + 2
    while n > 0 and not (L / n >= rc_eff):
        n -= 1
    return n


def expected_geometry(box, rc, mode):
    rc_eff = F32(rc) if mode else rc
    n = np.array([expected_cells(L, rc_eff) for L in box])
    if mode == 0:
        n = np.maximum(n, 1)
    return n


@pytest.mark.parametrize("box,rc", BOXES)
@pytest.mark.parametrize("mode", [0, 1])
def test_cell_count_invariant(box, rc, mode):
    box = np.array(box)
    n = expected_geometry(box, rc, mode)
    o = Oracle(1, scenarios.comp_1clj(), None, 1.0, rc, rc, box, periodic=True, cell_mode=mode)
    cpd, cl, hmin, hmax = o.cell_geometry()
    np.testing.assert_array_equal(cpd - 2, n)
    np.testing.assert_array_equal(cl, box / n)
    np.testing.assert_array_equal(hmin, -cl)
    np.testing.assert_array_equal(hmax, box + cl)
    rc_eff = F32(rc) if mode else rc
    assert np.all(cl >= rc_eff) and np.all(box / (n + 1) < rc_eff)


def test_float_cutoff_changes_the_count():
    """On the first box the two modes differ in x (9 vs 10 cells): a
    restatement that ignored the float cutoff fails here."""
    box = np.array(BOXES[0][0])
    assert math.floor(box[0] / 2.6) == 9 and math.floor(box[0] / F32(2.6)) == 10
    for mode, nx in ((0, 9), (1, 10)):
        o = Oracle(1, scenarios.comp_1clj(), None, 1.0, 2.6, 2.6, box, periodic=True,
                   cell_mode=mode)
        assert o.cell_geometry()[0][0] - 2 == nx


def test_rebuild_without_a_cell_is_refused():
    """rebuild exits when no cell of edge float(rc) fits (LinkedCells.cpp:160-164)."""
    with pytest.raises(ValueError):
        Oracle(1, scenarios.comp_1clj(), None, 1.0, 2.6, 2.6, np.array([2.5, 8.0, 8.0]),
               periodic=True, cell_mode=1)


@pytest.mark.parametrize("box,rc", BOXES[:5])
@pytest.mark.parametrize("mode", [0, 1])
def test_cell_index_containment(box, rc, mode):
    """getCellIndexOfPoint (LinkedCells.cpp:870-923): every point lies in
    the cell it is assigned to, with the geometry derived above."""
    box = np.array(box)
    n = expected_geometry(box, rc, mode)
    if mode == 0 and np.any(n < 2):
        pytest.skip("the ctor refuses fewer than 2 cells")
    cpd, cl = n + 2, box / n
    o = Oracle(1, scenarios.comp_1clj(), None, 1.0, rc, rc, box, periodic=True, cell_mode=mode)
    rng = np.random.default_rng(17)
    pts = rng.uniform(-cl, box + cl, size=(3000, 3))
    k = rng.integers(0, cpd - 1, size=(500, 3))
    grid = k * cl
    pts = np.vstack([pts, grid, np.nextafter(grid, np.inf), np.nextafter(grid, -np.inf),
                     np.zeros((1, 3)), np.nextafter(box, 0)[None]])
    idx = o.cell_index(pts)

    def border(i, d):
        if i == 0:
            return -cl[d]
        if i == 1:
            return 0.0
        if i == cpd[d] - 1:
            return box[d]
        if i == cpd[d]:
            return box[d] + cl[d]
        return i * cl[d] + (-cl[d])

    for p, ci in zip(pts, idx):
        ijk = (ci % cpd[0], (ci // cpd[0]) % cpd[1], ci // (cpd[0] * cpd[1]))
        for d in range(3):
            # CellBorderAndFlagManager::getBorder (CellBorderAndFlagManager.h:
            # 134-152) values, compared exactly: the reference's +-1 fix-up
            # guarantees border(i) <= x < border(i + 1)
            assert border(ijk[d], d) <= p[d] < border(ijk[d] + 1, d), (p, ijk, d)


def _site_positions(comp, r):
    """Identity orientation: site = centre + body offset."""
    lj = [r + np.array(s[:3]) for s in comp["lj"]]
    ch = [(r + np.array(s[:3]), s[4]) for s in comp["ch"]]
    return lj, ch


@pytest.mark.parametrize("dist,rc_lj,rc", [(2.7, 2.5, 3.0), (2.4, 2.5, 3.0), (3.1, 2.5, 3.0)])
def test_split_cutoff_two_waters(dist, rc_lj, rc):
    desc = np.array(scenarios.comp_water(rc_shift=rc_lj), np.float64)
    comp = scenarios.parse_components(desc)[0]
    r = np.array([[10.0, 10.0, 10.0], [10.0 + dist * 0.6, 10.0 + dist * 0.8, 10.0]])
    q = np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))
    cid = np.zeros(2, np.int32)
    o = Oracle(1, desc, None, 1.0, rc, rc_lj, np.full(3, 30.0), periodic=False)
    got = o.forces(r, q, cid)
    eps24, sig2, shift6 = (t[0, 0] for t in o.lj_tables())
    (lja, cha), (ljb, chb) = _site_positions(comp, r[0]), _site_positions(comp, r[1])
    F = np.zeros(3)
    u = 0.0
    nlj = ncc = 0
    if dist < rc_lj:
        for a in lja:
            for b in ljb:
                out = Oracle.site_pair(0, a, np.zeros(3), 0.0, b, np.zeros(3), 0.0, eps24, sig2,
                                       shift6)
                F += out[:3]
                u += out[9] / 6.0
                nlj += 1
    if dist < rc:
        for a, qa in cha:
            for b, qb in chb:
                out = Oracle.site_pair(1, a, np.zeros(3), qa, b, np.zeros(3), qb)
                F += out[:3]
                u += out[9]
                ncc += 1
    np.testing.assert_allclose(got["F"][0], F, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(got["F"][1], -F, rtol=1e-13, atol=1e-15)
    upot, _ = macro_to_upot_virial(got["macro"])
    assert abs(upot - u) <= 1e-13 * max(abs(u), 1e-300)
    assert got["counters"][1] == nlj and got["counters"][2] == ncc


def test_split_cutoff_periodic_vs_brute_force():
    """rcLJ = 0.8 rc on the multi-site mixture: the C08 walk with halo copies
    against the O(N^2) minimum-image sum."""
    s = scenarios.small_mixture(seed=23)
    o = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, 0.8 * s.rc, s.box, periodic=True)
    a = o.forces(s.r, s.q, s.cid)
    b = o.forces(s.r, s.q, s.cid, brute=True)
    for k in ("F", "M", "Vi"):
        assert np.abs(a[k] - b[k]).max() < 1e-12 * np.abs(b[k]).max()
    full = Oracle(s.ncomp, s.comp_desc, s.mix, s.eps_rf, s.rc, s.rc, s.box,
                  periodic=True).forces(s.r, s.q, s.cid)
    # the split removes LJ pairs only
    assert a["counters"][1] < full["counters"][1]
    np.testing.assert_array_equal(a["counters"][2:], full["counters"][2:])
