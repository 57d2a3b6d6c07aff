"""CPU pins of the oracle's subtracting_average_pressure with fresh cells
(or_mean_remove + or_set_fresh, restating CORE_POISSON.F90:419-490):

  * the mean is the fluid-volume mean (closed form, fresh cells included);
  * a fresh cell whose non-fresh fluid neighbours all hold one value gets
    that value; one whose 26 neighbours are all non-fresh fluid cells on a
    uniform grid reproduces a linear field at its centre;
  * every fresh cell is the mean of its non-blank, non-fresh 3x3x3
    neighbours after the mean removal, counted by brute force on tiny
    grids with fresh cells next to each other and next to blank cells
    (the reference's loop order cannot leak: a fresh cell never reads
    another fresh cell)."""
import numpy as np
import pytest

from oracle_lib import Oracle


def tiny(n=10, seed=0):
    xn = yn = zn = np.linspace(0.0, 1.0, n + 1)
    o = Oracle(xn, yn, zn)
    shp = o.shape
    rng = np.random.default_rng(seed)
    ib = np.zeros(shp, np.int8)
    ib[3:6, 4:7, 3:7] = 1                 # a blank block
    fr = np.zeros(shp, np.int8)
    inner = np.zeros(shp, bool)
    inner[1:-2, 1:-2, 1:-2] = True
    cand = inner & (ib == 0)
    pick = rng.random(shp) < 0.15
    fr[cand & pick] = 1
    fr[6, 5, 4] = 1 if ib[6, 5, 4] == 0 else 0      # next to the block
    fr[6, 5, 5] = 1                                 # two fresh cells side by side
    o.set_iblank(ib)
    o.set_fresh(fr)
    return o, ib, fr, rng


def test_mean_is_the_fluid_volume_mean():
    o, ib, fr, rng = tiny()
    p = rng.standard_normal(o.shape)
    fl = np.zeros(o.shape, bool)
    fl[1:-2, 1:-2, 1:-2] = ib[1:-2, 1:-2, 1:-2] == 0
    m = o.mean_remove(p.copy())
    assert m == pytest.approx(p[fl].mean(), rel=1e-13)       # uniform grid: equal volumes


def test_fresh_cells_take_a_constant_neighbourhood():
    o, ib, fr, rng = tiny(seed=1)
    p = np.full(o.shape, 3.25)
    p[fr == 1] = rng.standard_normal((fr == 1).sum())          # fresh cells hold anything
    fl = np.zeros(o.shape, bool)
    fl[1:-2, 1:-2, 1:-2] = ib[1:-2, 1:-2, 1:-2] == 0
    # the outer layer is not shifted by the mean removal: give it the shifted
    # constant so that every neighbourhood holds one value
    shifted = 3.25 - p[fl].mean()
    outer = np.ones(o.shape, bool)
    outer[1:-2, 1:-2, 1:-2] = False
    p[outer] = shifted
    q = p.copy()
    m = o.mean_remove(q)
    f = fr == 1
    # interior neighbours hold 3.25 - m, outer ones the numpy-shifted value:
    # the two means differ by their summation order only
    assert np.allclose(q[f], 3.25 - m, rtol=1e-12, atol=0)
    assert np.allclose(q[f], shifted, rtol=1e-12, atol=0)
    assert not q[1:-2, 1:-2, 1:-2][ib[1:-2, 1:-2, 1:-2] == 1].any()


def test_isolated_fresh_cell_reproduces_a_linear_field():
    n = 10
    xn = np.linspace(0.0, 1.0, n + 1)
    o = Oracle(xn, xn, xn)
    ib = np.zeros(o.shape, np.int8)
    fr = np.zeros(o.shape, np.int8)
    fr[5, 4, 6] = 1
    o.set_iblank(ib)
    o.set_fresh(fr)
    xc = np.zeros(n + 3)
    xc[1:n + 1] = 0.5 * (xn[1:] + xn[:-1])
    p = 0.3 + 1.1 * xc[None, None, :] - 0.7 * xc[None, :, None] + 2.0 * xc[:, None, None]
    q = p.copy()
    m = o.mean_remove(q)
    assert q[5, 4, 6] == pytest.approx(p[5, 4, 6] - m, rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("seed", [2, 3, 4])
def test_fresh_fill_brute_force(seed):
    o, ib, fr, rng = tiny(seed=seed)
    p = rng.standard_normal(o.shape)
    q = p.copy()
    m = o.mean_remove(q)
    # reference after the mean removal: interior fluid cells p - m, blank 0,
    # the outer layer untouched
    snap = p.copy()
    I = (slice(1, -2),) * 3
    core = snap[I]
    core[:] = np.where(ib[I] == 1, 0.0, core - m)
    nz, ny, nx = (s - 3 for s in o.shape)
    for k, j, i in zip(*np.nonzero(fr)):
        if not (1 <= k <= nz and 1 <= j <= ny and 1 <= i <= nx):
            continue
        vals = [snap[kk, jj, ii] for kk in (k - 1, k, k + 1) for jj in (j - 1, j, j + 1) for ii in (i - 1, i, i + 1)
                if ib[kk, jj, ii] != 1 and fr[kk, jj, ii] != 1]
        assert q[k, j, i] == pytest.approx(sum(vals) / len(vals), rel=1e-14, abs=1e-15)
    # every other cell holds the snapshot
    other = fr != 1
    assert np.array_equal(q[other], snap[other])


def test_fresh_cell_without_fluid_neighbours_keeps_its_value():
    """A fresh cell whose 26 neighbours are all blank or fresh (a body that
    moved more than a cell in one step) has nothing to average; the
    reference divides by zero there, the oracle and the library keep the
    mean-removed value instead."""
    n = 10
    xn = np.linspace(0.0, 1.0, n + 1)
    o = Oracle(xn, xn, xn)
    ib = np.zeros(o.shape, np.int8)
    fr = np.zeros(o.shape, np.int8)
    fr[4:7, 4:7, 4:7] = 1                  # a 3x3x3 block of fresh cells
    o.set_iblank(ib)
    o.set_fresh(fr)
    rng = np.random.default_rng(6)
    p = rng.standard_normal(o.shape)
    q = p.copy()
    m = o.mean_remove(q)
    assert np.isfinite(q).all()
    assert q[5, 5, 5] == pytest.approx(p[5, 5, 5] - m, rel=1e-13, abs=1e-14)   # the enclosed cell
    # a cell on the block's face still averages its fluid neighbours
    vals = [p[k, j, i] - m for k in (3, 4, 5) for j in (3, 4, 5) for i in (3, 4, 5) if fr[k, j, i] != 1]
    assert q[4, 4, 4] == pytest.approx(sum(vals) / len(vals), rel=1e-12)
