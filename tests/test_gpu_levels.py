"""GPU: the V-cycle's single-level operators through pp_level_data and
pp_level_apply against the numpy restatement of tests/levels.py (pinned by
tests/test_levels.py): coarse iblank and face flags bit-exact, level metrics,
the level residual, restriction (the general kernel and the fused
residual + by-2 restriction the V-cycle runs), prolongation + correction
and one zebra line pass set, in FP64 and in FP32 (the mixed solve's
V-cycles) on every level.

Grids: 48x40x32 (every level halves in x, y and z: the fused by-2 kernels),
128x48x18 (18 -> 9 -> 4: the last coarse cell of an odd count spans three
fine cells; the general kernels) and 65x49x41 (odd in every direction)."""
import numpy as np
import pytest

from flow_structure_interaction_fem_ibm_amd import pp
from levels import Hierarchy

pytestmark = pytest.mark.gpu

F64, F32 = pp.PP_PREC_FP64, pp.PP_PREC_MIXED
TOL = {F64: 1e-12, F32: 2e-5}


def announce(msg):
    print(f"[case] {msg}", flush=True)


def setup(nx, ny, nz):
    xn = pp.tanh_nodes(nx, 1.0, 1.3)
    yn = pp.tanh_nodes(ny, 1.0, 1.0)
    zn = pp.tanh_nodes(nz, 1.0, 1.4)
    ib, fr, gs, ex = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid((0.5, 0.47, 0.53), (0.23, 0.2, 0.25), 30.0, 24)])
    ghost = np.zeros(ib.shape, bool)
    ghost[gs.kg, gs.jg, gs.ig] = True
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    assert s.level_info() == H.dims
    return s, H


GRIDS = [(48, 40, 32), (128, 48, 18), (65, 49, 41)]


def close(got, ref, tol):
    scale = max(np.abs(ref).max(), 1e-300)
    return np.abs(got - ref).max() <= tol * scale


@pytest.mark.parametrize("grid", GRIDS)
def test_level_flags_and_metrics(grid):
    announce(f"level flags and metrics {grid}")
    s, H = setup(*grid)
    if grid == (128, 48, 18):
        assert H.dims[2] == (32, 12, 4)
    for l in range(H.nlev):
        fl, m = s.level_data(l)
        assert np.array_equal(fl, H.flags[l]), l
        for d in range(3):
            for k in ("w", "xc", "cm", "cp"):
                assert np.allclose(m[k][d], H.metrics[l][d][k], rtol=1e-14, atol=0), (l, d, k)


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("prec", [F64, F32])
def test_level_residual_restriction_prolongation(grid, prec):
    announce(f"level residual / restriction / prolongation {grid} prec={prec}")
    s, H = setup(*grid)
    rng = np.random.default_rng(4)
    tol = TOL[prec]
    for l in range(H.nlev):
        shp = H.dims[l][::-1]
        gh = (H.flags[l] & 2) != 0
        e = rng.standard_normal(shp) * (~H.blank[l] | gh)
        rhs = rng.standard_normal(shp) * ~H.blank[l]
        if prec == F32:   # the values the FP32 kernels see
            e, rhs = e.astype(np.float32).astype(np.float64), rhs.astype(np.float32).astype(np.float64)
        res = s.level_apply(pp.PP_LOP_RESIDUAL, l, e, rhs, precision=prec)
        ref = H.residual(l, e, rhs)
        assert close(res, ref, 10 * tol), ("residual", l)
        assert not res[H.blank[l]].any()
        if l == H.nlev - 1:
            break
        r = rng.standard_normal(shp)
        got = s.level_apply(pp.PP_LOP_RESTRICT, l, r, precision=prec)
        assert close(got, H.restrict(l, r), tol), ("restrict", l)
        # the V-cycle's own restriction step (fused by-2 kernel where it applies)
        got = s.level_apply(pp.PP_LOP_RESRESTRICT, l, e, rhs, precision=prec)
        assert close(got, H.restrict(l, H.residual(l, e, rhs)), 10 * tol), ("resrestrict", l)
        ec = rng.standard_normal(H.dims[l + 1][::-1])
        ef = rng.standard_normal(shp)
        got = s.level_apply(pp.PP_LOP_PROLONG, l, ec, ef, precision=prec)
        assert close(got, H.prolong(l, ec, ef), tol), ("prolong", l)


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("prec", [F64, F32])
def test_line_pass_every_level(grid, prec):
    announce(f"zebra line pass set on every level {grid} prec={prec}")
    s, H = setup(*grid)
    rng = np.random.default_rng(8)
    tol = 1e-10 if prec == F64 else 2e-4
    for l in range(H.nlev - 1):
        shp = H.dims[l][::-1]
        gh = (H.flags[l] & 2) != 0
        e = rng.standard_normal(shp) * (~H.blank[l] | gh)
        rhs = rng.standard_normal(shp) * ~H.blank[l]
        if prec == F32:
            e, rhs = e.astype(np.float32).astype(np.float64), rhs.astype(np.float32).astype(np.float64)
        for omega in (1.0, 1.3):
            got = s.level_apply(pp.PP_LOP_LINE_PASS, l, e, rhs, precision=prec, omega=omega)
            ref = H.line_pass(l, e, rhs, omega)
            assert close(got, ref, tol), (l, omega, np.abs(got - ref).max() / np.abs(ref).max())


def test_level_apply_rejects_bad_arguments():
    s, H = setup(48, 40, 32)
    z = np.zeros(H.dims[0][::-1])
    with pytest.raises(pp.PPError):
        s.level_apply(9, 0, z, z)
    with pytest.raises(pp.PPError):
        s.level_apply(pp.PP_LOP_PROLONG, H.nlev - 1, z, z)     # no coarser level
    with pytest.raises(pp.PPError):
        s.level_apply(pp.PP_LOP_LINE_PASS, 0, z, z, omega=2.5)
