"""GPU: the forward FP32 V-cycles run the finest level on a colour-split
copy (even cells of every row, then odd cells) by default.  The split
kernels do the arithmetic of the natural-layout ones in the same order, so
a mixed solve with PP_SMOOTH_LINE and one with PP_SMOOTH_LINE_NATURAL agree
bit for bit, residual histories and iteration counts included; so do the
hipGraph-replayed preconditioner and a second geometry on the same
context."""
import numpy as np
import pytest

import cases
from flow_structure_interaction_fem_ibm_amd import pp

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[case] {msg}", flush=True)


def both_layouts(s, b, **kw):
    out = {}
    for sm in (pp.PP_SMOOTH_LINE, pp.PP_SMOOTH_LINE_NATURAL):
        p = np.zeros(b.shape)
        st = s.solve(p, b, precision=pp.PP_PREC_MIXED, smoother=sm, **kw)
        assert st["status"] == "PP_OK", (sm, st["status"], st["rel_l2"])
        out[sm] = (p, st)
    return out


def assert_same(out):
    (p0, s0), (p1, s1) = out[pp.PP_SMOOTH_LINE], out[pp.PP_SMOOTH_LINE_NATURAL]
    assert s0["outer_iterations"] == s1["outer_iterations"]
    assert s0["vcycles"] == s1["vcycles"]
    assert s0["history"] == s1["history"]
    assert np.array_equal(p0, p1)


# (label, sphere_case arguments): x-line chunks of 2 (32 cells), 4 (64), 6
# (96 and 130) and 8 (256) cells per lane; y/z segments of 1, 2, 4 and 8
# cells (16-, 24-, 48- and 192-cell lines) and 64 segments (512-cell lines,
# 16-line tiles of 1024 threads); 130x98x82 halves to odd counts below
CASES = [
    ("96x64x48 stretched sphere", dict(n=96, ny=64, nz=48, r=0.2, stretch=1.5)),
    ("130x98x82 stretched sphere", dict(n=130, ny=98, nz=82, r=0.23, c=(0.45, 0.5, 0.55), stretch=1.2)),
    ("256x192x192 stretched sphere", dict(n=256, ny=192, nz=192, r=0.2, c=(0.4, 0.5, 0.5), stretch=2.0)),
    ("64x24x16 stretched sphere", dict(n=64, ny=24, nz=16, r=0.3, stretch=1.2)),
    ("32x512x64 stretched sphere", dict(n=32, ny=512, nz=64, r=0.25, stretch=1.5)),
]


@pytest.mark.parametrize("label,kw", CASES, ids=[c[0].split()[0] for c in CASES])
def test_split_and_natural_layouts_agree_bitwise(label, kw):
    announce(f"split vs natural finest level, {label}")
    xn, yn, zn, ib, gs, _ = cases.sphere_case(**kw)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    b = cases.random_field(ib.shape, 11, ib)
    out = both_layouts(s, b, rel_tol=1e-10, max_outer=100)
    assert_same(out)
    assert out[pp.PP_SMOOTH_LINE][1]["rel_l2"] <= 1e-10


def test_split_layout_with_graph_and_geometry_change():
    announce("split layout: hipGraph replay and a second geometry, 64x48x40")
    xn, yn, zn, ib, gs, _ = cases.sphere_case(64, r=0.2, stretch=1.5, ny=48, nz=40)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    b = cases.random_field(ib.shape, 3, ib)
    ref = both_layouts(s, b, rel_tol=1e-10, max_outer=100)
    assert_same(ref)
    for rep in range(3):
        p = np.zeros(ib.shape)
        st = s.solve(p, b, precision=pp.PP_PREC_MIXED, rel_tol=1e-10, max_outer=100, use_graph=1)
        assert st["history"] == ref[pp.PP_SMOOTH_LINE][1]["history"], rep
        assert np.array_equal(p, ref[pp.PP_SMOOTH_LINE][0]), rep
    # a moved body: new flags, split flags and ghost offsets
    ib2, _, gs2, _ = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid((0.55, 0.45, 0.5), (0.2, 0.2, 0.2), 0.0, 48)],
                                       mode=pp.PP_GEOM_ANALYTIC)
    s.set_geometry(ib2, gs2)
    b2 = cases.random_field(ib.shape, 4, ib2)
    assert_same(both_layouts(s, b2, rel_tol=1e-10, max_outer=100))


def test_natural_only_grid_still_solves():
    announce("48x40x36: x-line chunks of 3 cells keep the natural layout")
    xn, yn, zn, ib, gs, _ = cases.sphere_case(48, r=0.2, stretch=1.5, ny=40, nz=36)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    b = cases.random_field(ib.shape, 5, ib)
    assert_same(both_layouts(s, b, rel_tol=1e-10, max_outer=100))


@pytest.mark.parametrize("label,kw", CASES[:2], ids=[c[0].split()[0] for c in CASES[:2]])
def test_split_and_natural_adjoints_agree_bitwise(label, kw):
    announce(f"adjoint: split vs natural finest level, {label}")
    xn, yn, zn, ib, gs, _ = cases.sphere_case(**kw)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    w = cases.random_field(ib.shape, 21, ib)
    out = {}
    for sm in (pp.PP_SMOOTH_LINE, pp.PP_SMOOTH_LINE_NATURAL):
        lam, st = s.adjoint(w, precision=pp.PP_PREC_MIXED, smoother=sm, rel_tol=1e-10, max_outer=100)
        assert st["status"] == "PP_OK", (sm, st["status"], st["rel_l2"])
        out[sm] = (lam, st)
    (l0, s0), (l1, s1) = out[pp.PP_SMOOTH_LINE], out[pp.PP_SMOOTH_LINE_NATURAL]
    assert s0["outer_iterations"] == s1["outer_iterations"] and s0["vcycles"] == s1["vcycles"]
    assert s0["history"] == s1["history"]
    assert np.array_equal(l0, l1)
