"""GPU: the triangulated set_boundary on the device (pp_geom_build_device,
pp_geom_build_surfaces_device) against the host build (pp_geom_build,
itself bit-exact with the oracle's restatement in its integer outputs,
tests/test_geometry_tri.py): iblank, fresh cells, the ghost list, image-point
cells, bodies, elements, intercepts, image points, normals, probe lengths,
the Neumann and Dirichlet coefficients and the fresh-cell tables are
identical bit for bit, on a sphere, two rotated bodies on a stretched grid,
a moving body with fresh cells, a caller-given surface, the 256x192x192
ellipsoid, a step of the 384x256x256 moving sphere and the 64 spheres of the
512x512x384 grid."""
import time

import numpy as np
import pytest

from flow_structure_interaction_fem_ibm_amd import configs, pp

pytestmark = pytest.mark.gpu

TRI = pp.PP_GEOM_TRIANGULATED


def announce(msg):
    print(f"[case] {msg}", flush=True)


def same(host, dev):
    ib, fr, gs, ex = host
    ib2, fr2, gs2, ex2 = dev
    assert np.array_equal(ib, ib2), int((ib != ib2).sum())
    assert np.array_equal(fr, fr2), int((fr != fr2).sum())
    assert gs.n == gs2.n
    for name in ("ig", "jg", "kg", "ic", "jc", "kc", "coeff", "probe"):
        assert np.array_equal(getattr(gs, name), getattr(gs2, name)), name
    for name in ("bi", "ip", "nrm", "body", "elem", "coeffD"):
        assert np.array_equal(ex[name], ex2[name]), name
    assert (ex["fresh"] is None) == (ex2["fresh"] is None) and ex["fresh_error"] == ex2["fresh_error"]
    if ex["fresh"] is not None:
        for k, v in ex["fresh"].items():
            assert np.array_equal(v, ex2["fresh"][k]), k


def both(xn, yn, zn, bodies, iblank_prev=None, probe_norm=2.0):
    t0 = time.perf_counter()
    host = pp.build_geometry(xn, yn, zn, bodies, mode=TRI, iblank_prev=iblank_prev, probe_norm=probe_norm)
    t1 = time.perf_counter()
    dev = pp.build_geometry(xn, yn, zn, bodies, mode=TRI, iblank_prev=iblank_prev, probe_norm=probe_norm, device=0)
    t2 = time.perf_counter()
    same(host, dev)
    return host, t1 - t0, t2 - t1


def test_sphere_uniform():
    announce("device geometry: sphere 24^3")
    xn = np.arange(25) / 24.0
    (ib, _, gs, _), th, td = both(xn, xn, xn, [pp.Ellipsoid((0.5, 0.5, 0.5), (0.3, 0.3, 0.3), 0.0, 32)])
    assert gs.n > 0 and ib.sum() > 0


def test_two_rotated_bodies_stretched():
    announce("device geometry: two rotated ellipsoids 48x40x36 stretched")
    xn, yn, zn = pp.tanh_nodes(48, 2.0, 1.3), pp.tanh_nodes(40, 1.0, 0.9), pp.tanh_nodes(36, 1.0, 1.1)
    bodies = [pp.Ellipsoid((0.6, 0.5, 0.5), (0.3, 0.18, 0.2), 25.0, 40),
              pp.Ellipsoid((1.35, 0.45, 0.52), (0.2, 0.22, 0.15), -40.0, 24)]
    (ib, _, gs, ex), th, td = both(xn, yn, zn, bodies, probe_norm=1.7)
    assert set(np.unique(ex["body"])) == {1, 2}


def test_moving_body_fresh_cells():
    announce("device geometry: moving ellipsoid with fresh cells 40x36x32")
    xn, yn, zn = pp.tanh_nodes(40, 1.0, 1.2), pp.tanh_nodes(36, 1.0, 0.9), pp.tanh_nodes(32, 1.0, 1.4)
    ib0 = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid((0.45, 0.5, 0.5), (0.22, 0.2, 0.24), 25.0, 32)], mode=TRI)[0]
    (ib, fr, gs, ex), th, td = both(xn, yn, zn, [pp.Ellipsoid((0.47, 0.5, 0.49), (0.22, 0.2, 0.24), 25.0, 32)],
                                    iblank_prev=ib0)
    assert fr.sum() > 0 and ex["fresh"] is not None and len(ex["fresh"]["i"]) == fr.sum()


def test_caller_given_surface():
    announce("device geometry: unstructured surface 32x28x24")
    xn, yn, zn = pp.tanh_nodes(32, 1.0, 1.1), pp.tanh_nodes(28, 1.0, 0.9), pp.tanh_nodes(24, 1.0, 1.2)
    xyz, tri = pp.canonical_surface(pp.Ellipsoid((0.5, 0.5, 0.5), (0.25, 0.2, 0.22), 10.0, 24))
    surf = [(xyz, tri, (0.0, 0.0, 0.0))]
    host = pp.build_geometry_surfaces(xn, yn, zn, surf)
    dev = pp.build_geometry_surfaces(xn, yn, zn, surf, device=0)
    same(host, dev)


def test_bench_ellipsoid_256x192x192():
    announce("device geometry: 256x192x192 prolate ellipsoid")
    xn, yn, zn = np.linspace(0.0, 4.0, 257), pp.tanh_nodes(192, 3.0, 2.0), pp.tanh_nodes(192, 3.0, 2.0)
    body = pp.Ellipsoid((1.5, 1.5, 1.5), (0.5, 0.25, 0.25), 20.0, 96)
    (ib, _, gs, _), th, td = both(xn, yn, zn, [body])
    print(f"[result] 256x192x192: {gs.n} ghosts, host {th * 1e3:.0f} ms, device {td * 1e3:.0f} ms", flush=True)


def test_moving_sphere_step_384x256x256():
    announce("device geometry: 384x256x256 moving sphere steps 0 and 1")
    xn, yn, zn = configs.config3_grid()
    ib0, _, _, _ = pp.build_geometry(xn, yn, zn, [configs.config3_body(0)], mode=TRI, device=0)
    (ib, fr, gs, ex), th, td = both(xn, yn, zn, [configs.config3_body(1)], iblank_prev=ib0)
    print(f"[result] 384x256x256 step 1: {gs.n} ghosts, {int(fr.sum())} fresh, host {th * 1e3:.0f} ms, "
          f"device {td * 1e3:.0f} ms", flush=True)


def test_64_spheres_512x512x384():
    announce("device geometry: 512x512x384, 64 spheres")
    xn, yn, zn = configs.config5_grid()
    (ib, _, gs, _), th, td = both(xn, yn, zn, configs.config5_bodies(xn, yn, zn))
    print(f"[result] 512x512x384 64 spheres: {gs.n} ghosts, host {th:.2f} s, device {td:.2f} s", flush=True)


def test_device_build_refuses_bad_input():
    announce("device geometry: argument checks")
    xn = np.arange(13) / 12.0
    with pytest.raises(ValueError):
        pp.build_geometry(xn, xn, xn, [pp.Ellipsoid((0.5, 0.5, 0.5), (0.3, 0.3, 0.3), 0.0, 16)], device=0)
    with pytest.raises(pp.PPError):
        pp.build_geometry(xn, xn, xn, [pp.Ellipsoid((0.5, 0.5, 0.5), (-0.3, 0.3, 0.3), 0.0, 16)], mode=TRI, device=0)
