"""The Fortran ISO_C_BINDING module compiles with flang, links against the
library, and its BIND(C) types match the C structs (sizes, and the values
pp_default_opts writes).  CPU only: no call reaches a GPU."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from flow_structure_interaction_fem_ibm_amd import pp

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(pp.__file__)
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_module_matches_the_c_abi(tmp_path):
    lib = pp.lib()  # builds nothing; the library must already be built
    so = os.path.join(PKG, "libpp_hip.so")
    mod = os.path.join(PKG, "fortran", "pp_hip_mod.F90")
    prog = os.path.join(HERE, "fortran", "layout_check.F90")
    exe = str(tmp_path / "layout_check")
    r = subprocess.run([FLANG, "-O1", "-J", str(tmp_path), mod, prog, "-o", exe, so, f"-Wl,-rpath,{PKG}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {l.split(" ", 1)[0]: l.split(" ", 1)[1].split() for l in r.stdout.splitlines() if " " in l}
    o = pp.default_opts()
    f = out["opts"]
    assert int(f[0]) == o.precision and float(f[1]) == pytest.approx(o.rel_tol) and float(f[2]) == o.abs_linf_tol
    assert [int(v) for v in f[3:10]] == [o.max_outer, o.cycles_per_outer, o.nu_pre, o.nu_post, o.mean_remove,
                                         o.project_rhs, o.use_graph]
    assert float(f[10]) == pytest.approx(o.omega) and int(f[11]) == o.profile and int(f[12]) == o.krylov
    assert int(f[13]) == o.nu_coarse and int(f[14]) == o.smoother
    # a 2x2x2 grid is rejected before any device call, with the library's message
    assert int(out["create"][0]) == pp.PP_BAD_ARG
    assert "at least 4 cells" in " ".join(out["error"])
    sizes = [int(v) for v in out["sizes"]]
    assert sizes == [C.sizeof(pp.SolveOpts), C.sizeof(pp.GridDesc), C.sizeof(pp.SolveStats), C.sizeof(pp.Geometry)]
    assert int(out["bodysize"][0]) == C.sizeof(pp.BodyDesc)
    assert int(out["adstatsize"][0]) == C.sizeof(pp.ADStats)
    assert int(out["surfsize"][0]) == C.sizeof(pp.SurfaceDesc)
    # the same geometry through the ctypes binding
    xn = 1.2 * np.arange(25) / 24.0
    body = pp.Ellipsoid((0.61, 0.58, 0.6), (0.3, 0.2, 0.22), 20.0, 40)
    ib, _, gs, ex = pp.build_geometry(xn, xn, xn, [body], mode=pp.PP_GEOM_TRIANGULATED, probe_norm=2.0)
    n = len(gs.ig)
    assert n > 0 and [int(v) for v in out["geom"]] == [pp.PP_OK]
    assert [int(v) for v in out["counts"]] == [pp.PP_OK, n, int((ib[1:-1, 1:-1, 1:-1] == 1).sum()), 0]
    g = out["get"]
    assert int(g[0]) == pp.PP_OK and int(g[1]) == int(ib.astype(np.int64).sum())
    # Fortran indices are the same 1-based cell indices the C side writes
    assert [int(v) for v in g[2:8]] == [int(a.astype(np.int64).sum()) for a in (gs.ig, gs.jg, gs.kg, gs.ic, gs.jc, gs.kc)]
    assert float(g[8]) == pytest.approx(float(gs.coeff.sum()), rel=1e-14, abs=1e-12)
    assert float(g[9]) == pytest.approx(float(gs.probe.sum()), rel=1e-14)
    assert [int(v) for v in out["elems"]] == [pp.PP_OK, int(ex["body"].sum()), int(ex["elem"].sum())]
    d = out["dirichlet"]
    assert int(d[0]) == pp.PP_OK and float(d[1]) == pytest.approx(float(ex["coeffD"].sum()), rel=1e-14)
    assert int(out["destroy"][0]) == pp.PP_OK
