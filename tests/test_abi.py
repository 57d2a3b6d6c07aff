"""The C-ABI library loads and exports every function include/pp_hip.h
declares (no GPU calls)."""
import ctypes
import os
import re

from flow_structure_interaction_fem_ibm_amd import pp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "pp_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_binding_set():
    assert set(declared_functions()) == set(pp.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pp.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_default_options_and_error_text():
    o = pp.default_opts()
    assert o.precision == pp.PP_PREC_MIXED and o.rel_tol == 1e-8 and o.nu_pre == 2 and o.nu_post == 2
    lib = pp.lib()
    rc = lib.pp_create(None, None)
    assert rc == pp.PP_BAD_ARG
    assert b"null" in lib.pp_last_error()


def test_create_rejects_bad_grids_without_touching_a_device():
    import numpy as np
    lib = pp.lib()
    xn = np.linspace(0, 1, 3)          # 2 cells: too small
    g = pp._grid_desc(xn, xn, xn)
    h = ctypes.c_void_p()
    assert lib.pp_create(ctypes.byref(h), ctypes.byref(g)) == pp.PP_BAD_ARG
    xn = np.linspace(0, 1, 9)
    g = pp._grid_desc(xn, xn, xn)
    g.bc[0] = 1                        # Dirichlet: not implemented
    assert lib.pp_create(ctypes.byref(h), ctypes.byref(g)) == pp.PP_BAD_ARG
    assert b"PP_BC_NEUMANN" in lib.pp_last_error()


def test_create_accepts_the_grids_the_reference_coarsens():
    """Odd cell counts coarsen too (the last coarse cell takes three fine
    cells), so grids the reference's MG_initial halves by dropping a cell
    (CORE_MG.F90:98-126) pass pp_create's checks: 150^3, 130x98x82,
    600x64x64, 65x49x41, 1024 cells along a line.  Here they fail only for
    want of a device.  A line of 1025 cells is refused."""
    import numpy as np
    lib = pp.lib()
    h = ctypes.c_void_p()
    for dims in ((150, 150, 150), (130, 98, 82), (600, 64, 64), (65, 49, 41), (1024, 8, 6)):
        g = pp._grid_desc(*(np.linspace(0, 1, n + 1) for n in dims))
        rc = lib.pp_create(ctypes.byref(h), ctypes.byref(g))
        assert rc != pp.PP_BAD_ARG, (dims, lib.pp_last_error())
        if rc == pp.PP_OK:
            lib.pp_destroy(h)
    g = pp._grid_desc(np.linspace(0, 1, 1026), np.linspace(0, 1, 9), np.linspace(0, 1, 9))
    assert lib.pp_create(ctypes.byref(h), ctypes.byref(g)) == pp.PP_BAD_ARG
    assert b"1024 cells per line" in lib.pp_last_error()
