"""The C-ABI library (libcoldplate_hip.so) on a host without a GPU: it loads,
exports every function the public header declares, its structs have the
layout the ctypes mirrors assume, and host-only entry points (defaults,
argument checks, error reporting) work.  No kernel is launched."""
import ctypes as C
import os
import re

import pytest


@pytest.fixture(scope="module")
def L(pkg):
    if not os.path.exists(pkg._capi.LIB_PATH):
        pytest.fail("libcoldplate_hip.so is not built: run __graft_entry__.build()")
    return pkg._capi.lib()


def test_every_declared_symbol_is_exported(pkg, L):
    syms = pkg._capi.declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_header_struct_fields_match_ctypes_mirrors(pkg):
    """Field names and order of each C struct equal the ctypes mirror's."""
    txt = open(pkg._capi.HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    T = pkg._types
    for cname, mirror in (("cps_simple_config", T.SimpleConfig), ("cps_step_stats", T.StepStats),
                          ("cps_patch", T.Patch), ("cps_thermal_config", T.ThermalConfig),
                          ("cps_solve_stats", T.SolveStats), ("cps_thermal_metrics", T.ThermalMetrics),
                          ("cps_iter_log", T.IterLog)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), txt, flags=re.S).group(1)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            parts = decl.split(None, 1)
            names += [n.strip() for n in parts[1].split(",")]
        assert names == [f for f, _ in mirror._fields_], cname


def test_config_defaults_match_reference(pkg, L):
    T = pkg._types
    c = T.SimpleConfig()
    L.cps_simple_config_default(C.byref(c))
    # SIMPLE.h:152-153, 176-177, 374-375, 397-398
    assert (c.rho, c.eta, c.uv_alpha, c.p_alpha) == (997.0, 0.00089, 0.7, 0.3)
    assert (c.momentum_tol, c.max_momentum_iter, c.p_tol, c.max_pressure_iter) == (1e-6, 400, 1e-6, 30)
    py = pkg.default_config()
    for k, _ in T.SimpleConfig._fields_:
        if k in ("nx", "ny", "hx", "hy", "u_in"):
            continue
        assert getattr(c, k) == getattr(py, k), k
    t = T.ThermalConfig()
    L.cps_thermal_config_default(C.byref(t))
    # thermal_solver.cpp:160-168, 182-183
    assert (t.k_s, t.k_f, t.rho, t.cp, t.t_in) == (400.0, 0.6, 998.0, 4180.0, 30.0)
    assert (t.k_mode, t.qk, t.rtol, t.max_iter) == (1, 1.0, 1e-6, 1000)


def test_abi_version_and_argument_errors(pkg, L):
    T = pkg._types
    assert L.cps_abi_version() >= 1
    h = C.c_void_p()
    assert L.cps_simple_create(None, None, None, None, C.byref(h)) == T.CPS_ERR_ARG
    assert b"null" in L.cps_last_error()
    assert L.cps_thermal_create(None, None, None, None, None, None, None, None, C.byref(h)) == T.CPS_ERR_ARG
    assert L.cps_simple_step(None, 1, None) == T.CPS_ERR_ARG
    assert L.cps_thermal_solve(None, None, None) == T.CPS_ERR_ARG
    assert L.cps_create(0, None) == T.CPS_ERR_ARG
    with pytest.raises(pkg._capi.CpsError):
        pkg._capi.check(L.cps_synchronize(None))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_device_is_reported_not_faked(pkg, L):
    """Without a device the library fails loudly (no CPU fallback)."""
    h = C.c_void_p()
    st = L.cps_create(0, C.byref(h))
    assert st == pkg._types.CPS_ERR_HIP
    with pytest.raises(pkg._capi.CpsError):
        pkg.SIMPLE(pkg.configs.channel_200x50(), __import__("numpy").ones((50, 200)))
