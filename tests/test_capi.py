"""The C-ABI library loads and exports every entry point include/hcx.h declares."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hcx.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(hcx_[a-zA-Z0-9_]+)\s*\(", src)
    return sorted(set(n for n in names if not n.startswith("hcx_complex")))


def test_header_functions_exported(hcx_mod):
    from hcx import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in names if n not in _capi.SIGNATURES]
    assert not unbound, f"declared but not bound in hcx/_capi.py: {unbound}"


def test_abi_version_and_errors(hcx_mod):
    lib = hcx_mod.lib()
    assert lib.hcx_abi_version() == 1
    # invalid argument path sets a message and a status, no crash
    st = lib.hcx_system_create_synthetic(None, None)
    assert st == 1
    assert b"invalid" in lib.hcx_last_error()


def test_default_options_are_reference_constants(hcx_mod):
    o = hcx_mod.default_opts()
    # kernel_HC_Solver_generalized_3views_4pts.cu:78,116,121-131,214-215
    assert o.dt0 == 0.05 and o.end_zone == 0.05 and o.t_tol == 1e-7
    assert o.corr_tol == 1e-6 and o.inf_tol == 1e14
    assert o.unique_tol == 1e-4  # definitions.hpp:24
    assert o.real_tol == 1e-8  # definitions.hpp:26 (DP)
    assert o.real_ref_tol == 1e-4  # definitions.hpp:25 (SP), the reference's own real test
    assert o.replicas == 0  # replica lane groups chosen by the library


def test_no_device_fails_loudly_without_gpu(hcx_mod):
    """Without a HIP device the product raises instead of computing elsewhere."""
    if hcx_mod.device_count() > 0:
        return
    import numpy as np
    import pytest
    s = hcx_mod.System.synthetic(3, 0, 3, 0, 1)
    hx, ht = s.tables()
    with pytest.raises(hcx_mod.HcxError, match="no HIP device"):
        hcx_mod.Problem(3, s.info["num_coeffs"], hx, ht, np.ones((8, 3)))
