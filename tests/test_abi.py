"""The C-ABI library loads and exports every symbol include/rydhip.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, 'include', 'rydhip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|int64_t|size_t)\s+(ryd_\w+)\s*\(', text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ('ryd_plan_create', 'ryd_hv', 'ryd_lanczos', 'ryd_cheb_moments'):
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    from energy_level_amd import _native as nat
    assert os.path.exists(nat.LIB_PATH), 'build() first'
    L = ctypes.CDLL(nat.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # the binding table covers the header too
    assert set(_declared()) <= set(nat.SIGNATURES)


def test_abi_version_and_error_path(pkg):
    from energy_level_amd import _native as nat
    L = nat.lib()
    assert L.ryd_abi_version() == 12
    # invalid model is rejected before any device work
    handle = ctypes.c_void_p()
    m = nat.Model()
    m.n = 0
    rc = L.ryd_plan_create(ctypes.byref(m), ctypes.byref(handle))
    assert rc == nat.RYD_ERR_INVALID
    assert 'n must be' in nat.last_error()
    with pytest.raises(ValueError):
        nat.check(rc)


def test_argument_checks_before_device_work(pkg):
    """Entry points reject null or out-of-range arguments before any device
    call (so these run on a host without a GPU)."""
    from energy_level_amd import _native as nat
    L = nat.lib()
    g = (ctypes.c_double * 4)(0.0, 0.0, 0.0, 0.0)
    # ryd_evolve_taylor: null plan, then a real dtype on a null plan
    assert L.ryd_evolve_taylor(None, nat.RYD_C128, None, 1.0, 1, 20, g, None, None) == nat.RYD_ERR_INVALID
    assert 'null argument' in nat.last_error()
    assert L.ryd_hv(None, nat.RYD_F64, None, None, None) == nat.RYD_ERR_INVALID
    # ryd_plan_create_csr validates the host arrays before any device call
    h = ctypes.c_void_p()
    ptr = (ctypes.c_int64 * 3)(0, 1, 2)
    col = (ctypes.c_int64 * 2)(0, 5)
    val = (ctypes.c_double * 2)(1.0, 2.0)
    assert L.ryd_plan_create_csr(2, ptr, col, val, None, ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert 'column index out of range' in nat.last_error()
    bad = (ctypes.c_int64 * 3)(0, 2, 1)
    assert L.ryd_plan_create_csr(2, bad, col, val, None, ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert L.ryd_plan_create_csr(0, ptr, col, val, None, ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert L.ryd_evolve_forward(None, nat.RYD_C64, None, 1, None, None, None, None, 4, None,
                                None, None) == nat.RYD_ERR_INVALID


def test_csr_batch_argument_checks_before_device_work(pkg):
    """ryd_batch_create_csr validates the concatenated CSR on the host: a
    column outside its own instance, complex values with a real dtype, a
    decreasing row pointer; the lock and tile-size calls refuse a null batch."""
    import numpy as np
    from energy_level_amd import _native as nat
    L = nat.lib()
    h = ctypes.c_void_p()
    dims = np.array([2, 1], np.int64)
    ip = np.array([0, 1, 2, 3], np.int64)
    re = np.array([1.0, 1.0, 2.0])
    im = np.array([0.5, -0.5, 0.0])
    # row 1 of instance 0 points at column 2 (outside dim 2): refused
    bad_cols = np.array([1, 2, 0], np.int64)
    assert L.ryd_batch_create_csr(2, dims.ctypes.data, ip.ctypes.data, bad_cols.ctypes.data,
                                  re.ctypes.data, None, nat.RYD_F64, 8,
                                  ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert 'outside its instance' in nat.last_error()
    cols = np.array([1, 0, 0], np.int64)
    assert L.ryd_batch_create_csr(2, dims.ctypes.data, ip.ctypes.data, cols.ctypes.data,
                                  re.ctypes.data, im.ctypes.data, nat.RYD_F64, 8,
                                  ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert 'complex dtype' in nat.last_error()
    ip_bad = np.array([0, 2, 1, 3], np.int64)
    assert L.ryd_batch_create_csr(2, dims.ctypes.data, ip_bad.ctypes.data, cols.ctypes.data,
                                  re.ctypes.data, None, nat.RYD_F64, 8,
                                  ctypes.byref(h)) == nat.RYD_ERR_INVALID
    zero = np.array([0, 1], np.int64)
    assert L.ryd_batch_create_csr(2, zero.ctypes.data, ip.ctypes.data, cols.ctypes.data,
                                  re.ctypes.data, None, nat.RYD_F64, 8,
                                  ctypes.byref(h)) == nat.RYD_ERR_INVALID
    assert L.ryd_batch_set_lock(None, None) == nat.RYD_ERR_INVALID
    assert L.ryd_batch_basis_bytes(None) == -1
    assert L.ryd_batch_attach_basis(None, None, 0) == nat.RYD_ERR_INVALID
