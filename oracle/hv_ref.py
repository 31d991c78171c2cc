"""ctypes wrapper of the C restatement ``hv_ref.c`` — TEST INFRASTRUCTURE ONLY.

Used by tests/ (parity at sizes numpy handles slowly) and by bench.py's
``cpu_baseline`` leg (the strong multi-threaded CPU matrix-free H·v).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libhvref.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/libhvref.so is not built (run make -C oracle)')
        L = ctypes.CDLL(path)
        L.ref_diag.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, _dp, _lp,
                               ctypes.c_int64, _dp]
        L.ref_hv_full_f64.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.ref_hv_full_f32.argtypes = [ctypes.c_int, _dp, _dp, _fp, _fp]
        L.ref_hv_full_c128.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ref_hv_fib_f64.argtypes = [ctypes.c_int, _dp, _dp, _lp, _dp, _dp]
        L.ref_hv_fib_f32.argtypes = [ctypes.c_int, _dp, _dp, _lp, _fp, _fp]
        L.ref_fib_codes.argtypes = [ctypes.c_int, _lp]
        L.ref_fib_dim.argtypes = [ctypes.c_int]
        L.ref_fib_dim.restype = ctypes.c_int64
        L.ref_fib_unrank.argtypes = [ctypes.c_int, ctypes.c_int64]
        L.ref_fib_unrank.restype = ctypes.c_int64
        L.ref_hv_full_f64_rows.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int64,
                                           ctypes.c_int64]
        L.ref_indep_codes.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), _lp,
                                      ctypes.c_int64]
        L.ref_indep_codes.restype = ctypes.c_int64
        L.ref_hv_codes_f64.argtypes = [ctypes.c_int, _dp, _dp, _lp, ctypes.c_int64, _dp, _dp]
        L.ref_num_threads.restype = ctypes.c_int
        L.ref_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def num_threads():
    return int(lib().ref_num_threads())


def set_threads(t):
    lib().ref_set_threads(int(t))


def diag(n, delta, pairs, codes=None):
    """Diagonal per basis state (reference channels.py:58-77, I1 on)."""
    pi = np.ascontiguousarray([int(p['i']) for p in pairs], np.int32)
    pj = np.ascontiguousarray([int(p['j']) for p in pairs], np.int32)
    V = np.ascontiguousarray([float(p.get('V', 0.0)) for p in pairs], np.float64)
    delta = np.ascontiguousarray(delta, np.float64)
    dim = (1 << n) if codes is None else len(codes)
    out = np.empty(dim, np.float64)
    cptr = None
    if codes is not None:
        codes = np.ascontiguousarray(codes, np.int64)
        cptr = _p(codes, _lp)
    lib().ref_diag(int(n), _p(delta, _dp), len(pairs), _p(pi, _ip), _p(pj, _ip),
                   _p(V, _dp), cptr, int(dim), _p(out, _dp))
    return out


def hv_full(n, Om, d, x):
    """Full-basis H·v, dtype follows x (float64, float32 or complex128)."""
    Om = np.asarray(Om, complex)
    ore = np.ascontiguousarray(Om.real, np.float64)
    oim = np.ascontiguousarray(Om.imag, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    L = lib()
    if x.dtype == np.float64:
        L.ref_hv_full_f64(int(n), _p(ore, _dp), _p(d, _dp), _p(x, _dp), _p(y, _dp))
    elif x.dtype == np.float32:
        L.ref_hv_full_f32(int(n), _p(ore, _dp), _p(d, _dp), _p(x, _fp), _p(y, _fp))
    elif x.dtype == np.complex128:
        L.ref_hv_full_c128(int(n), _p(ore, _dp), _p(oim, _dp), _p(d, _dp),
                           x.ctypes.data_as(_dp), y.ctypes.data_as(_dp))
    else:
        raise TypeError(x.dtype)
    return y


def fib_dim(n):
    return int(lib().ref_fib_dim(int(n)))


def fib_codes(n):
    out = np.empty(fib_dim(n), np.int64)
    lib().ref_fib_codes(int(n), _p(out, _lp))
    return out


def hv_fib(n, Om, d, codes, x):
    ore = np.ascontiguousarray(np.asarray(Om, complex).real, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    codes = np.ascontiguousarray(codes, np.int64)
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    if x.dtype == np.float64:
        lib().ref_hv_fib_f64(int(n), _p(ore, _dp), _p(d, _dp), _p(codes, _lp), _p(x, _dp), _p(y, _dp))
    elif x.dtype == np.float32:
        lib().ref_hv_fib_f32(int(n), _p(ore, _dp), _p(d, _dp), _p(codes, _lp), _p(x, _fp), _p(y, _fp))
    else:
        raise TypeError(x.dtype)
    return y


def hv_full_rows(n, Om, d_rows, x, t0, t1):
    """Rows [t0, t1) of the full-basis FP64 H·v (d_rows: diagonal of those rows)."""
    ore = np.ascontiguousarray(np.asarray(Om, complex).real, np.float64)
    d_rows = np.ascontiguousarray(d_rows, np.float64)
    y = np.empty(int(t1 - t0), np.float64)
    lib().ref_hv_full_f64_rows(int(n), _p(ore, _dp), _p(d_rows, _dp), _p(x, _dp), _p(y, _dp),
                               int(t0), int(t1))
    return y


def _neighbour_masks(n, edges):
    nb = np.zeros(max(int(n), 1), np.uint64)
    for i, j in edges:
        nb[int(i)] |= np.uint64(1) << np.uint64(int(j))
        nb[int(j)] |= np.uint64(1) << np.uint64(int(i))
    return nb


def indep_count(n, edges):
    """Number of independent sets (depth-first count, no list)."""
    nb = _neighbour_masks(n, edges)
    return int(lib().ref_indep_codes(int(n), nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                     None, 0))


def indep_codes(n, edges):
    """Ascending codes of the independent sets (reference basis.py:24-47 order)."""
    cnt = indep_count(n, edges)
    nb = _neighbour_masks(n, edges)
    out = np.empty(cnt, np.int64)
    got = lib().ref_indep_codes(int(n), nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                _p(out, _lp), cnt)
    assert got == cnt
    return out


def hv_codes(n, Om, d, codes, x):
    """Projected FP64 H·v on an ascending code list (real Ω)."""
    ore = np.ascontiguousarray(np.asarray(Om, complex).real, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    codes = np.ascontiguousarray(codes, np.int64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().ref_hv_codes_f64(int(n), _p(ore, _dp), _p(d, _dp), _p(codes, _lp), len(codes),
                           _p(x, _dp), _p(y, _dp))
    return y
