"""ctypes binding of ``librydhip.so`` (the C ABI declared in include/rydhip.h).

The library is loaded after ``torch`` so that the process has a single HIP
runtime (torch's ``libamdhip64.so.7`` satisfies the library's dependency).
There is no CPU fallback: if the library or a GPU is missing, every entry
point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'librydhip.so')

RYD_OK, RYD_ERR_INVALID, RYD_ERR_MEMORY, RYD_ERR_DEVICE, RYD_ERR_UNSUPPORTED = range(5)
RYD_F32, RYD_F64, RYD_C64, RYD_C128 = range(4)
ABI_VERSION = 12
RYD_BASIS_FULL, RYD_BASIS_CODES, RYD_BASIS_CHAIN_FIB, RYD_BASIS_GRAPH, RYD_BASIS_CSR = range(5)

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)


class Model(ctypes.Structure):
    """Mirror of ``ryd_model``."""
    _fields_ = [('n', ctypes.c_int32),
                ('om_re', _dp), ('om_im', _dp), ('delta', _dp),
                ('n_pairs', ctypes.c_int32),
                ('pair_i', _ip), ('pair_j', _ip), ('pair_v', _dp),
                ('basis_kind', ctypes.c_int32),
                ('codes', _lp), ('n_codes', ctypes.c_int64),
                ('pair_jr', _dp), ('pair_ji', _dp), ('pair_ar', _dp), ('pair_ai', _dp),
                ('n_edges', ctypes.c_int32), ('edge_i', _ip), ('edge_j', _ip)]


# name -> (restype, argtypes); kept in the order of include/rydhip.h
SIGNATURES = {
    'ryd_abi_version': (ctypes.c_int, []),
    'ryd_last_error': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    'ryd_plan_create': (ctypes.c_int, [ctypes.POINTER(Model), ctypes.POINTER(_vp)]),
    'ryd_plan_create_csr': (ctypes.c_int, [ctypes.c_int64, _vp, _vp, _vp, _vp,
                                           ctypes.POINTER(_vp)]),
    'ryd_plan_destroy': (ctypes.c_int, [_vp]),
    'ryd_plan_dim': (_i64, [_vp]),
    'ryd_plan_rank_mode': (ctypes.c_int, [_vp]),
    'ryd_workspace_bytes': (ctypes.c_size_t, [_vp]),
    'ryd_hv': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    'ryd_apply': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl,
                                 _vp, _vp, ctypes.c_int, _vp, _vp]),
    'ryd_lanczos': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _i64, _i64, _vp, _vp,
                                   _vp, _dbl, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    'ryd_lanczos_state_bytes': (ctypes.c_size_t, []),
    'ryd_reortho_workspace_bytes': (ctypes.c_size_t, [_i64, _i64]),
    'ryd_lanczos_reortho': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _vp, _i64, _dbl, _vp, _vp,
                                           _vp, _dbl, ctypes.c_int, _vp, _vp, _vp]),
    'ryd_lanczos_reortho_locked': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _i64, _vp, _i64,
                                                  _dbl, _vp, _vp, _vp, _dbl, ctypes.c_int, _vp,
                                                  _vp, _vp]),
    'ryd_lanczos_locked': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _i64, _i64, _vp,
                                          _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _vp, ctypes.c_int,
                                          _vp, _vp]),
    'ryd_lanczos_status': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(_i64), _vp]),
    'ryd_lanczos_stored': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _i64, _i64, _i64, _vp, _vp,
                                          _vp, _dbl, _vp, _vp]),
    'ryd_ritz_stored': (ctypes.c_int, [ctypes.c_int, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _i64,
                                       ctypes.c_int, _vp, _vp, _vp, _vp]),
    'ryd_cheb_moments': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _i64, _dbl,
                                        _dbl, _vp, _vp, _vp]),
    'ryd_cheb_moments_noalloc': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _i64, _dbl,
                                                _dbl, _vp, _vp]),
    'ryd_site_occupation': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    'ryd_pauli_expect': (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                        _vp, _vp]),
    'ryd_site_correlations': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t,
                                             _vp]),
    'ryd_sector_codes': (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp]),
    'ryd_sector_rank': (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp]),
    'ryd_evolve_forward': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _vp, _vp, _vp, _vp,
                                          ctypes.c_int, _vp, _vp, _vp]),
    'ryd_evolve_taylor': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _dbl, ctypes.c_int, ctypes.c_int,
                                         _vp, _vp, _vp]),
    'ryd_loss_axpy': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    'ryd_evolve_adjoint': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _dbl, _dbl, _dbl, _dbl,
                                          ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    'ryd_evolve_forward_store': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _vp, _vp, _vp, _vp,
                                                ctypes.c_int, _vp, _vp, _vp, _vp]),
    'ryd_evolve_partials_bytes': (_i64, [_vp, ctypes.c_int, ctypes.c_int]),
    'ryd_evolve_adjoint_stored': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _i64, _vp, _vp, _vp, _vp,
                                                 ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    'ryd_batch_create': (ctypes.c_int, [ctypes.POINTER(Model), ctypes.c_int32, ctypes.c_int,
                                        ctypes.c_int32, ctypes.POINTER(_vp)]),
    'ryd_batch_create_csr': (ctypes.c_int, [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                            ctypes.c_int32, ctypes.POINTER(_vp)]),
    'ryd_batch_set_lock': (ctypes.c_int, [_vp, _vp]),
    'ryd_batch_destroy': (ctypes.c_int, [_vp]),
    'ryd_batch_offset': (_i64, [_vp, ctypes.c_int32]),
    'ryd_batch_run': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    'ryd_batch_run_steps': (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    'ryd_batch_stop': (ctypes.c_int, [_vp, _vp]),
    'ryd_batch_autotune': (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32), _vp]),
    'ryd_ritz_lowest': (ctypes.c_int, [ctypes.c_int32, _vp, _vp, _i64, _vp, _i64, _vp, _dbl, _dbl,
                                       _vp, _vp, _vp, _i64]),
    'ryd_batch_basis_bytes': (_i64, [_vp]),
    'ryd_batch_attach_basis': (ctypes.c_int, [_vp, _vp, _i64]),
    'ryd_batch_set_tile_bits': (ctypes.c_int, [_vp, ctypes.c_int32]),
    'ryd_batch_results': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'ryd_batch_ritz': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'ryd_dots': (ctypes.c_int, [ctypes.c_int, _vp, _vp, _i64, _vp, _vp, _vp]),
    'ryd_cast': (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, _vp, _i64, _vp]),
}

_LIB = None


class NativeError(RuntimeError):
    """HIP runtime failure inside librydhip."""


def lib():
    """The loaded library; raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    import torch  # noqa: F401  (one HIP runtime per process: torch's)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('librydhip.so is not built (run __graft_entry__.build()): %s' % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.ryd_abi_version() != ABI_VERSION:
        raise RuntimeError('librydhip ABI version mismatch')
    _LIB = L
    return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().ryd_last_error(buf, len(buf))
    return buf.value.decode(errors='replace')


def check(rc):
    """Map a status code onto the reference's exception types
    (ValueError / MemoryError / RuntimeError, matrix_free.py:135-175)."""
    if rc == RYD_OK:
        return
    msg = last_error()
    if rc == RYD_ERR_INVALID:
        raise ValueError(msg)
    if rc == RYD_ERR_MEMORY:
        raise MemoryError(msg)
    if rc == RYD_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise NativeError(msg)


def require_gpu():
    """RuntimeError when no HIP device is visible (reference matrix_free.py:135-136)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('HIP device unavailable')


def stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return _vp(s.cuda_stream)


def ptr(t):
    return None if t is None else _vp(t.data_ptr())


def dtype_code(torch_dtype):
    import torch
    table = {torch.float32: RYD_F32, torch.float64: RYD_F64,
             torch.complex64: RYD_C64, torch.complex128: RYD_C128}
    if torch_dtype not in table:
        raise ValueError('unsupported state dtype %r' % (torch_dtype,))
    return table[torch_dtype]
