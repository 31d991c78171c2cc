"""Device plan: immutable model tables on one GPU plus a per-object workspace.

``Plan.hv`` replaces the reference's matrix-free action
(tomisnot/energy-level src/core/manybody/channels.py:210-221 on the CPU and
matrix_free.py:200-204 on the GPU); no transition table is ever built.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .model import RydModel

__all__ = ['Plan']


class Plan:
    """``ryd_plan`` owner.  Vectors are torch tensors on ``self.device``."""

    def __init__(self, model: RydModel, basis='full', codes=None, device='cuda', edges=None):
        import torch
        nat.require_gpu()
        L = nat.lib()
        self.model = model
        self.n = int(model.n)
        dv = torch.device(device)
        if dv.type != 'cuda':
            raise RuntimeError('HIP device unavailable (device=%s)' % device)
        # always carry an explicit index so tensors compare equal to it
        self.device = torch.device('cuda', torch.cuda.current_device() if dv.index is None
                                   else dv.index)
        kinds = {'full': nat.RYD_BASIS_FULL, 'codes': nat.RYD_BASIS_CODES,
                 'chain_fib': nat.RYD_BASIS_CHAIN_FIB, 'graph': nat.RYD_BASIS_GRAPH}
        if basis not in kinds:
            raise ValueError('unknown basis %r' % (basis,))
        self.basis = basis
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            self._keep.append(a)
            return a

        om = np.asarray(model.Om, complex)
        ore, oim = arr(om.real, np.float64), arr(om.imag, np.float64)
        dl = arr(model.delta, np.float64)
        pi, pj = arr(model.pair_i, np.int32), arr(model.pair_j, np.int32)
        pv = arr(model.pair_v, np.float64)
        pJ = np.asarray(model.pair_J, np.complex128)
        pA = np.asarray(model.pair_A, np.complex128)
        m = nat.Model()
        if model.has_pair_flips:
            m.pair_jr = arr(pJ.real, np.float64).ctypes.data_as(nat._dp)
            m.pair_ji = arr(pJ.imag, np.float64).ctypes.data_as(nat._dp)
            m.pair_ar = arr(pA.real, np.float64).ctypes.data_as(nat._dp)
            m.pair_ai = arr(pA.imag, np.float64).ctypes.data_as(nat._dp)
        m.n = self.n
        m.om_re = ore.ctypes.data_as(nat._dp)
        m.om_im = oim.ctypes.data_as(nat._dp)
        m.delta = dl.ctypes.data_as(nat._dp)
        m.n_pairs = len(pv)
        m.pair_i = pi.ctypes.data_as(nat._ip)
        m.pair_j = pj.ctypes.data_as(nat._ip)
        m.pair_v = pv.ctypes.data_as(nat._dp)
        m.basis_kind = kinds[basis]
        if codes is not None:
            c = arr(codes, np.int64)
            m.codes = c.ctypes.data_as(nat._lp)
            m.n_codes = len(c)
        else:
            m.codes = None
            m.n_codes = 0
        if basis == 'graph':
            e = np.asarray(edges if edges is not None else [], np.int64).reshape(-1, 2)
            m.n_edges = len(e)
            m.edge_i = arr(e[:, 0], np.int32).ctypes.data_as(nat._ip)
            m.edge_j = arr(e[:, 1], np.int32).ctypes.data_as(nat._ip)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(L.ryd_plan_create(ctypes.byref(m), ctypes.byref(handle)))
        self._h = handle
        self._keep = []
        self.dim = int(L.ryd_plan_dim(handle))
        self.is_real = model.is_real
        self._ws = None

    # -- lifetime ------------------------------------------------------------
    def close(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            nat.lib().ryd_plan_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def rank_mode(self):
        """'full', 'number', 'search', 'graph' or 'chain' (how codes are ranked)."""
        m = int(nat.lib().ryd_plan_rank_mode(self._h))
        return {0: 'full', 1: 'number', 2: 'search', 3: 'graph', 4: 'chain'}.get(m, 'unknown')

    def lanczos_one_launch(self, dtype):
        """True when ryd_lanczos runs every step of a call in one launch:
        full basis without J/A channels, D <= 8192 (complex128 4096)."""
        import torch
        lim = 4096 if dtype == torch.complex128 else 8192
        return self.basis == 'full' and not self.model.has_pair_flips and self.dim <= lim

    def workspace(self):
        """Zero-filled workspace of ``ryd_workspace_bytes`` (reused by this plan)."""
        import torch
        if self._ws is None:
            nb = int(nat.lib().ryd_workspace_bytes(self._h))
            self._ws = torch.zeros(nb, dtype=torch.uint8, device=self.device)
        return self._ws

    # -- checks ----------------------------------------------------------------
    def check_vec(self, x, name='x'):
        import torch
        if not isinstance(x, torch.Tensor) or x.device != self.device:
            raise ValueError('%s must be a tensor on %s' % (name, self.device))
        if x.dim() != 1 or x.numel() != self.dim or not x.is_contiguous():
            raise ValueError('%s must be a contiguous vector of length %d' % (name, self.dim))
        code = nat.dtype_code(x.dtype)
        if not self.is_real and code in (nat.RYD_F32, nat.RYD_F64):
            raise ValueError('complex Rabi phases need a complex dtype')
        return code

    # -- operator --------------------------------------------------------------
    def hv(self, x, out=None):
        """y = H x (same dtype as x)."""
        import torch
        code = self.check_vec(x)
        y = torch.empty_like(x) if out is None else out
        self.check_vec(y, 'out')
        nat.check(nat.lib().ryd_hv(self._h, code, nat.ptr(x), nat.ptr(y), nat.stream_ptr()))
        return y

    def apply(self, x, y, s=1.0, ch=1.0, cx=0.0, cz=0.0, z=None, probe=None, sums=None,
              accumulate=False):
        """y = ch·H(s·x) + cx·(s·x) + cz·z with optional FP64 sums (device tensor[4])."""
        code = self.check_vec(x)
        self.check_vec(y, 'y')
        if z is not None:
            self.check_vec(z, 'z')
        ws = self.workspace() if sums is not None else None
        nat.check(nat.lib().ryd_apply(self._h, code, nat.ptr(x), nat.ptr(z), nat.ptr(y),
                                      float(s), float(ch), float(cx), float(cz), nat.ptr(probe),
                                      nat.ptr(sums), int(bool(accumulate)), nat.ptr(ws),
                                      nat.stream_ptr()))
        return y

    def dots(self, a, b, sums):
        """sums[0:3] = {Re<a,b>, |b|^2, |a|^2} (device FP64)."""
        code = nat.dtype_code(b.dtype)
        nat.check(nat.lib().ryd_dots(code, nat.ptr(a), nat.ptr(b), int(b.numel()), nat.ptr(sums),
                                     nat.ptr(self.workspace()), nat.stream_ptr()))
        return sums


class CsrPlan(Plan):
    """``ryd_plan`` over a caller's sparse Hermitian matrix H (scipy sparse or
    dense; rows and columns in the order of the vectors), copied once to the
    device.  Serves hv, apply, the Lanczos and Chebyshev drivers; entry points
    that need site codes refuse it."""

    def __init__(self, H, device='cuda'):
        import scipy.sparse as sp
        import torch
        nat.require_gpu()
        L = nat.lib()
        Hc = sp.csr_matrix(H, copy=True)
        if Hc.ndim != 2 or Hc.shape[0] != Hc.shape[1] or Hc.shape[0] < 1:
            raise ValueError('H must be a non-empty square matrix')
        Hc.sum_duplicates()
        Hc.sort_indices()
        self.model = None
        self.n = 0
        self.basis = 'csr'
        dv = torch.device(device)
        if dv.type != 'cuda':
            raise RuntimeError('HIP device unavailable (device=%s)' % device)
        self.device = torch.device('cuda', torch.cuda.current_device() if dv.index is None
                                   else dv.index)
        D = int(Hc.shape[0])
        indptr = np.ascontiguousarray(Hc.indptr, np.int64)
        idx = np.ascontiguousarray(Hc.indices, np.int64)
        data = np.asarray(Hc.data)
        re = np.ascontiguousarray(data.real, np.float64)
        im = (np.ascontiguousarray(data.imag, np.float64)
              if np.iscomplexobj(data) and np.any(data.imag != 0) else None)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(L.ryd_plan_create_csr(D, indptr.ctypes.data, idx.ctypes.data,
                                            re.ctypes.data,
                                            None if im is None else im.ctypes.data,
                                            ctypes.byref(handle)))
        self._h = handle
        self._keep = []
        self.dim = D
        self.nnz = int(Hc.nnz)
        self.is_real = im is None
        self._ws = None


def cast(x, dtype):
    """Precision cast of a state vector through the library (F32<->F64, C64<->C128)."""
    import torch
    y = torch.empty(x.shape, dtype=dtype, device=x.device)
    nat.check(nat.lib().ryd_cast(nat.dtype_code(x.dtype), nat.ptr(x), nat.dtype_code(dtype),
                                 nat.ptr(y), int(x.numel()), nat.stream_ptr()))
    return y
