"""Batched ground states of many instances of uneven N in one graph launch.

Replaces the reference's dense batched eigh (tomisnot/energy-level
src/core/manybody/equilibrium.py:251-311, uniform dim only, O(dim^3)) and its
serial ensemble loop (ensemble.py:51-175, which rejects uneven N at
:114-116) with a matrix-free batched Lanczos (``ryd_batch_*``): FP32 or FP64
vectors, FP64 alpha/beta, the whole recurrence captured in one hipGraph and a
tile size chosen by a timing autotuner.  E0 and the gap come from the FP64
tridiagonal of each instance (lowest genuine Ritz values).
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from . import _native as nat
from .model import RydModel, from_positions, chain_positions, c6_coupling, from_calib

__all__ = ['BatchLanczos', 'BatchCsrLanczos', 'CsrInstance', 'csr_instances', 'disorder_ensemble',
           'batched_ground_state_hip', 'batched_ground_state_csr_hip', 'certified_ground_states']

BMAXN = 24            # largest instance n (occupation rows hold BMAXN sites + the norm)


def _basis_fits(device, nbytes, cached_only):
    """True when ``nbytes`` more device memory fit (blocks torch's allocator
    holds unused count as free) with 4 GiB to spare; ``cached_only``: they
    must come from those cached blocks (a new allocation of tens of GB, and
    its release, cost the driver about a second on MI355X, more than the
    replay it saves)."""
    import torch
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    if cached_only and nbytes > cached:
        return False
    free, _ = torch.cuda.mem_get_info(device)
    return nbytes + (4 << 30) <= free + cached


class BatchLanczos:
    """``ryd_batch`` owner for a list of :class:`RydModel` (full basis, real drive).

    ``store_basis``: lend the batch a torch buffer for its whole Krylov basis
    (``ryd_batch_attach_basis``) so that Ritz vectors combine the stored rows
    instead of replaying the recurrence.  True: whenever memory allows;
    'auto': only out of blocks torch's allocator already holds unused (a
    repeated solve reuses the buffer of the previous one); False: the batch's
    three rotating rows."""

    def __init__(self, models, steps=150, dtype='float32', device='cuda', store_basis='auto'):
        import torch
        nat.require_gpu()
        if dtype not in ('float32', 'float64'):
            raise ValueError('dtype must be float32/float64')
        self.models = list(models)
        self.count = len(self.models)
        if self.count < 1:
            raise ValueError('empty batch')
        self.steps = int(steps)
        self.dtype = dtype
        self.tdt = torch.float32 if dtype == 'float32' else torch.float64
        dv = torch.device(device)
        self.device = torch.device('cuda', torch.cuda.current_device() if dv.index is None
                                   else dv.index)
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        Arr = nat.Model * self.count
        ms = Arr()
        for k, m in enumerate(self.models):
            om = np.asarray(m.Om, complex)
            if np.any(om.imag != 0):
                raise ValueError('batched instances need real drives')
            ms[k].n = int(m.n)
            ms[k].om_re = arr(om.real, np.float64).ctypes.data_as(nat._dp)
            ms[k].om_im = None
            ms[k].delta = arr(m.delta, np.float64).ctypes.data_as(nat._dp)
            ms[k].n_pairs = len(m.pair_v)
            ms[k].pair_i = arr(m.pair_i, np.int32).ctypes.data_as(nat._ip)
            ms[k].pair_j = arr(m.pair_j, np.int32).ctypes.data_as(nat._ip)
            ms[k].pair_v = arr(m.pair_v, np.float64).ctypes.data_as(nat._dp)
            ms[k].basis_kind = nat.RYD_BASIS_FULL
            ms[k].codes = None
            ms[k].n_codes = 0
        h = ctypes.c_void_p()
        code = nat.RYD_F32 if dtype == 'float32' else nat.RYD_F64
        with torch.cuda.device(self.device):
            nat.check(nat.lib().ryd_batch_create(ms, self.count, code, self.steps, ctypes.byref(h)))
        self._h = h
        L = nat.lib()
        self.offsets = np.array([L.ryd_batch_offset(h, i) for i in range(self.count + 1)], np.int64)
        self.total = int(self.offsets[-1])
        self.tile_bits = 12
        self.autotune_ms = None
        self._lend_basis(store_basis)

    def _lend_basis(self, store_basis):
        """Lend a torch buffer for the whole Krylov basis (``store_basis`` as
        in the class docstring)."""
        import torch
        self._basis = None
        if store_basis not in (True, False, 'auto'):
            raise ValueError("store_basis must be True, False or 'auto'")
        if not store_basis:
            return
        L = nat.lib()
        need = int(L.ryd_batch_basis_bytes(self._h))
        if _basis_fits(self.device, need, cached_only=store_basis == 'auto'):
            try:
                buf = torch.empty(need, dtype=torch.uint8, device=self.device)
            except torch.OutOfMemoryError:
                buf = None
            if buf is not None:
                nat.check(L.ryd_batch_attach_basis(self._h, nat.ptr(buf), need))
                self._basis = buf

    @property
    def stores_basis(self):
        """True when the whole Krylov basis is kept (Ritz vectors without replay)."""
        return self._basis is not None

    def close(self):
        if getattr(self, '_h', None) is not None and self._h.value:
            nat.lib().ryd_batch_destroy(self._h)
            self._h = None
        # the lent basis goes back to torch's cache only after the batch is gone
        self._basis = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def start_vectors(self, seed=0):
        """Concatenated N(0,1) start vectors from a seeded device generator."""
        import torch
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        return torch.randn(self.total, dtype=torch.float64, device=self.device,
                           generator=g).to(self.tdt)

    def set_tile_bits(self, bits):
        """Use tiles of 2^bits codes (10, 11 or 12), e.g. the autotuner's
        choice of another batch over the same instances."""
        nat.check(nat.lib().ryd_batch_set_tile_bits(self._h, int(bits)))
        self.tile_bits = int(bits)

    def autotune(self, v0):
        tb = ctypes.c_int32(0)
        ms = (ctypes.c_double * 3)()
        nat.check(nat.lib().ryd_batch_autotune(self._h, nat.ptr(v0), ctypes.byref(tb), ms))
        self.tile_bits = int(tb.value)
        self.autotune_ms = {10 + k: float(ms[k]) for k in range(3)}
        return self.tile_bits

    def run(self, v0, use_graph=True):
        if v0.numel() != self.total or v0.dtype != self.tdt or v0.device != self.device:
            raise ValueError('start vectors must be a %s tensor of length %d on %s'
                             % (self.tdt, self.total, self.device))
        self._hstop = None
        nat.check(nat.lib().ryd_batch_run(self._h, nat.ptr(v0.contiguous()), int(bool(use_graph)),
                                          nat.stream_ptr()))

    def coefficients(self):
        alpha = np.zeros((self.count, self.steps))
        nrm = np.zeros((self.count, self.steps + 1))
        stop = np.zeros(self.count, np.int32)
        nat.check(nat.lib().ryd_batch_results(self._h, alpha.ctypes.data, nrm.ctypes.data,
                                              stop.ctypes.data))
        return alpha, nrm, stop

    def instance_index(self):
        """int16 tensor of the instance of every element of the concatenation."""
        import torch
        if getattr(self, '_inst_el', None) is None:
            sizes = torch.as_tensor(np.diff(self.offsets), device=self.device)
            self._inst_el = torch.repeat_interleave(
                torch.arange(self.count, dtype=torch.int16, device=self.device), sizes)
        return self._inst_el

    def ritz(self, v0, ycoef):
        """Σ_j ycoef[b, j]·q_j of every instance (q_j the normalised Lanczos
        vectors of the last run from ``v0``: the stored rows, or replayed)
        and the unnormalised
        site weights of each result (occupation rows: BMAXN sites + |x|²)."""
        import torch
        ycoef = np.ascontiguousarray(ycoef, np.float64)
        if ycoef.shape != (self.count, self.steps):
            raise ValueError('ycoef must be shaped (count, steps)')
        ritz = torch.empty(self.total, dtype=self.tdt, device=self.device)
        occ = torch.empty((self.count, BMAXN + 1), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ryd_batch_ritz(self._h, nat.ptr(v0.contiguous()), ycoef.ctypes.data,
                                           nat.ptr(ritz), nat.ptr(occ), nat.stream_ptr()))
        return ritz, occ

    def occupations(self, v0):
        """<n_i> of each instance's lowest Ritz vector (list of arrays, site
        order), from the stored rows or replaying the recurrence from the
        same start vectors ``v0``."""
        self.ground_states()
        _, occ = self.ritz(v0, self.restart_coefficients(np.zeros(self.count, np.int64)))
        occ = occ.cpu().numpy()
        return [occ[b, :int(m.n)] / occ[b, BMAXN] for b, m in enumerate(self.models)]

    def _host_stopped(self):
        hs = getattr(self, '_hstop', None)
        return np.zeros(self.count, bool) if hs is None else hs < self.steps

    def _steps_of(self, nrm, stop):
        """Valid steps per instance: all, up to the breakdown, or up to the
        step where :meth:`run_parts` stopped it."""
        hs = getattr(self, '_hstop', None)
        k = np.full(self.count, self.steps, np.int32) if hs is None else hs.astype(np.int32)
        tiny = 1e-6 if self.dtype == 'float32' else 1e-13
        for b in np.nonzero(stop)[0]:
            bad = np.nonzero(~(nrm[b, 1:int(k[b]) + 1] > tiny))[0]
            if bad.size:
                k[b] = int(bad[0]) + 1
        return k

    def run_parts(self, v0, tol, merge_rel, need0, need1, part=150):
        """The recurrence from ``v0`` in parts of ``part`` steps (no graph).
        After each part, an instance whose still-needed Ritz pairs (``need0``,
        ``need1``: bool per instance) have bounds <= tol·max(1,|E|) stops at
        that step, so a long cycle costs only the steps its instances use;
        :meth:`ground_states` then reads each at its own step count."""
        if v0.numel() != self.total or v0.dtype != self.tdt or v0.device != self.device:
            raise ValueError('start vectors must be a %s tensor of length %d on %s'
                             % (self.tdt, self.total, self.device))
        v0 = v0.contiguous()
        self._hstop = np.full(self.count, self.steps, np.int32)
        need0, need1 = np.asarray(need0, bool), np.asarray(need1, bool)
        j = 0
        while j < self.steps:
            j1 = min(self.steps, j + int(part))
            nat.check(nat.lib().ryd_batch_run_steps(self._h, nat.ptr(v0) if j == 0 else None,
                                                    j, j1, nat.stream_ptr()))
            j = j1
            if j >= self.steps:
                break
            _, _, stop = self.coefficients()
            live = (need0 | need1) & (stop == 0)
            if not np.any(live):
                break
            self._hstop[live] = j
            gs = self.ground_states(tol, live, merge_rel)
            E1 = gs['E0'] + gs['gap']
            with np.errstate(invalid='ignore'):
                ok1 = gs['ritz_bound_gap'] <= tol * np.maximum(1.0, np.abs(E1))
            done = live & (gs['converged_E0'] | ~need0) & (ok1 | ~need1)
            self._hstop[live & ~done] = self.steps
            if np.any(done):
                nat.check(nat.lib().ryd_batch_stop(self._h, np.ascontiguousarray(done, np.int32).ctypes.data))

    def ground_states(self, tol=None, active=None, merge_rel=None):
        """E0, gap (lowest two genuine Ritz values, ``ryd_ritz_lowest``), their
        Ritz bounds and, with ``tol``, converged flags per instance:
        ``converged_E0`` (the E0 bound <= tol·max(1,|E0|)) and ``converged``
        (both bounds, the rule of torch_lanczos_ground_state,
        matrix_free.py:258-266).  ``active`` (bool mask) limits the work to
        those instances; the others are left at zero.  The two Ritz vectors'
        coefficients are kept for :meth:`restart_coefficients`."""
        alpha, nrm, stop = self.coefficients()
        idx = np.arange(self.count) if active is None else np.nonzero(active)[0]
        k = self._steps_of(nrm, stop)
        E = np.full((self.count, 2), np.nan)
        bnd = np.full((self.count, 2), np.inf)
        if getattr(self, '_ycoef', None) is None:
            self._ycoef = np.zeros((self.count, 2, self.steps))
        if idx.size:
            a = np.ascontiguousarray(alpha[idx])
            o = np.ascontiguousarray(nrm[idx, 1:])
            kk = np.ascontiguousarray(k[idx])
            dstop = (stop != 0) & ~self._host_stopped()
            bl = np.where(dstop[idx], 0.0, nrm[idx, kk])
            bl = np.ascontiguousarray(bl, np.float64)
            Es = np.zeros((idx.size, 2))
            Bs = np.zeros((idx.size, 2))
            Ys = np.zeros((idx.size, 2, self.steps))
            nat.check(nat.lib().ryd_ritz_lowest(
                int(idx.size), kk.ctypes.data, a.ctypes.data, self.steps, o.ctypes.data,
                self.steps, bl.ctypes.data, float(tol if tol is not None else 0.0),
                float(merge_rel if merge_rel is not None else
                      (1e-6 if self.dtype == 'float32' else 1e-10)), Es.ctypes.data, Bs.ctypes.data, Ys.ctypes.data, self.steps))
            E[idx], bnd[idx] = Es, Bs
            self._ycoef[idx] = Ys
        out = dict(E0=np.where(np.isnan(E[:, 0]), 0.0, E[:, 0]), gap=E[:, 1] - E[:, 0],
                   ritz_bound=np.where(np.isinf(bnd[:, 0]), 0.0, bnd[:, 0]),
                   ritz_bound_gap=np.where((stop != 0) & ~self._host_stopped(), 0.0, bnd[:, 1]),
                   steps=np.where(np.isin(np.arange(self.count), idx), k, 0),
                   stopped=(stop != 0) & ~self._host_stopped())
        if tol is not None:
            E0, E1 = out['E0'], out['E0'] + np.nan_to_num(out['gap'])
            c0 = out['ritz_bound'] <= tol * np.maximum(1.0, np.abs(E0))
            out['converged_E0'] = c0
            out['converged'] = c0 & (out['ritz_bound_gap'] <= tol * np.maximum(1.0, np.abs(E1)))
        return out

    def restart_coefficients(self, which):
        """ycoef rows from the last ``ground_states`` call: for instance b,
        ``which[b]`` = 0 → its lowest Ritz vector, 1 → the sum of the two
        lowest (the restart seed; the lowest alone when there is no second),
        -1 → zeros."""
        w = np.asarray(which)
        y = self._ycoef
        out = np.where((w >= 0)[:, None], y[:, 0], 0.0)
        two = (w == 1)[:, None] & np.isfinite(y[:, 1]).all(axis=1)[:, None]
        return np.ascontiguousarray(out + np.where(two, y[:, 1], 0.0))


class CsrInstance:
    """One Hermitian instance for :class:`BatchCsrLanczos`: any sparse or
    dense square matrix (rows and columns in the order of its vectors),
    kept as a sorted CSR without explicit zeros.  ``n`` is its qubit count
    when dim = 2^n (n <= 24: occupations, site i = bit n-1-i of the row
    index), else 0."""

    def __init__(self, H):
        import scipy.sparse as sp
        A = sp.csr_matrix(H)
        if A.ndim != 2 or A.shape[0] != A.shape[1]:
            raise ValueError('instances must be square, got %r' % (A.shape,))
        A.eliminate_zeros()
        A.sum_duplicates()
        A.sort_indices()
        self.H = A
        self.dim = int(A.shape[0])
        self.complex = bool(np.iscomplexobj(A.data) and np.any(A.data.imag != 0))
        bits = max(self.dim - 1, 0).bit_length()
        self.n = bits if (1 << bits) == self.dim and bits <= BMAXN else 0


def csr_instances(mats):
    """:class:`CsrInstance` of every matrix (instances pass through).  The
    conversions run on up to 16 host threads: a dense matrix is scanned for
    its non-zeros, which for a stack of dense matrices (64 × 16 MB at
    dim 1024, complex128) is memory-bound host work that numpy and scipy
    do with the GIL released."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    todo = [k for k, m in enumerate(mats) if not isinstance(m, CsrInstance)]
    out = list(mats)
    workers = min(16, os.cpu_count() or 1, len(todo))
    if workers <= 1:
        for k in todo:
            out[k] = CsrInstance(mats[k])
        return out
    with ThreadPoolExecutor(workers) as ex:
        for k, inst in zip(todo, ex.map(lambda k: CsrInstance(mats[k]), todo)):
            out[k] = inst
    return out


class BatchCsrLanczos(BatchLanczos):
    """``ryd_batch_create_csr`` owner: the batched Lanczos of
    :class:`BatchLanczos` over :class:`CsrInstance` items (any Hermitian
    matrices, any dims).  ``dtype`` 'float32' / 'float64' is the precision;
    the vectors are complex when any instance is.  The tile size is fixed
    (``autotune`` and ``set_tile_bits`` do nothing); ``set_lock`` runs the
    recurrence in the complement of per-instance vectors.  ``force_complex``
    keeps complex vectors for real instances too (a batch shrunk out of a
    mixed one continues from complex restart vectors)."""

    def __init__(self, instances, steps=150, dtype='float64', device='cuda', store_basis='auto',
                 force_complex=False):
        import torch
        nat.require_gpu()
        if dtype not in ('float32', 'float64'):
            raise ValueError('dtype must be float32/float64')
        self.models = csr_instances(instances)
        self.count = len(self.models)
        if self.count < 1:
            raise ValueError('empty batch')
        self.steps = int(steps)
        self.dtype = dtype
        cplx = bool(force_complex) or any(m.complex for m in self.models)
        self.tdt = {(False, 'float32'): torch.float32, (False, 'float64'): torch.float64,
                    (True, 'float32'): torch.complex64, (True, 'float64'): torch.complex128}[
            (cplx, dtype)]
        dv = torch.device(device)
        self.device = torch.device('cuda', torch.cuda.current_device() if dv.index is None
                                   else dv.index)
        dims = np.array([m.dim for m in self.models], np.int64)
        nnz = np.array([m.H.nnz for m in self.models], np.int64)
        base = np.concatenate([[0], np.cumsum(nnz)[:-1]])
        indptr = np.ascontiguousarray(np.concatenate(
            [[0]] + [m.H.indptr[1:].astype(np.int64) + base[k] for k, m in enumerate(self.models)]),
            np.int64)
        indices = np.ascontiguousarray(np.concatenate(
            [m.H.indices.astype(np.int64) for m in self.models]), np.int64)
        data = np.concatenate([m.H.data for m in self.models])
        val_re = np.ascontiguousarray(data.real, np.float64)
        val_im = np.ascontiguousarray(data.imag, np.float64) if cplx else None
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(nat.lib().ryd_batch_create_csr(
                self.count, dims.ctypes.data, indptr.ctypes.data,
                indices.ctypes.data if indices.size else None,
                val_re.ctypes.data if val_re.size else None,
                val_im.ctypes.data if val_im is not None and val_im.size else None,
                nat.dtype_code(self.tdt), self.steps, ctypes.byref(h)))
        self._h = h
        L = nat.lib()
        self.offsets = np.array([L.ryd_batch_offset(h, i) for i in range(self.count + 1)], np.int64)
        self.total = int(self.offsets[-1])
        self.tile_bits = None
        self.autotune_ms = None
        self._lock = None
        self._lend_basis(store_basis)

    def close(self):
        super().close()
        self._lock = None

    def start_vectors(self, seed=0):
        """Concatenated N(0,1) start vectors (complex: both parts) from a
        seeded device generator."""
        import torch
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        dt = torch.complex128 if self.tdt.is_complex else torch.float64
        return torch.randn(self.total, dtype=dt, device=self.device, generator=g).to(self.tdt)

    def set_tile_bits(self, bits):
        pass

    def autotune(self, v0):
        return None

    def set_lock(self, g):
        """Lock vectors (concatenated unit vectors in the batch's layout, a
        device tensor kept alive here) or None."""
        if g is not None:
            if g.numel() != self.total or g.dtype != self.tdt or g.device != self.device:
                raise ValueError('lock vectors must be a %s tensor of length %d on %s'
                                 % (self.tdt, self.total, self.device))
            g = g.contiguous()
        nat.check(nat.lib().ryd_batch_set_lock(self._h, nat.ptr(g)))
        self._lock = g

    def project_off(self, v, g):
        """v with each instance's component along its g removed (device op)."""
        import torch
        inst = self.instance_index().long()
        c = torch.zeros(self.count, dtype=v.dtype, device=v.device)
        c.index_add_(0, inst, torch.conj(g) * v)
        return v - c[inst] * g


def _make_batch(items, steps, dtype, device, store_basis, force_complex=False):
    """The batch kind for these instances: full-basis Rydberg models, or CSR."""
    if isinstance(items[0], RydModel):
        return BatchLanczos(items, steps=steps, dtype=dtype, device=device,
                            store_basis=store_basis)
    return BatchCsrLanczos(items, steps=steps, dtype=dtype, device=device,
                           store_basis=store_basis, force_complex=force_complex)


def _any_complex(items):
    return any(getattr(m, 'complex', False) for m in items)


def disorder_ensemble(count=512, nmin=10, nmax=20, sigma=0.1, deltas=None, c6=10.0, omega=1.0,
                      seed=0):
    """Chain instances with seeded N ~ U{nmin..nmax}, Gaussian jitter σ on all
    three coordinates (reference model.py:117-120), Δ swept over ``deltas``
    (instance i takes deltas[i % len]), V = C6/r^6 on all pairs."""
    rng = np.random.default_rng(int(seed))
    deltas = np.linspace(-2.0, 4.0, 32) if deltas is None else np.asarray(deltas, float)
    out = []
    for i in range(int(count)):
        n = int(rng.integers(nmin, nmax + 1))
        pos = chain_positions(n) + rng.normal(0.0, sigma, size=(n, 3))
        calib = from_positions(pos, c6_coupling(c6), terms=('I1',))
        Delta = deltas[i % len(deltas)]
        out.append(from_calib(calib, omega * np.ones(n), -Delta * np.ones(n)))
    return out


def _masked(B, x, keep):
    """x with the instances where ``keep`` is False set to zero (device op)."""
    import torch
    m = torch.as_tensor(np.asarray(keep, bool), device=B.device)[B.instance_index().long()]
    return torch.where(m, x, torch.zeros((), dtype=x.dtype, device=x.device))


def _segments(B, x, local):
    """The concatenated segments of x (B's layout) of B's instances ``local``."""
    import torch
    return torch.cat([x[int(B.offsets[i]):int(B.offsets[i + 1])] for i in local])


def _reset_records(out, count):
    """Per-pair certification records of a stage (E0 and E1 separately, each
    from the first cycle whose bound passes)."""
    out['_rec'] = dict(c0=np.zeros(count, bool), c1=np.zeros(count, bool),
                       E0=np.zeros(count), E1=np.zeros(count),
                       b0=np.full(count, np.inf), b1=np.full(count, np.inf))


def _stage(models, idx, v, dtype, steps, tol, merge_rel, cycles, out, use_graph, tile_bits,
           device, occ_rows=None, seeds=None, finish_last=True, store_basis='auto', vecs=None,
           want=2, lock=None, parts=None):
    """Batched Lanczos cycles over the instances ``idx`` (indices into
    ``models``) from their concatenated start vectors ``v``, until each
    instance's lowest ``want`` genuine Ritz pairs have bounds <=
    tol·max(1,|E|) or ``cycles`` are spent.  Each pair is certified by the
    first cycle whose bound passes (``out['_rec']``, kept across the stages
    of one precision): once both pairs are nearly converged, a restart from
    their sum starts a recurrence whose next norm is tiny, and the rounding
    it amplifies keeps the two bounds from passing in the same cycle.

    A finished instance records E0, gap, bounds and flags in ``out`` (at its
    index) and then its lowest Ritz vector's site weights (``occ_rows``) and,
    with ``vecs`` (a dict), that vector itself, or, with ``seeds`` (a dict),
    the sum of its two lowest Ritz vectors as the seed of a next stage.  The
    others restart from that sum: while they are at least half of the batch
    they stay in it (a finished instance gets a zero start vector, which the
    batch kernels skip: its norm is zero, so its stop flag is set before the
    first step); below that the cycle continues on a new batch of those
    instances only, so that the launches shrink with the work.  With
    ``lock`` (a dict of vectors, CSR batches) every batch runs in the
    complement of each instance's vector (``BatchCsrLanczos.set_lock``; the
    start and restart vectors are projected off it), so its lowest pair is
    the next level.  With ``finish_last`` False the instances still
    unconverged after the last cycle are not recorded.  Returns (cycles run,
    indices still unconverged, their restart vectors concatenated, or None)."""
    import torch
    fc = _any_complex(models)
    idx = np.asarray(idx, np.int64)
    R = out['_rec']

    def new_batch(sel):
        B = _make_batch([models[i] for i in sel], steps, dtype, device, store_basis, fc)
        if tile_bits is not None:
            B.set_tile_bits(tile_bits)
        G = None
        if lock is not None:
            G = torch.cat([lock[int(i)] for i in sel]).to(B.tdt)
            inst = B.instance_index().long()
            nrm = torch.zeros(B.count, dtype=torch.float64, device=B.device)
            nrm.index_add_(0, inst, (torch.abs(G) ** 2).to(torch.float64))
            G = G / torch.sqrt(nrm)[inst].to(G.dtype)
            B.set_lock(G)
        return B, G

    B, G = new_batch(idx)
    if G is not None:
        v = B.project_off(v.to(B.tdt), G)
    try:
        act = np.ones(len(idx), bool)
        done = 0
        for cyc in range(int(cycles)):
            if parts and type(B) is BatchLanczos and B.stores_basis and steps >= 2 * parts:
                B.run_parts(v, tol, merge_rel, act & ~R['c0'][idx],
                            act & (~R['c1'][idx] if want == 2 else False), parts)
            else:
                B.run(v, use_graph=use_graph)
            done += 1
            gs = B.ground_states(tol, act, merge_rel)
            E1 = gs['E0'] + gs['gap']
            ok0 = act & gs['converged_E0']
            with np.errstate(invalid='ignore'):
                ok1 = act & (gs['ritz_bound_gap'] <= tol * np.maximum(1.0, np.abs(E1)))
            n0 = ok0 & ~R['c0'][idx]
            R['c0'][idx[n0]] = True
            R['E0'][idx[n0]] = gs['E0'][n0]
            R['b0'][idx[n0]] = gs['ritz_bound'][n0]
            n1 = ok1 & ~R['c1'][idx]
            R['c1'][idx[n1]] = True
            R['E1'][idx[n1]] = E1[n1]
            R['b1'][idx[n1]] = gs['ritz_bound_gap'][n1]
            cert = R['c0'][idx] & (R['c1'][idx] if want == 2 else True)
            last = cyc == cycles - 1
            fin = act & (cert | (last and finish_last))
            g = idx[fin]
            c0f, c1f = R['c0'][g], R['c1'][g]
            e0 = np.where(c0f, R['E0'][g], gs['E0'][fin])
            out['E0'][g] = e0
            out['gap'][g] = np.where(c1f, R['E1'][g], E1[fin]) - e0
            out['ritz_bound'][g] = np.where(c0f, R['b0'][g], gs['ritz_bound'][fin])
            out['ritz_bound_gap'][g] = np.where(c1f, R['b1'][g], gs['ritz_bound_gap'][fin])
            out['converged_E0'][g] = c0f
            out['converged'][g] = (c0f & c1f) if want == 2 else c0f
            out['steps_' + dtype][idx[act]] += gs['steps'][act]
            rest = act & ~fin
            which = np.where(rest, 1, np.where(fin, 1 if seeds is not None else 0, -1))
            if not (np.any(rest) or seeds is not None or occ_rows is not None
                    or vecs is not None):
                return done, idx[:0], None
            x, occ = B.ritz(v, B.restart_coefficients(which))
            if G is not None:
                x = B.project_off(x, G)
            if occ_rows is not None and np.any(fin):
                occ_np = occ.cpu().numpy()
                for l in np.nonzero(fin)[0]:
                    occ_rows[idx[l]] = occ_np[l].copy()
            if seeds is not None:
                for l in np.nonzero(fin)[0]:
                    seeds[int(idx[l])] = x[int(B.offsets[l]):int(B.offsets[l + 1])].clone()
            elif vecs is not None:
                for l in np.nonzero(fin)[0]:
                    vecs[int(idx[l])] = x[int(B.offsets[l]):int(B.offsets[l + 1])].clone()
            if not np.any(rest):
                return done, idx[:0], None
            if last:
                loc = np.nonzero(rest)[0]
                return done, idx[loc], _segments(B, x, loc)
            if 2 * int(rest.sum()) <= len(idx):
                loc = np.nonzero(rest)[0]
                v = _segments(B, x, loc)
                idx = idx[loc]
                B.close()
                B, G = new_batch(idx)
                if G is not None:
                    v = B.project_off(v.to(B.tdt), G)
                act = np.ones(len(idx), bool)
            else:
                v = _masked(B, x, rest)
                act = rest
        return done, idx[:0], None
    finally:
        B.close()


def certified_ground_states(models, steps=150, dtype='float32', refine=True, bulk_tol=1e-6,
                            tol=1e-9, bulk_cycles=4, refine_schedule=((300, 1), (700, 1), (1500, 2)),
                            seed=0, device='cuda', autotune=True, use_graph=True, observables=True,
                            tile_bits=None, seed_noise=1e-3, store_basis='auto', vectors=None,
                            refine_parts=150):
    """Batched ground states with a convergence test per instance.

    Bulk: ``steps``-step batched Lanczos cycles in ``dtype`` (one captured
    graph) from seeded start vectors; an instance leaves when its two lowest
    genuine Ritz pairs have bounds <= tol·max(1,|E|), the others restart from
    the sum of their two lowest Ritz vectors (see :func:`_stage`).  With an
    FP32 bulk and ``refine``, the bulk stops at ``bulk_tol`` and FP64 cycles
    take over to ``tol`` from the bulk Ritz vectors (plus a seeded random
    part of relative size ``seed_noise``, so that a direction the bulk lost
    is present again), following ``refine_schedule`` = ((steps, cycles), ...):
    a short cycle first (an instance the bulk brought to its FP32 limit
    needs no more), then longer ones for what is left, the last ones long
    and continued (few restarts) for the slowest instances, which by then
    form a small batch of their own.  E0, the gap and their bounds are then
    FP64 results.  Ghost copies merge within 1e-6·|T| in FP32 cycles and
    1e-10·|T| in FP64 ones (their rounding).  An FP64 bulk (or no
    refinement) runs ``bulk_cycles`` cycles and then the schedule's cycle
    counts as further cycles of ``steps`` steps.  Defaults measured on the
    512-instance ensemble (N 10-20, 150-step cycles, each pair certified
    independently, lent basis): E1, not E0, sets the FP64 time (the slow
    instances have a close E2), so no instance finishes in a 60-step first
    cycle; four FP32 cycles to 1e-6 and ((300, 1), (700, 1), (1500, 2))
    certify all 512 in 1.83 s (three FP32 cycles 2.34 s, five 1.86 s; each
    timed after a warm-up that leaves its basis sizes in torch's cache),
    against 2.90 s for two FP32 cycles and
    ((60, 1), (300, 1), (1000, 3)), 2.53 s for two FP32 cycles and the new
    schedule, 2.62-3.05 s for ((300, 1), (700, 1), (2000, 1)),
    ((300, 1), (800, 1), (1800, 2)) and ((350, 1), (900, 1), (2000, 1)).
    With the FP64 cycles in parts of 150 steps (``refine_parts``), instances
    leave a cycle at the part where they pass: 1.53 s (201,600 FP64
    instance-steps instead of 311,000).
    Earlier, without per-pair certification: FP32 to 1e-5 with ((50, 1),
    (200, 2), (1000, 2)) left 4 instances unconverged.

    Returns E0, gap, ritz_bound (E0), ritz_bound_gap (E1), converged (both
    pairs, the final stage), converged_E0, bulk_converged,
    steps_float32/steps_float64 (steps run per instance), cycles_* and, with
    ``observables``, n_atom (<n_i> of each instance's final lowest Ritz
    vector).  ``tile_bits`` skips the autotuner.  ``store_basis`` is passed
    to every :class:`BatchLanczos` of the stages (True lets the first call
    take the basis memory into torch's cache; a repeated call with 'auto'
    then reuses it: 3.72 s instead of 3.88 s on the ensemble above).
    ``refine_parts``: FP64 cycles of at least twice that many steps run in
    parts of that many steps (:meth:`BatchLanczos.run_parts`, full-basis
    batches with a stored basis), and an instance whose needed pairs pass
    stops at the end of a part instead of running to the cycle's end (None:
    whole cycles).
    ``models`` are :class:`RydModel` instances (the full-basis batch) or
    :class:`CsrInstance` items (any Hermitian matrices, ``BatchCsrLanczos``);
    ``vectors`` (a dict) receives each instance's final lowest Ritz vector
    (device, the batch's layout and element type)."""
    import torch
    count = len(models)
    two_stage = bool(refine) and dtype == 'float32'
    out = dict(E0=np.zeros(count), gap=np.full(count, np.nan), ritz_bound=np.zeros(count),
               ritz_bound_gap=np.zeros(count), converged=np.zeros(count, bool),
               converged_E0=np.zeros(count, bool), steps_float32=np.zeros(count, np.int64),
               steps_float64=np.zeros(count, np.int64))
    occ_rows = [None] * count if observables else None
    merge = {'float32': 1e-6, 'float64': 1e-10}
    schedule = [(int(st), int(cy)) for st, cy in refine_schedule if int(cy) > 0]
    B = _make_batch(models, steps, dtype, device, False, _any_complex(models))
    try:
        v0 = B.start_vectors(seed)
        if tile_bits is not None:
            B.set_tile_bits(tile_bits)
        elif autotune:
            tile_bits = B.autotune(v0)
        out['autotune_ms'] = B.autotune_ms
        out['tile_bits'] = B.tile_bits
    finally:
        B.close()
    all_idx = np.arange(count)
    # one stage (an FP64 bulk, or no refinement) takes the whole cycle budget
    budget = bulk_cycles if two_stage else bulk_cycles + sum(cy for _, cy in schedule)
    seeds = {} if two_stage else None
    _reset_records(out, count)
    done, _, _ = _stage(models, all_idx, v0, dtype, steps, bulk_tol if two_stage else tol,
                        merge[dtype], budget, out, use_graph, tile_bits, device,
                        None if two_stage else occ_rows, seeds, store_basis=store_basis,
                        vecs=None if two_stage else vectors)
    del v0
    out['cycles_' + dtype] = done
    out['bulk_converged'] = out['converged'].copy()
    if two_stage:
        out['converged'][:] = False
        out['converged_E0'][:] = False
        gen = torch.Generator(device=seeds[0].device)
        gen.manual_seed(int(seed) + 1)
        parts = []
        for b in range(count):
            x = seeds.pop(b)
            x = x.to(torch.complex128 if x.is_complex() else torch.float64)
            d = x.numel()
            noise = torch.randn(d, dtype=x.dtype, device=x.device, generator=gen)
            parts.append(x / torch.linalg.vector_norm(x) + (seed_noise / np.sqrt(d)) * noise)
        vrest = torch.cat(parts)
        del parts
        done64, rest = 0, all_idx
        _reset_records(out, count)
        for k, (st, cy) in enumerate(schedule):
            if not len(rest):
                break
            d, rest, vrest = _stage(models, rest, vrest, 'float64', st, tol, merge['float64'], cy,
                                    out, use_graph, tile_bits, device, occ_rows,
                                    finish_last=(k == len(schedule) - 1),
                                    store_basis=store_basis, vecs=vectors, parts=refine_parts)
            done64 += d
        out['cycles_float64'] = done64
    if observables:
        out['n_atom'] = [np.asarray(occ_rows[b][:int(m.n)]) / occ_rows[b][BMAXN]
                         for b, m in enumerate(models)]
    out['tolerance'] = tol if (two_stage or dtype == 'float64') else bulk_tol
    return out


def batched_ground_state_hip(models, steps=150, dtype='float32', seed=0, device='cuda',
                             autotune=True, use_graph=True, observables=True, refine=True,
                             tol=1e-9, bulk_tol=1e-6, bulk_cycles=4,
                             refine_schedule=((300, 1), (700, 1), (1500, 2))):
    """Dict like batched_ground_state (tomisnot/energy-level
    src/core/manybody/equilibrium.py:251-311): E0 (B,), gap (B,), n_atom (a
    list, one array of <n_i> per instance since N may differ) and backend,
    for uneven-N instances; plus converged (B,), converged_E0 and the Ritz
    bounds of E0 and E1.  Instances run batched Lanczos cycles until their
    two lowest Ritz pairs converge (FP32 bulk to ``bulk_tol``, then an FP64
    refinement to ``tol``, see :func:`certified_ground_states`); an instance
    that does not converge within the cycle budget is flagged and reported
    in ``warning``.  ``observables=False`` skips <n_i>."""
    import warnings
    out = certified_ground_states(models, steps=steps, dtype=dtype, refine=refine,
                                  bulk_tol=bulk_tol, tol=tol, bulk_cycles=bulk_cycles,
                                  refine_schedule=refine_schedule, seed=seed, device=device,
                                  autotune=autotune, use_graph=use_graph,
                                  observables=observables)
    out.update(backend='hip-batched-lanczos', dtype=dtype)
    bad = np.nonzero(~out['converged'])[0]
    if bad.size:
        bad0 = np.nonzero(~out['converged_E0'])[0]
        out['warning'] = ('%d of %d instances did not converge to tol=%g within the cycle budget '
                          '(%d of them in E0; the others in E1, so their gap is an estimate '
                          'within ritz_bound_gap); indices %s'
                          % (bad.size, len(models), out['tolerance'], bad0.size, bad[:16].tolist()))
        warnings.warn(out['warning'], RuntimeWarning)
    return out


def _locked_second_level(items, vecs, tol, seed=1, device='cuda', store_basis='auto',
                         schedule=((150, 3), (300, 2), (1000, 3))):
    """The second level of every CSR instance counted with multiplicity:
    batched FP64 Lanczos cycles in the complement of the instance's ground
    vector (``_stage`` with ``lock``; seeded start vectors projected off it)
    until the lowest Ritz pair there has a bound <= tol·max(1,|E|), as eigh's
    w[1] (equilibrium.py:272; a degenerate ground level gives E1 = E0).
    ``schedule`` = ((steps, cycles), ...) as the refinement of
    :func:`certified_ground_states`: short cycles first, long ones for the
    instances whose next level has a close neighbour.  Returns (E1, bound,
    converged), NaN / inf where an instance has no complement (dim 1)."""
    import torch
    count = len(items)
    out = dict(E0=np.full(count, np.nan), gap=np.full(count, np.nan),
               ritz_bound=np.full(count, np.inf), ritz_bound_gap=np.full(count, np.inf),
               converged=np.zeros(count, bool), converged_E0=np.zeros(count, bool),
               steps_float64=np.zeros(count, np.int64))
    _reset_records(out, count)
    rest = np.array([i for i in range(count) if items[i].dim > 1], np.int64)
    if not rest.size:
        return out['E0'], out['ritz_bound'], out['converged_E0']
    dt = torch.complex128 if _any_complex(items) else torch.float64
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    v = torch.randn(int(sum(items[i].dim for i in rest)), dtype=dt, device=dev, generator=gen)
    sched = [(int(st), int(cy)) for st, cy in schedule if int(cy) > 0]
    for k, (st, cy) in enumerate(sched):
        if not len(rest):
            break
        _, rest, v = _stage(items, rest, v, 'float64', st, tol, 1e-10, cy, out, True, None, device,
                            finish_last=(k == len(sched) - 1), store_basis=store_basis, want=1,
                            lock=vecs)
    return out['E0'], out['ritz_bound'], out['converged_E0']


def batched_ground_state_csr_hip(mats, steps=150, dtype='float64', tol=1e-9, seed=0,
                                 device='cuda', observables=True, multiplicity=True,
                                 store_basis='auto'):
    """Ground states of many Hermitian instances of any kind and size (dense
    or sparse matrices, real or complex, uneven dims) in batched launches:
    ``certified_ground_states`` over :class:`CsrInstance` items (E0 and its
    Ritz vector certified to ``tol``·max(1,|E|), as torch_lanczos_ground_state's
    rule, matrix_free.py:258-266), then, with ``multiplicity``, one locked
    batched run per instance set for E1 counted with multiplicity
    (:func:`_locked_second_level`), so gap = E1 - E0 is eigh's w[1] - w[0]
    (equilibrium.py:272) also for a degenerate ground level.  ``dtype``
    'float32' runs an FP32 (complex64) bulk before the FP64 refinement.
    Returns E0, gap, converged (E0 and, with ``multiplicity``, E1), the
    bounds and, with ``observables``, n_atom (instances with dim 2^n)."""
    items = csr_instances(mats)
    vecs = {}
    out = certified_ground_states(items, steps=steps, dtype=dtype, tol=tol, seed=seed,
                                  device=device, autotune=False, observables=observables,
                                  store_basis=store_basis, vectors=vecs)
    if multiplicity:
        E1, b1, c1 = _locked_second_level(items, vecs, tol, seed=seed + 7, device=device,
                                          store_basis=store_basis)
        out['gap_distinct'] = out['gap']
        out['gap'] = E1 - out['E0']
        out['ritz_bound_gap'] = b1
        out['converged_E1'] = c1
        out['converged'] = out['converged_E0'] & (c1 | np.array([m.dim < 2 for m in items]))
    out['backend'] = 'hip-batched-csr-lanczos'
    return out


def batched_ground_state_dense_hip(H_batch, backend='auto', vram_budget=1.2e9, device='cuda'):
    """batched_ground_state (tomisnot/energy-level
    src/core/manybody/equilibrium.py:251-311) with its signature: H_batch is a
    (B, dim, dim) Hermitian stack with dim = 2^n; → dict(E0 (B,), gap (B,),
    n_atom (B, n), backend).  The stack goes to one CSR batch
    (:func:`batched_ground_state_csr_hip`: every instance in the same
    launches, E0 certified to 1e-9·max(1,|E0|), E1 counted with multiplicity
    so that a degenerate ground level gives gap 0 as eigh's w[1] - w[0]
    does); <n_i> of the ground state with site 0 the most significant bit,
    as bit_matrix.  ``backend`` and ``vram_budget`` are accepted for the
    signature.  An instance whose E0 or E1 misses the tolerance within the
    cycle budget keeps its best estimate, with a warning."""
    Hb = np.asarray(H_batch)
    if Hb.ndim != 3 or Hb.shape[1] != Hb.shape[2]:
        raise ValueError('H_batch must be shaped (B, dim, dim), got %r' % (Hb.shape,))
    B, dim, _ = Hb.shape
    n = int(round(np.log2(dim))) if dim > 0 else 0
    if dim < 1 or (1 << n) != dim:
        raise ValueError('dim=%d is not 2^N' % dim)
    if B == 0:
        return dict(E0=np.zeros(0), gap=np.zeros(0), n_atom=np.zeros((0, n)),
                    backend='hip-batched-csr-lanczos')
    out = batched_ground_state_csr_hip([Hb[b] for b in range(B)], device=device)
    bad = np.nonzero(~out['converged'])[0]
    if bad.size:
        warnings.warn('%d of %d instances did not converge to 1e-9 within the cycle budget; '
                      'indices %s' % (bad.size, B, bad[:16].tolist()), RuntimeWarning)
    n_atom = np.array([np.asarray(a, float) for a in out['n_atom']]).reshape(B, n)
    return dict(E0=np.asarray(out['E0']), gap=np.asarray(out['gap']), n_atom=n_atom,
                backend=out['backend'])
