"""Piecewise-constant time evolution on the HIP engine, forward and adjoint.

Propagator convention of the reference (tomisnot/energy-level
src/core/manybody/engine_ed.py:145): psi(t + dt) = exp(-i·CONV·H·dt) psi,
CONV = 2π·1e-3 (H in MHz, t in ns).  Each slice applies a Chebyshev series
(Jacobi–Anger)

    exp(-iτH) = e^{-iτc} [J_0(τr) + 2 Σ_{k>=1} (-i)^k J_k(τr) T_k((H - c)/r)],

τ = CONV·dt, with [c - r, c + r] a Gershgorin enclosure of the slice's
spectrum.  The adjoint is the Clenshaw-form reverse recurrence of that same
series (``ryd_evolve_adjoint_stored``), so gradients are those of the
discrete propagator; (c, r) are constants of the slice (no gradient flows
through them).  Checkpoints are kept every ``checkpoint`` slices and the
slices between them are recomputed once during the backward pass.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .model import RydModel, from_calib
from .plan import Plan

__all__ = ['CONV', 'cheb_coefficients', 'terms_for', 'PulseProblem', 'pulse_loss_and_grad',
           'PulseEvolution', 'propagate', 'z2_index', 'slice_bounds', 'segment_intervals',
           'segment_drive', 'propagate_segments_hip', 'taylor_bound', 'propagate_lossy']

TAYLOR_DEGREE = 20      # terms per Taylor step; truncation below 1/21! ≈ 2e-20 at ‖h·A‖ <= 1

CONV = 2.0 * np.pi * 1e-3


def cheb_coefficients(tau, c, r, K):
    """a_0..a_K of exp(-iτ(c + r x)) = Σ a_k T_k(x) (Jacobi–Anger), as complex128."""
    from scipy.special import jv
    k = np.arange(K + 1)
    a = (2.0 - (k == 0)) * ((-1j) ** k) * jv(k, tau * r)
    return np.exp(-1j * tau * c) * a


def terms_for(tau_r, eps=1e-15):
    """Smallest K with the series tail below eps (|J_k(z)| ≤ (z/2)^k / k!)."""
    from math import lgamma, log
    z = abs(tau_r)
    k = max(1, int(np.ceil(z)) + 1)
    while True:
        if z == 0 or k * log(z / 2.0 if z > 0 else 1.0) - lgamma(k + 1) < log(eps):
            return k
        k += 1


def _diag_range(model: RydModel, dadd):
    """Bounds of the diagonal delta·n + dadd·N + V·n n over all codes."""
    lo = np.minimum(0.0, np.asarray(model.delta) + dadd).sum() + np.minimum(0.0, model.pair_v).sum()
    hi = np.maximum(0.0, np.asarray(model.delta) + dadd).sum() + np.maximum(0.0, model.pair_v).sum()
    return lo, hi


def slice_bounds(model: RydModel, om, dadd):
    """(c, r) enclosing the spectrum of om·F + D + dadd·N (Gershgorin)."""
    lo, hi = _diag_range(model, dadd)
    rad = abs(om) * float(np.sum(np.abs(model.Om))) / 2.0
    lo, hi = lo - rad, hi + rad
    c = 0.5 * (lo + hi)
    r = max(0.5 * (hi - lo), 1e-12)
    return c, r


class PulseProblem:
    """Per-slice uniform drive: H_k = Ω_k·F + D + (−Δ_k)·N on the full basis.

    ``model`` supplies the flip pattern F (its Ω, normally 1 on every site),
    the fixed diagonal D (its delta, normally 0, and V pairs).  Slices have
    duration ``dt`` and ``K+1`` Chebyshev terms.  ``units='reference'`` is the
    reference propagator exp(−i·CONV·H·dt) (dt in ns, H in MHz);
    ``'natural'`` is exp(−iH·dt).  A slice whose phase span τ·r would leave
    a series tail above ``eps`` is integrated as several equal sub-steps of
    the same drive (gradients are summed back onto the slice).

    The adjoint keeps the state at every ``checkpoint``-th slice; in the
    backward sweep each segment between checkpoints is recomputed once with
    its Chebyshev vectors kept in HBM (at most ``store_budget_bytes``, by
    default half the free device memory; a longer segment is recomputed in
    chunks from its checkpoint).
    """

    def __init__(self, model: RydModel, dt, K=15, dtype='complex64', device='cuda',
                 checkpoint=20, plan=None, units='reference', eps=1e-13, store_budget_bytes=None):
        import torch
        nat.require_gpu()
        self.model = model
        self.plan = plan if plan is not None else Plan(model, device=device)
        self.dt = float(dt)
        if units not in ('reference', 'natural'):
            raise ValueError("units must be 'reference' (exp(-i·CONV·H·dt)) or 'natural' (exp(-iH·dt))")
        self.units = units
        self.tau = CONV * self.dt if units == 'reference' else self.dt
        self.eps = float(eps)
        self.K = int(K)
        self.tdt = {'complex64': torch.complex64, 'complex128': torch.complex128}[dtype]
        self.dtype = dtype
        self.checkpoint = max(1, int(checkpoint))
        self.store_budget_bytes = store_budget_bytes
        self.dim = self.plan.dim

    def _substeps(self, tau_r):
        """Sub-steps so that a (K+1)-term series of span tau_r/s has tail <= eps."""
        s = 1
        while terms_for(tau_r / s, self.eps) > self.K:
            s += 1
        return s

    def _slice_params(self, Om, Delta):
        """Per sub-step arrays and the owner slice of each sub-step."""
        Om = np.asarray(Om, float)
        Delta = np.asarray(Delta, float)
        dadd = -Delta
        oms, das, cs, rs, coef, owner = [], [], [], [], [], []
        for k, (om, da) in enumerate(zip(Om, dadd)):
            c, r = slice_bounds(self.model, om, da)
            sub = self._substeps(self.tau * r)
            a = cheb_coefficients(self.tau / sub, c, r, self.K)
            for _ in range(sub):
                oms.append(om)
                das.append(da)
                cs.append(c)
                rs.append(r)
                coef.append(np.stack([a.real, a.imag], axis=1))
                owner.append(k)
        coef = np.ascontiguousarray(np.asarray(coef, np.float64))
        self._owner = np.asarray(owner, np.int64)
        self._nslices = len(Om)
        return (np.ascontiguousarray(oms, dtype=float), np.ascontiguousarray(das, dtype=float),
                np.asarray(cs), np.asarray(rs), coef)

    def _forward_slices(self, psi, params, k0, k1, work):
        Om, dadd, cs, rs, coef = params
        if k1 <= k0:
            return
        sl = slice(k0, k1)
        om_a = np.ascontiguousarray(Om[sl])
        da_a = np.ascontiguousarray(dadd[sl])
        c_a = np.ascontiguousarray(cs[sl])
        r_a = np.ascontiguousarray(rs[sl])
        cf = np.ascontiguousarray(coef[sl])
        nat.check(nat.lib().ryd_evolve_forward(
            self.plan.handle, nat.dtype_code(psi.dtype), nat.ptr(psi), k1 - k0,
            om_a.ctypes.data, da_a.ctypes.data, c_a.ctypes.data, r_a.ctypes.data, self.K,
            cf.ctypes.data, nat.ptr(work), nat.stream_ptr()))

    def propagate(self, psi0, Om, Delta):
        """ψ(T) from ψ0 (device tensor, copied)."""
        import torch
        params = self._slice_params(Om, Delta)
        psi = psi0.to(self.tdt).clone()
        work = torch.empty(4 * self.dim, dtype=self.tdt, device=self.plan.device)
        self._forward_slices(psi, params, 0, len(params[0]), work)
        return psi

    def _segments(self):
        """Checkpoint segments as sub-step ranges [s0, s1): a checkpoint at
        every ``checkpoint``-th slice (sub-steps of one slice stay together)."""
        owner = self._owner
        starts = [0] + [i for i in range(1, len(owner))
                        if owner[i] != owner[i - 1] and owner[i] % self.checkpoint == 0]
        return list(zip(starts, starts[1:] + [len(owner)]))

    def forward_checkpoints(self, psi0, params):
        """ψ(T) and the checkpoint states {first sub-step of a segment: ψ}."""
        import torch
        psi = psi0.to(self.tdt).clone()
        work = torch.empty(4 * self.dim, dtype=self.tdt, device=self.plan.device)
        ckpts = {}
        for s0, s1 in self._segments():
            ckpts[s0] = psi.clone()
            self._forward_slices(psi, params, s0, s1, work)
        return psi, ckpts

    def _store_steps(self, longest):
        """Sub-steps whose Chebyshev vectors fit the store budget (≥ 1)."""
        import torch
        per = self.K * self.dim * (8 if self.tdt == torch.complex64 else 16)
        budget = self.store_budget_bytes
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self.plan.device)
            budget = free // 2
        return max(1, min(longest, int(budget // per)))

    def adjoint(self, g, ckpts, params):
        """Backward sweep from g = dL/d(conj ψ(T)) (device, overwritten with
        dL/d(conj ψ0)).  Each checkpoint segment is recomputed once with its
        Chebyshev vectors kept (``ryd_evolve_forward_store``) and then run
        backwards (``ryd_evolve_adjoint_stored``); the gradient dots stay on
        the device until the end.  Returns the (sub-steps, K, 4) FP64 sums."""
        import torch
        P = self.plan
        Omv, dadd, cs, rs, coef = params
        M = len(Omv)
        L = nat.lib()
        code = nat.dtype_code(self.tdt)
        segs = self._segments()
        chunk = self._store_steps(max(s1 - s0 for s0, s1 in segs))
        tstore = torch.empty(chunk * self.K * self.dim, dtype=self.tdt, device=P.device)
        work = torch.empty(4 * self.dim, dtype=self.tdt, device=P.device)
        nbytes = int(L.ryd_evolve_partials_bytes(P.handle, code, self.K))
        if nbytes <= 0:
            raise NotImplementedError('the adjoint runs on the fused full-basis kernels '
                                      '(Omega/delta/V, n <= 22)')
        partials = torch.empty(nbytes // 8, dtype=torch.float64, device=P.device)
        sums = torch.zeros(M * self.K * 4, dtype=torch.float64, device=P.device)

        def arr(x, lo, hi):
            return np.ascontiguousarray(x[lo:hi])

        for s0, s1 in reversed(segs):
            # chunks of the segment from its end; a chunk that does not start
            # at the checkpoint is reached by a plain forward from it
            c1 = s1
            while c1 > s0:
                c0 = max(s0, c1 - chunk)
                psi = ckpts[s0].clone()
                if c0 > s0:
                    self._forward_slices(psi, params, s0, c0, work)
                om_a, da_a, c_a, r_a = (arr(Omv, c0, c1), arr(dadd, c0, c1), arr(cs, c0, c1),
                                        arr(rs, c0, c1))
                cf = np.ascontiguousarray(coef[c0:c1])
                nat.check(L.ryd_evolve_forward_store(
                    P.handle, code, nat.ptr(psi), c1 - c0, om_a.ctypes.data, da_a.ctypes.data,
                    c_a.ctypes.data, r_a.ctypes.data, self.K, cf.ctypes.data, nat.ptr(tstore),
                    nat.ptr(work), nat.stream_ptr()))
                nat.check(L.ryd_evolve_adjoint_stored(
                    P.handle, code, nat.ptr(g), c1 - c0, om_a.ctypes.data, da_a.ctypes.data,
                    c_a.ctypes.data, r_a.ctypes.data, self.K, cf.ctypes.data, nat.ptr(tstore),
                    nat.ptr(work), nat._vp(sums.data_ptr() + 8 * c0 * self.K * 4), nat.ptr(partials),
                    nat.stream_ptr()))
                c1 = c0
        return sums.view(M, self.K, 4).cpu().numpy()

    def grad_from_sums(self, all_sums, rs):
        """(dL/dΩ_k, dL/dΔ_k) per slice from the adjoint sums."""
        w = np.full(self.K, 2.0)
        w[0] = 1.0
        fac = 2.0 / rs
        d_om_sub = fac * (all_sums[:, :, 0] @ w)
        d_dadd_sub = fac * (all_sums[:, :, 1] @ w)
        d_om = np.bincount(self._owner, weights=d_om_sub, minlength=self._nslices)
        d_dadd = np.bincount(self._owner, weights=d_dadd_sub, minlength=self._nslices)
        return d_om, -d_dadd

    def loss_and_grad(self, psi0, Om, Delta, target_index):
        """L = 1 − |<target|ψ(T)>|² and (dL/dΩ_k, dL/dΔ_k) by the adjoint."""
        import torch
        params = self._slice_params(Om, Delta)
        psi, ckpts = self.forward_checkpoints(psi0, params)
        o = psi[int(target_index)].item()
        loss = 1.0 - abs(o) ** 2
        # cotangent dL/d(conj psi_T) = -o e_target
        g = torch.zeros(self.dim, dtype=self.tdt, device=self.plan.device)
        g[int(target_index)] = -o
        all_sums = self.adjoint(g, ckpts, params)
        d_om, d_de = self.grad_from_sums(all_sums, params[3])
        return loss, d_om, d_de, psi


def pulse_loss_and_grad(problem, psi0, Om, Delta, target_index):
    return problem.loss_and_grad(psi0, Om, Delta, target_index)


def propagate(problem, psi0, Om, Delta):
    return problem.propagate(psi0, Om, Delta)


def _autograd_function():
    import torch

    class PulseEvolution(torch.autograd.Function):
        """loss = 1 − |<target|ψ(T)>|² with a hand-written adjoint backward.

        ``PulseEvolution.apply(Om, Delta, problem, psi0, target_index)`` with
        Om, Delta 1-D float64 tensors (per-slice drive); returns a 0-d tensor."""

        @staticmethod
        def forward(ctx, Om, Delta, problem, psi0, target_index):
            # forward sweep only: the loss and the checkpoint states; the
            # adjoint runs in backward
            params = problem._slice_params(Om.detach().cpu().numpy(),
                                           Delta.detach().cpu().numpy())
            psi, ckpts = problem.forward_checkpoints(psi0, params)
            o = psi[int(target_index)].item()
            ctx.problem = problem
            ctx.params = params
            ctx.ckpts = ckpts
            ctx.o = o
            ctx.target = int(target_index)
            ctx.dtypes = (Om.dtype, Delta.dtype)
            return torch.tensor(1.0 - abs(o) ** 2, dtype=Om.dtype)

        @staticmethod
        def backward(ctx, grad_out):
            problem = ctx.problem
            g = torch.zeros(problem.dim, dtype=problem.tdt, device=problem.plan.device)
            g[ctx.target] = -ctx.o
            all_sums = problem.adjoint(g, ctx.ckpts, ctx.params)
            ctx.ckpts = None
            gO, gD = problem.grad_from_sums(all_sums, ctx.params[3])
            gO = torch.as_tensor(gO, dtype=ctx.dtypes[0])
            gD = torch.as_tensor(gD, dtype=ctx.dtypes[1])
            return grad_out * gO, grad_out * gD, None, None, None

    return PulseEvolution


class _Lazy:
    _fn = None

    def __getattr__(self, name):
        if _Lazy._fn is None:
            _Lazy._fn = _autograd_function()
        return getattr(_Lazy._fn, name)


PulseEvolution = _Lazy()


def z2_index(n):
    """Code of |r g r g ...> (site 0 excited; site 0 is the most significant bit)."""
    return sum(1 << (n - 1 - i) for i in range(0, int(n), 2))


# ---------------------------------------------------------------------------
# piecewise-constant segments with per-site drives (reference engine_ed.py:130-147)
# ---------------------------------------------------------------------------
def segment_intervals(a, b, segments):
    """Split [a, b] at every segment start strictly inside it (the rule of
    tomisnot/energy-level src/core/manybody/runtime.py:56-60)."""
    cuts = sorted({float(s['t0']) for s in segments or [] if a < float(s['t0']) < b})
    edges = [float(a)] + cuts + [float(b)]
    return list(zip(edges[:-1], edges[1:]))


def _segment_arrays(segment, n):
    om = np.asarray(segment.get('Om', np.zeros(n)), float).reshape(-1)
    ph = np.asarray(segment.get('phase', np.zeros(n)), float).reshape(-1)
    de = np.asarray(segment.get('delta', np.zeros(n)), float).reshape(-1)
    if om.size != n or ph.size != n or de.size != n:
        raise ValueError('segment_drive: Om/phase/delta of every segment need length n')
    return om * np.exp(1j * ph), de


def segment_drive(t, segments, n, interp='left'):
    """(Om complex[n], delta[n]) at time t: the latest segment start <= t
    ('left'), or the mean of that segment and the next ('midpoint'); the
    rule of runtime.py:90-114."""
    if not segments:
        return None
    if interp not in ('left', 'midpoint'):
        raise ValueError("interp must be 'left' or 'midpoint', got %r" % (interp,))
    ordered = sorted(segments, key=lambda s: float(s['t0']))
    idx = 0
    for k, seg in enumerate(ordered):
        if float(seg['t0']) <= t:
            idx = k
    om, de = _segment_arrays(ordered[idx], n)
    if interp == 'midpoint' and idx < len(ordered) - 1:
        om2, de2 = _segment_arrays(ordered[idx + 1], n)
        om, de = 0.5 * (om + om2), 0.5 * (de + de2)
    return om, de


def taylor_bound(model: RydModel, gamma):
    """Row-sum (Gershgorin) bound of ‖H − (i/2)·Σ γ_i n_i‖ for the Taylor
    step count: |diagonal| ≤ Σ|δ| + Σ|V| + Σγ/2, off-diagonal ≤ Σ|Ω|/2 +
    Σ(|J| + |A|) per row."""
    g = np.abs(np.asarray(gamma, float))
    return float(np.sum(np.abs(model.delta)) + np.sum(np.abs(model.pair_v)) + 0.5 * np.sum(g)
                 + 0.5 * np.sum(np.abs(model.Om)) + np.sum(np.abs(model.pair_J))
                 + np.sum(np.abs(model.pair_A)))


def propagate_lossy(plan, model, psi, dt, gamma, work=None):
    """psi ← exp(−i·CONV·dt·(H − (i/2)·Σ_i g_i n_i)) psi on the device
    (``ryd_evolve_taylor``), g_i = γ_i·1e-3/CONV: the non-Hermitian drive of
    engine_ed._propagate_segments (engine_ed.py:130-147, expm_multiply with
    the diagonal −(i/2)·γ_i·1e-3/CONV·n_i of channels.py:58-77; γ in 1/µs,
    so |ψ|² of an excited atom decays as exp(−γ·t·1e-3), t in ns), by Taylor
    steps with ‖h·(H − (i/2)G)‖ ≤ 1."""
    import torch
    gamma = np.asarray(gamma, np.float64)
    if gamma.shape != (plan.n,):
        raise ValueError('gamma needs one entry per atom')
    gamma = np.ascontiguousarray(gamma * 1e-3 / CONV)      # MHz, like delta
    code = plan.check_vec(psi)
    if code not in (nat.RYD_C64, nat.RYD_C128):
        raise ValueError('lossy evolution needs a complex vector')
    tau = CONV * float(dt)
    nsub = max(1, int(np.ceil(tau * taylor_bound(model, gamma))))
    if work is None or work.numel() < 2 * plan.dim or work.dtype != psi.dtype:
        work = torch.empty(2 * plan.dim, dtype=psi.dtype, device=psi.device)
    nat.check(nat.lib().ryd_evolve_taylor(plan.handle, code, nat.ptr(psi), tau, nsub,
                                          TAYLOR_DEGREE, gamma.ctypes.data, nat.ptr(work),
                                          nat.stream_ptr()))
    return psi


def propagate_segments_hip(calib, segments, t, psi0=None, terms=None, interp='left',
                           dtype='complex128', device='cuda', return_states=False,
                           basis_probabilities_max_dim=4096, basis_indices=None,
                           blockade_edges=None, plan=None, pauli_strings=None, corr=False,
                           on_state=None, psi0_plan_order=False, gamma=None):
    """Exact piecewise-constant evolution, output at the times ``t`` (ns):
    every output interval is split at the segment starts and each piece
    [a, b] is propagated by exp(−i·CONV·H·(b−a)) with the drive of
    ``segment_drive((a+b)/2)`` (engine_ed.py:130-147), here by the device
    Chebyshev series (accurate to 1e-13 per piece; one plan per distinct
    drive, as the reference caches one CSR per drive).

    The basis is the full 2^n one, or the projected basis of a
    ``HamiltonianPlan`` (``plan`` with ``projected``; or ``basis_indices`` /
    ``blockade_edges`` directly), as engine_ed.py:191-193 restricts to
    ``plan['states']``: graph bases are enumerated on the device.

    ``psi0``: None (|g…g>, engine_ed.py:165), a vector of the basis
    dimension (in the order of ``basis_indices``), or a full 2^n vector,
    restricted to the basis and normalised by its full norm (weight outside
    the basis is an error, as state.py:112-140).

    ``pauli_strings`` (specs as ``normalize_pauli_strings``) and ``corr``
    add the per-output-time reductions of engine_ed.py:259-281 on the device:
    ``pauli`` (len(t), strings) complex and ``corr`` (len(t), n, n).
    ``on_state(k, plan, x)`` is called at every output time k with the device
    state in the plan's (ascending) basis order.  ``gamma`` (per atom, MHz)
    adds the loss diagonal −(i/2)Γ_i n_i; each piece is then propagated by
    the device Taylor series (``propagate_lossy``) instead of Chebyshev.

    Returns {t, pops_atom (len(t), n), norms (len(t),), psi (final state,
    host complex128, basis order), method='manybody', backend='hip-chebyshev',
    basis_states (codes, projected runs only)[, states][, basis_probabilities]
    [, pauli, pauli_labels][, corr]}."""
    import torch
    from .lanczos import site_occupations, restricted_plan
    from .basis import sector_codes
    nat.require_gpu()
    n = len(calib['atoms'])
    t = np.asarray(t, float)
    if t.ndim != 1 or t.size < 1 or np.any(np.diff(t) < 0):
        raise ValueError('t must be a non-decreasing 1-D time grid')
    if dtype not in ('complex64', 'complex128'):
        raise ValueError('dtype must be complex64 or complex128')
    if plan is not None and plan.get('projected'):
        if basis_indices is None and blockade_edges is None:
            st = plan.get('states')
            blockade_edges = plan.get('blockade_edges')
            if st is not None and not (hasattr(st, 'cpu') and blockade_edges is not None):
                basis_indices = st
    projected = basis_indices is not None or blockade_edges is not None
    tdt = torch.complex64 if dtype == 'complex64' else torch.complex128
    dv = torch.device(device)
    from .observables import normalize_pauli_strings, PauliPlanHip, corr_matrix_hip
    specs = normalize_pauli_strings(pauli_strings, n) if pauli_strings is not None else None
    plans = {}
    pops, norms, states, probs, pauli_series, corr_series = [], [], [], [], [], []
    order = None            # plan order -> caller order of basis_indices
    pauli_plan = None

    def plan_for(om, de):
        nonlocal order
        key = (tuple(np.round(om.real, 15)), tuple(np.round(om.imag, 15)), tuple(de))
        if key not in plans:
            model = from_calib(calib, om, de, None, terms)
            if projected:
                P, order = restricted_plan(model, basis_indices, blockade_edges, device=dv)
            else:
                P = Plan(model, device=dv)
            plans[key] = (model, P)
        return plans[key]

    def record(p, x):
        nonlocal pauli_plan
        norms.append(float(torch.linalg.vector_norm(x.to(torch.complex128)).item()))
        pops.append(np.asarray(site_occupations(p, x)))
        if specs:
            if pauli_plan is None:
                pauli_plan = PauliPlanHip(p, specs)
            pauli_series.append(pauli_plan.expect(x))
        if corr:
            corr_series.append(corr_matrix_hip(p, x)[0])
        if on_state is not None:
            on_state(len(norms) - 1, p, x)
        if return_states or (x.numel() <= basis_probabilities_max_dim):
            v = x.to(torch.complex128).cpu().numpy()
            if order is not None:
                back = np.empty_like(v)
                back[order] = v
                v = back
            if return_states:
                states.append(v)
            if x.numel() <= basis_probabilities_max_dim:
                probs.append(np.abs(v) ** 2)

    try:
        first = segment_drive(float(t[0]), segments, n, interp) if segments else None
        if first is None:
            raise ValueError('segments are required')
        _, p0 = plan_for(*first)
        dim = p0.dim
        codes = sector_codes(p0).cpu().numpy() if projected else None
        if psi0 is None:
            psi = torch.zeros(dim, dtype=tdt, device=dv)
            if projected and (codes.size == 0 or codes[0] != 0):
                raise ValueError('|g...g> is not in the projected basis; give psi0')
            psi[0] = 1.0
        else:
            v = psi0.detach().cpu().numpy() if torch.is_tensor(psi0) else np.asarray(psi0)
            v = np.asarray(v, complex).reshape(-1)
            if v.size == dim:
                if order is not None and not psi0_plan_order:
                    v = v[order]
            elif projected and v.size == (1 << n):
                norm = np.linalg.norm(v)
                inside = v[codes]
                if norm > 0 and abs(np.linalg.norm(inside) - norm) > 1e-12 * norm:
                    raise ValueError('initial state has weight outside the projected basis')
                v = inside / norm if norm > 0 else inside
            else:
                raise ValueError('psi0 must have %d entries%s'
                                 % (dim, ' (or 2^n)' if projected else ''))
            psi = torch.as_tensor(v, device=dv).to(tdt)
        record(p0, psi)
        lossy = gamma is not None and bool(np.any(np.asarray(gamma, float) != 0.0))
        work = None
        for k in range(1, len(t)):
            for a, b in segment_intervals(t[k - 1], t[k], segments):
                if b <= a:
                    continue
                om, de = segment_drive(0.5 * (a + b), segments, n, interp)
                model, P = plan_for(om, de)
                if lossy:
                    if work is None:
                        work = torch.empty(2 * P.dim, dtype=tdt, device=dv)
                    psi = propagate_lossy(P, model, psi.contiguous(), b - a, gamma, work)
                    continue
                prob = PulseProblem(model, dt=b - a, K=15, dtype=dtype, plan=P,
                                    units='reference')
                psi = prob.propagate(psi, [1.0], [0.0])
            record(p0, psi)
        fin = psi.to(torch.complex128).cpu().numpy()
        if order is not None:
            back = np.empty_like(fin)
            back[order] = fin
            fin = back
        out = dict(t=t, pops_atom=np.asarray(pops), norms=np.asarray(norms), psi=fin,
                   method='manybody', backend='hip-chebyshev')
        if projected:
            out['basis_states'] = (np.asarray(basis_indices.cpu() if hasattr(basis_indices, 'cpu')
                                              else basis_indices, np.int64)
                                   if basis_indices is not None else codes)
        if specs is not None:
            out['pauli'] = np.asarray(pauli_series, complex).reshape(len(t), len(specs))
            out['pauli_labels'] = [s_['label'] for s_ in specs]
        if corr:
            out['corr'] = np.asarray(corr_series)
        if return_states:
            out['states'] = np.asarray(states)
        if probs:
            out['basis_probabilities'] = np.asarray(probs)
        return out
    finally:
        for _, P in plans.values():
            P.close()
