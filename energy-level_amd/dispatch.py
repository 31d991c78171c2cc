"""The many-body equilibrium dispatch for the HIP backends, shaped like the
reference's ``solve_equilibrium`` (tomisnot/energy-level
src/core/manybody/eq_route.py:340-490) so that its callers switch backend
names and keep every result key.

Backends (declared like ``EQ_BACKENDS``, backends.py:112-118):

* ``hip-lanczos`` (gs): matrix-free device Lanczos on the full basis, or on
  the states of a projected ``system['plan']`` (the reference's gpu-lanczos
  refuses projected plans, eq_route.py:689-691); keys of
  ``torch_lanczos_ground_state`` plus ``n_atom`` (eq_route.py:689-699).
* ``hip-kpm`` (thermal, dos): matrix-free KPM on the full basis or a
  projected plan's basis; no host CSR is assembled (the reference builds one
  through ``_system_H``, eq_route.py:73-74, :743, :774).
* ``hip-batch`` (gs): the batched solver on this one system
  (eq_route.py:702-709 returns E0, gap, n_atom).

The result carries ``want``, ``backend``, ``n_sites`` (the reference's
``base``), the solver keys, and what ``_equilibrium_result``
(eq_route.py:145-160) adds: ``requested_backend``, ``resolved_backend`` and,
when ``vec`` is returned, ``basis`` = {bit_order: 'site0-msb', indices, dim,
n_sites}.

The source-spectrum recurrence of ``spectrum.lanczos`` (spectrum.py:34-103)
and its post-processing ``quadrature`` / ``lorentz_spectrum``
(spectrum.py:106-135) are mirrored at the end.
"""
from __future__ import annotations

import time

import numpy as np

__all__ = ['EQ_BACKENDS_HIP', 'normalize_backend', 'solve_equilibrium_hip',
           'spectrum_lanczos_hip', 'quadrature', 'lorentz_spectrum', 'source_spectrum_hip',
           'resolve_source']

EQ_BACKENDS_HIP = {
    'hip-lanczos': ('gs', 'spectrum'),
    'hip-kpm': ('thermal', 'dos'),
    'hip-batch': ('gs',),
}
_ALIASES = {'lanczos-hip': 'hip-lanczos', 'kpm-hip': 'hip-kpm', 'batch-hip': 'hip-batch'}


def normalize_backend(name, want):
    """Canonical backend name; ValueError for an unknown name or an
    unsupported ``want`` (as normalize_equilibrium_backend, backends.py:169-174)."""
    name = str(name or '').lower()
    name = _ALIASES.get(name, name)
    if name not in EQ_BACKENDS_HIP:
        raise ValueError('unknown HIP backend %r; valid: %s' % (name, sorted(EQ_BACKENDS_HIP)))
    if want not in EQ_BACKENDS_HIP[name]:
        raise ValueError('backend %r does not support want=%r' % (name, want))
    return name


def _projected(system):
    """(states, blockade_edges) of a projected plan, (None, None) for the full
    basis.  ``states`` may be a host list/array or a device tensor
    (``build_plan_hip``); the edges let the device enumerate and rank the
    basis instead of searching the list."""
    plan = system.get('plan') or {}
    if not plan.get('projected'):
        return None, None
    states = plan.get('states')
    edges = plan.get('blockade_edges')
    if states is None and edges is None:
        raise ValueError('a projected plan needs its states or blockade_edges')
    return states, edges


def _host_states(states):
    if states is None:
        return None
    if hasattr(states, 'cpu'):
        return states.cpu().numpy().astype(np.int64)
    return np.asarray(states, dtype=np.int64)


def solve_equilibrium_hip(system, want='gs', *, backend='hip-lanczos', betas=None, num=400,
                          nvec=200, seed=0, dtype='float64', return_state=False,
                          memory_budget_mb=None, margin=1e-3, krylov_dim=None, sector=None):
    """``solve_equilibrium`` for the HIP backends (see module docstring).

    ``system``: {'calib', 'Om', 'delta'[, 'gamma', 'terms', 'plan', 'device',
    'krylov_dim', 'n_sites', 'charge_sector', 'source', 'psi0', 'spectrum']};
    defaults and checks follow eq_route.py:340-490 (gamma must be 0; betas
    non-empty, finite, positive; ``sector`` / ``charge_sector`` fix the
    excitation number for gs and spectrum only).  Every result carries
    ``data_schema`` (results.py:11-137)."""
    from .results import describe_result
    requested = backend
    backend = normalize_backend(backend, want)
    system = dict(system)
    if 'calib' not in system:
        raise ValueError("HIP backends need system['calib']")
    calib = system['calib']
    n = int(system.get('n_sites') or len(calib['atoms']))
    gamma = system.get('gamma')
    if gamma is not None and np.any(np.asarray(gamma) != 0):
        raise ValueError('equilibrium needs a Hermitian Hamiltonian (gamma = 0)')
    device = system.get('device', 'cuda')
    base = dict(want=want, backend=backend, n_sites=n)
    states, edges = _projected(system)
    projected = states is not None or edges is not None
    charge = system.get('charge_sector')
    if charge is not None:
        # the product-level spelling of sector (eq_route.py:421-430)
        if isinstance(charge, (bool, np.bool_)) or int(charge) != charge:
            raise ValueError('charge_sector must be an integer')
        if sector is not None and int(sector) != int(charge):
            raise ValueError('sector and charge_sector disagree')
        sector = int(charge)
    if sector is not None:
        if want not in ('gs', 'spectrum'):
            raise ValueError('a fixed sector serves gs and spectrum only; thermal/DOS would '
                             'silently ignore it')
        if projected:
            raise ValueError('give either a projected plan or a sector, not both')
        from .basis import check_number_conserving, number_sector_basis
        check_number_conserving(calib, system.get('Om'), system.get('terms'))
        states = number_sector_basis(n, sector)
        projected = True
    if want == 'spectrum':
        if projected and sector is None:
            raise ValueError('want=spectrum takes the full basis or a number sector')
        result = _solve_spectrum(system, base, calib, n, sector, states, device,
                                 memory_budget_mb, seed, krylov_dim)
    elif want == 'gs':
        if backend == 'hip-lanczos':
            from .lanczos import hip_lanczos_ground_state
            # a projected plan's basis: the blockade graph enumerated on the
            # device when its edges are given (and agree with the states),
            # else the states as a code list
            dev_states = hasattr(states, 'cpu')
            out = hip_lanczos_ground_state(
                calib, system.get('Om'), system.get('delta'), gamma, system.get('terms'),
                basis_indices=None if (dev_states and edges is not None) else states,
                blockade_edges=edges,
                m=int(krylov_dim or system.get('krylov_dim', 48)), seed=seed, dtype=dtype,
                device=device, return_state=return_state, memory_budget_mb=memory_budget_mb)
            result = dict(base, **out)
        else:
            if projected:
                raise ValueError('hip-batch takes the full basis')
            if return_state:
                raise ValueError('the batched ground-state backend does not return states')
            from .batch import BMAXN, batched_ground_state_hip
            from .model import from_calib
            model = from_calib(calib, system.get('Om'), system.get('delta'), gamma,
                               system.get('terms'))
            if model.has_pair_flips or not model.is_real or model.n > BMAXN:
                # the batch kernels take real Ω/Δ/V instances of up to BMAXN
                # sites; the reference's batch takes any dense H, so one system
                # of another kind is one device Lanczos solve (same keys)
                from .lanczos import hip_lanczos_ground_state
                rl = hip_lanczos_ground_state(
                    calib, system.get('Om'), system.get('delta'), gamma, system.get('terms'),
                    m=int(krylov_dim or system.get('krylov_dim', 48)), seed=seed,
                    device=device, memory_budget_mb=memory_budget_mb)
                result = dict(base, E0=float(rl['E0']), gap=float(rl['gap']),
                              n_atom=np.asarray(rl['n_atom']))
            else:
                rb = batched_ground_state_hip([model], dtype=dtype if dtype in ('float32', 'float64')
                                              else 'float64', seed=seed, device=device)
                result = dict(base, E0=float(rb['E0'][0]), gap=float(rb['gap'][0]),
                              n_atom=rb['n_atom'][0], converged=bool(rb['converged'][0]),
                              ritz_bound=float(rb['ritz_bound'][0]),
                              ritz_bound_gap=float(rb['ritz_bound_gap'][0]),
                              tolerance=float(rb['tolerance']))
                if 'warning' in rb:
                    result['warning'] = rb['warning']
    else:
        from .kpm import kpm_moments_hip, spectral_bounds, thermal_from_moments, dos_from_moments
        from .model import from_calib
        from .plan import Plan
        if dtype not in ('float32', 'float64'):
            raise ValueError('hip-kpm dtype must be float32/float64')
        model = from_calib(calib, system.get('Om'), system.get('delta'), gamma, system.get('terms'))
        if projected:
            # the projected H of compile_sparse(plan) (eq_route.py:73-74), matrix-free
            from .lanczos import restricted_plan
            dev_states = hasattr(states, 'cpu')
            P, _ = restricted_plan(model, None if (dev_states and edges is not None) else states,
                                   edges, device=device)
        else:
            P = Plan(model, device=device)
        try:
            emin, emax, c, r = spectral_bounds(P, margin=margin)
            mu, c, r = kpm_moments_hip(plan=P, num=num, nvec=nvec, seed=seed, dtype=dtype,
                                       c=c, r=r)
            if want == 'thermal':
                b = np.atleast_1d(np.asarray([1.0] if betas is None else betas, float))
                if b.ndim != 1 or not len(b) or not np.isfinite(b).all() or np.any(b <= 0):
                    raise ValueError('betas must be a non-empty array of finite positive numbers')
                th = thermal_from_moments(P.dim, b, mu, c, r, num)
                result = dict(base, betas=th['betas'], E=th['E'], C=th['C'], F=th['F'])
            else:
                # energy grid of eq_route.py:776-779 over the spectral range
                span = max(emax - emin, 1e-9)
                energies = np.linspace(emin + span * 1e-6, emax - span * 1e-6, 400)
                result = dict(base, energies=energies,
                              rho=dos_from_moments(mu, c, r, energies, num))
        finally:
            P.close()
    result.setdefault('requested_backend', requested)
    result.setdefault('resolved_backend', result.get('backend'))
    if result.get('vec') is not None:
        indices = _host_states(states)
        if indices is None and edges is not None:
            from .basis import independent_states_hip
            indices = independent_states_hip(n, edges, device=device).cpu().numpy()
        result['basis'] = dict(bit_order='site0-msb', indices=indices, dim=len(result['vec']),
                               n_sites=n)
    if sector is not None and want == 'gs':
        result.setdefault('sector_idx', np.asarray(states))
        result.setdefault('dim_sub', len(states))
    return describe_result(result, calib)


def _solve_spectrum(system, base, calib, n, sector, idx, device, memory_budget_mb, seed,
                    krylov_dim):
    """``want='spectrum'`` (``_solve_spectrum``, eq_route.py:260-337): the
    source state phi = O|0> from ``system['source']`` (default the initial
    state itself when ``system['psi0']`` is given), |0> the given initial
    state or the device ground state, then the device Lanczos spectrum of
    phi with shift E0.  The basis is all 2^N codes or the number sector."""
    import torch
    from .lanczos import hip_lanczos_ground_state, restricted_plan
    from .manybody import state_vector_np
    from .model import from_calib
    from .plan import Plan
    opts = dict(system.get('spectrum') or {})
    source = system.get('source')
    if source is None and system.get('psi0') is not None:
        source = {'kind': 'psi0'}
    if source is None:
        raise ValueError('want=spectrum needs system["source"] or system["psi0"] '
                         '(vector/psi0/diagonal/site_weights/sparse/full_sparse)')
    gamma = system.get('gamma')
    dim = len(idx) if idx is not None else (1 << n)
    psi0 = system.get('psi0')
    if psi0 is not None and not isinstance(psi0, (np.ndarray, dict)) and not (
            isinstance(psi0, str) and psi0 == 'ground'):
        psi0 = state_vector_np(psi0, n, idx)
    if isinstance(psi0, np.ndarray) and psi0.ndim == 1:
        if psi0.size == (1 << n) and psi0.size != dim:
            # a full-space vector for a sector run: only when it lies inside
            full = np.asarray(psi0, complex)
            psi0 = full[np.asarray(idx, int)]
            if np.linalg.norm(psi0) <= 0:
                raise ValueError('the initial state has no weight in the fixed sector')
            if abs(np.linalg.norm(psi0) - np.linalg.norm(full)) > 1e-10:
                raise ValueError('the initial state has weight outside the fixed sector; it is '
                                 'not projected silently')
        gs = {'vec': np.asarray(psi0, complex), 'E0': float('nan')}
    else:
        g = hip_lanczos_ground_state(calib, system.get('Om'), system.get('delta'), gamma,
                                     system.get('terms'), basis_indices=idx,
                                     m=int(krylov_dim or system.get('krylov_dim', 48)),
                                     seed=seed, device=device, return_state=True,
                                     memory_budget_mb=memory_budget_mb)
        gs = {'vec': g['vec'], 'E0': float(g['E0'])}
    vec = np.asarray(gs['vec'], complex).reshape(-1)
    if vec.size != dim:
        raise ValueError('the initial state has %d entries, the basis %d' % (vec.size, dim))
    phi, info = resolve_source(source, vec, n, idx)
    model = from_calib(calib, system.get('Om'), system.get('delta'), gamma, system.get('terms'))
    if idx is None:
        P, order = Plan(model, device=device), None
    else:
        P, order = restricted_plan(model, idx, device=device)
    try:
        if order is not None:
            phi = phi[np.asarray(order)]
        real = P.is_real and not np.any(phi.imag)
        phi_t = torch.from_numpy(np.ascontiguousarray(phi.real if real else phi))
        e0 = gs['E0']
        out = source_spectrum_hip(
            P, phi_t, shift=e0 if np.isfinite(e0) else 0.0,
            eta=float(opts.get('eta') or 0.05),
            n_krylov=int(opts.get('n_krylov') or 0) or None,
            omega_min=opts.get('omega_min'), omega_max=opts.get('omega_max'),
            n_omega=int(opts.get('n_omega') or 800), reortho=bool(opts.get('reortho', True)),
            memory_budget_mb=memory_budget_mb)
    finally:
        P.close()
    errors = dict(out.pop('error_report') or {})
    errors.update({'source_kind': info.get('kind'),
                   'sector': None if sector is None else int(sector), 'dim_basis': dim})
    for key, value in info.items():
        if key not in ('kind', 'dim'):
            errors.setdefault(key, value)
    e0 = gs['E0']
    # no scalar gap for a spectrum (the reference reports None, not 0.0)
    return dict(base, backend='hip-lanczos', E0=float(e0) if np.isfinite(e0) else None, gap=None,
                evals=out['evals'], spectrum_weights=out['spectrum_weights'],
                energies=out['omega'], spectral=out['spectral'], source_norm=out['source_norm'],
                eta=out['eta'], krylov=out['krylov'], error_report=errors,
                **({'sector_idx': np.asarray(idx), 'dim_sub': dim} if idx is not None else {}))


def _budget_bytes(memory_budget_mb):
    """Bytes of a MiB budget; None for no limit; ValueError unless finite and
    positive (resources.budget_bytes, resources.py:5-11)."""
    if memory_budget_mb is None:
        return None
    value = float(memory_budget_mb)
    if not np.isfinite(value) or value <= 0:
        raise ValueError('memory_budget_mb must be finite and positive')
    return int(value * 2 ** 20)


def spectrum_lanczos_hip(plan, phi, n_krylov=200, *, reortho=True, memory_budget_mb=512.0,
                         tiny=1e-12):
    """Lanczos tridiagonal from a source vector ``phi`` on the device, with
    the keys and rules of ``spectrum.lanczos`` (spectrum.py:34-103): alphas
    (k,), betas (k-1,), norm2, n, converged (beta <= ``tiny`` or n = dim),
    ortho_loss (|<q_0, q>| of the first and the last vector), reortho, note,
    seconds.

    ``reortho`` (the reference default): two Gram-Schmidt sweeps of every new
    vector against all stored ones (``ryd_lanczos_reortho``: the basis stays
    in device memory, every coefficient on the device).  As the reference
    does, the k·dim·16·2-byte workspace of that basis is weighed against
    ``memory_budget_mb``: over budget the run falls back to the plain
    three-term recurrence and says so in ``note``."""
    import ctypes
    import torch
    from . import _native as nat
    from .lanczos import lanczos_run
    t0 = time.perf_counter()
    v = phi.reshape(-1)
    if v.numel() == 0:
        raise ValueError('source vector is empty')
    if v.numel() != plan.dim:
        raise ValueError('source vector has %d entries, the basis %d' % (v.numel(), plan.dim))
    cplx = v.is_complex() or not plan.is_real
    v = v.to(plan.device, torch.complex128 if cplx else torch.float64).contiguous()
    sums = torch.zeros(4, dtype=torch.float64, device=plan.device)
    plan.dots(None, v, sums)
    norm = float(np.sqrt(sums[1].item()))
    if not np.isfinite(norm) or norm <= 0:
        raise ValueError('source vector has zero or non-finite norm')
    k = int(n_krylov)
    if k < 1:
        raise ValueError('n_krylov must be positive')
    k = min(k, plan.dim)
    note = ''
    if reortho:
        need = plan.dim * k * 16 * 2
        limit = _budget_bytes(memory_budget_mb)
        if limit is not None and need > limit:
            reortho = False
            note = ('full reorthogonalisation workspace about %.0f MiB > budget %.0f MiB: '
                    'three-term recurrence without reorthogonalisation'
                    % (need / 2 ** 20, float(memory_budget_mb)))
    if not reortho:
        # |u_{j+1}| of the device recurrence is beta_j: the same absolute
        # breakdown threshold as the unit-vector recurrence
        R = lanczos_run(plan, v, k, 1, 0.0, want=1, tiny=tiny, store=False)
        j = int(R.j)
        alphas = np.asarray(R.a[:j], float)
        betas = np.asarray(R.b[:max(j - 1, 0)], float)
        converged = bool(R.stopped or j >= plan.dim)
        # without a stored basis the reference reports no overlap (0.0)
        return {'alphas': alphas, 'betas': betas, 'norm2': norm ** 2, 'n': j,
                'converged': converged, 'ortho_loss': 0.0, 'reortho': False,
                'note': note or 'three-term recurrence without reorthogonalisation',
                'seconds': round(time.perf_counter() - t0, 3)}
    L = nat.lib()
    code = nat.dtype_code(v.dtype)
    basis = torch.empty((k, plan.dim), dtype=v.dtype, device=plan.device)
    basis[0].copy_(v)
    w = torch.empty(plan.dim, dtype=v.dtype, device=plan.device)
    alpha = torch.zeros(k, dtype=torch.float64, device=plan.device)
    nrm = torch.zeros(k + 1, dtype=torch.float64, device=plan.device)
    state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=plan.device)
    gsws = torch.empty(max(1, int(L.ryd_reortho_workspace_bytes(plan.dim, k))), dtype=torch.uint8,
                       device=plan.device)
    nat.check(L.ryd_lanczos_reortho(plan.handle, code, nat.ptr(basis), plan.dim, nat.ptr(w), k,
                                    1.0 / norm, nat.ptr(alpha), nat.ptr(nrm), nat.ptr(state),
                                    float(tiny), 2, nat.ptr(gsws), nat.ptr(plan.workspace()),
                                    nat.stream_ptr()))
    stopped, steps = ctypes.c_int(0), ctypes.c_int64(0)
    nat.check(L.ryd_lanczos_status(nat.ptr(state), ctypes.byref(stopped), ctypes.byref(steps),
                                   nat.stream_ptr()))
    n = int(steps.value) if stopped.value else k
    a = alpha.cpu().numpy()
    nr = nrm.cpu().numpy()
    alphas = np.asarray(a[:n], float)
    betas = np.asarray(nr[1:n], float)
    # |<q_0, q>| with q the last vector: q_{n-1} when the recurrence ended by
    # breakdown or exhausted the basis, else the next one, w / beta_{n-1}
    if stopped.value or n >= plan.dim:
        ortho = float(abs(torch.vdot(basis[0], basis[n - 1]).item()))
    else:
        ortho = float(abs(torch.vdot(basis[0], w).item()) / nr[n]) if nr[n] > 0 else 0.0
    return {'alphas': alphas, 'betas': betas, 'norm2': norm ** 2, 'n': n,
            'converged': bool(stopped.value or n >= plan.dim), 'ortho_loss': ortho,
            'reortho': True, 'note': note, 'seconds': round(time.perf_counter() - t0, 3)}


def quadrature(alphas, betas, norm2, shift=0.0):
    """Tridiagonal → quadrature nodes (relative to ``shift``) and weights
    norm2·|e_0ᵀ v|² summing to norm2 (spectrum.py:106-121)."""
    alphas = np.asarray(alphas, float)
    betas = np.asarray(betas, float)
    if alphas.size == 0:
        return np.zeros(0), np.zeros(0)
    if betas.size:
        T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
        evals, evecs = np.linalg.eigh(T)
    else:
        evals, evecs = alphas.copy(), np.ones((1, 1))
    return evals - float(shift), float(norm2) * np.abs(evecs[0, :]) ** 2


def lorentz_spectrum(alphas, betas, norm2, omega, eta, shift=0.0):
    """S(ω) = −(1/π) Im <φ|(ω + shift + iη − H)⁻¹|φ> by the continued fraction
    of the tridiagonal (spectrum.py:124-135)."""
    alphas = np.asarray(alphas, float)
    betas = np.asarray(betas, float)
    omega = np.asarray(omega, float)
    if alphas.size == 0:
        return np.zeros_like(omega)
    z = omega + float(shift) + 1j * float(eta)
    g = z - alphas[-1]
    for j in range(alphas.size - 2, -1, -1):
        g = z - alphas[j] - betas[j] ** 2 / g
    return -float(norm2) * np.imag(1.0 / g) / np.pi


def source_spectrum_hip(plan, phi, *, shift=0.0, eta=0.05, n_krylov=None, omega_min=None,
                        omega_max=None, n_omega=800, reortho=True, memory_budget_mb=512.0):
    """``source_spectrum`` (spectrum.py:138-180): device Lanczos from ``phi``,
    quadrature nodes/weights, the Lorentzian S(ω) on a grid (by default the
    node range widened by max(5 %, 4η)) and its error report
    (spectral_integral, integral_ratio = ∫S/‖φ‖², which approaches 1 when the
    window is wide enough)."""
    dim = plan.dim
    k = int(n_krylov) if n_krylov else min(dim, 200)
    lz = spectrum_lanczos_hip(plan, phi, k, reortho=reortho, memory_budget_mb=memory_budget_mb)
    nodes, weights = quadrature(lz['alphas'], lz['betas'], lz['norm2'], shift)
    if omega_min is None or omega_max is None:
        lo, hi = (float(nodes.min()), float(nodes.max())) if nodes.size else (0.0, 1.0)
        pad = max(0.05 * (hi - lo), 4.0 * float(eta))
        omega_min = lo - pad if omega_min is None else float(omega_min)
        omega_max = hi + pad if omega_max is None else float(omega_max)
    omega = np.linspace(float(omega_min), float(omega_max), int(n_omega))
    eta = float(eta)
    limit = _budget_bytes(memory_budget_mb)
    if limit is not None and omega.size * 32 > limit:
        raise ValueError('broadened spectrum grid needs %.2f MiB, over the budget %.2f MiB'
                         % (omega.size * 32 / 2 ** 20, limit / 2 ** 20))
    spectral = lorentz_spectrum(lz['alphas'], lz['betas'], lz['norm2'], omega, eta, shift)
    integral = float(np.trapezoid(spectral, omega)) if omega.size > 1 else 0.0
    return {
        'omega': omega, 'spectral': spectral, 'evals': nodes, 'spectrum_weights': weights,
        'source_norm': float(lz['norm2']), 'eta': eta, 'shift': float(shift),
        'krylov': {'n': lz['n'], 'n_requested': k, 'converged': lz['converged'],
                   'ortho_loss': lz['ortho_loss'], 'reortho': lz['reortho'], 'note': lz['note'],
                   'seconds': lz['seconds']},
        'error_report': {'source_norm': float(lz['norm2']), 'spectral_integral': integral,
                         'integral_ratio': (integral / lz['norm2']) if lz['norm2'] else None,
                         'eta_mhz': eta, 'krylov_n': lz['n'], 'ortho_loss': lz['ortho_loss'],
                         'reortho': lz['reortho']},
    }


def resolve_source(source, psi0, n_sites, idx=None):
    """Source specification → (phi, info) in the basis of ``psi0`` (all 2^N
    codes, or the codes ``idx``), the forms of ``_resolve_source``
    (eq_route.py:163-252): vector, psi0, diagonal, site_weights (optionally
    subtract_mean), sparse (COO in the current basis) and full_sparse (COO on
    the 2^N codes, projected onto ``idx`` with the dropped weight reported as
    projected_leak)."""
    import scipy.sparse as sp
    if not isinstance(source, dict):
        raise ValueError('source must be a mapping (kind = vector/diagonal/sparse/full_sparse)')
    kind = str(source.get('kind') or '')
    n_sites = int(n_sites)
    dim = len(idx) if idx is not None else (1 << n_sites)
    psi = np.asarray(psi0, dtype=complex).reshape(-1)
    if psi.size != dim:
        raise ValueError('initial state has %d entries, the basis %d' % (psi.size, dim))

    def coo(block):
        rows = np.asarray(block.get('rows', ()), dtype=int).ravel()
        cols = np.asarray(block.get('cols', ()), dtype=int).ravel()
        vals = np.asarray(block.get('values', ()), dtype=complex).ravel()
        if not (len(rows) == len(cols) == len(vals)):
            raise ValueError('source rows/cols/values differ in length')
        if len(rows) == 0:
            raise ValueError('source sparse structure is empty')
        size = block.get('dim')
        return sp.coo_matrix((vals, (rows, cols)),
                             shape=(int(size or rows.max() + 1), int(size or cols.max() + 1))).tocsr()

    if kind == 'vector':
        v = np.asarray(source.get('values', ()), dtype=complex).reshape(-1)
        if v.size != dim:
            raise ValueError('source vector has %d entries, the basis %d' % (v.size, dim))
        return v, {'kind': kind, 'dim': dim}
    if kind in ('psi0', 'state'):
        return psi.copy(), {'kind': 'psi0', 'dim': dim}
    if kind == 'diagonal':
        d = np.asarray(source.get('values', ()), dtype=complex).reshape(-1)
        if d.size != dim:
            raise ValueError('source diagonal has %d entries, the basis %d' % (d.size, dim))
        return d * psi, {'kind': kind, 'dim': dim}
    if kind in ('site_weights', 'weights'):
        w = np.asarray(source.get('weights', ()), dtype=float).reshape(-1)
        if w.size != n_sites:
            raise ValueError('source site_weights has %d entries, the atoms %d' % (w.size, n_sites))
        codes = np.arange(1 << n_sites, dtype=np.int64) if idx is None else np.asarray(idx, np.int64)
        d = np.zeros(codes.size)
        for i in range(n_sites):                 # site 0 is the most significant bit
            d += w[i] * ((codes >> (n_sites - 1 - i)) & 1)
        d = d.astype(complex)
        info = {'kind': 'site_weights', 'dim': dim, 'sum_weights': float(w.sum()),
                'weight_sq': float((w * w).sum())}
        if source.get('subtract_mean'):
            mean = complex(np.vdot(psi, d * psi))
            d = d - mean
            info['subtracted_mean'] = mean.real
        return d * psi, info
    if kind == 'sparse':
        O = coo(source)
        if O.shape != (dim, dim):
            raise ValueError('source sparse shape %r differs from the basis (%d, %d)'
                             % (O.shape, dim, dim))
        return O @ psi, {'kind': kind, 'dim': dim}
    if kind == 'full_sparse':
        full = 1 << n_sites
        if full > (1 << 16):
            raise ValueError('full_sparse needs 2^N <= 65536; give a large system the source '
                             'in its current (sector) basis')
        O = coo(source)
        if O.shape != (full, full):
            raise ValueError('source full_sparse shape %r differs from (%d, %d)'
                             % (O.shape, full, full))
        if idx is None:
            return O @ psi, {'kind': kind, 'dim': full}
        expand = np.zeros(full, dtype=complex)
        expand[np.asarray(idx, int)] = psi
        phi_full = O @ expand
        phi = phi_full[np.asarray(idx, int)]
        nf = float(np.linalg.norm(phi_full))
        leak = 0.0 if nf == 0 else 1.0 - float(np.linalg.norm(phi)) ** 2 / nf ** 2
        return phi, {'kind': kind, 'dim': dim, 'projected_leak': max(0.0, leak)}
    raise ValueError('unknown source.kind=%r (vector/psi0/diagonal/site_weights/sparse/'
                     'full_sparse)' % kind)
