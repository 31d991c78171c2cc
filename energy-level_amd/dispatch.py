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
           'spectrum_lanczos_hip', 'quadrature', 'lorentz_spectrum']

EQ_BACKENDS_HIP = {
    'hip-lanczos': ('gs',),
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
                          memory_budget_mb=None, margin=1e-3, krylov_dim=None):
    """``solve_equilibrium`` for the HIP backends (see module docstring).

    ``system``: {'calib', 'Om', 'delta'[, 'gamma', 'terms', 'plan', 'device',
    'krylov_dim', 'n_sites']}; defaults and checks follow eq_route.py:340-490
    (gamma must be 0; betas non-empty, finite, positive)."""
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
    if want == 'gs':
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
    return result


def spectrum_lanczos_hip(plan, phi, n_krylov=200, tiny=1e-12):
    """Lanczos tridiagonal from a source vector ``phi`` on the device, with
    the keys of ``spectrum.lanczos`` (spectrum.py:34-103): alphas (k,),
    betas (k-1,), norm2, n, converged, ortho_loss, reortho, note, seconds.

    The device recurrence is the plain three-term one (no full
    reorthogonalisation: that costs O(k²·dim) and k stored vectors);
    ``ortho_loss`` = |<q_0, q_n>| reports how far orthogonality was lost, as
    the reference does after its run."""
    import torch
    from .lanczos import lanczos_run, lanczos_vector
    t0 = time.perf_counter()
    v = phi.reshape(-1)
    if v.numel() == 0:
        raise ValueError('source vector is empty')
    norm = float(torch.linalg.vector_norm(v.to(torch.complex128 if v.is_complex()
                                                 else torch.float64)).item())
    if not np.isfinite(norm) or norm <= 0:
        raise ValueError('source vector has zero or non-finite norm')
    k = int(n_krylov)
    if k < 1:
        raise ValueError('n_krylov must be positive')
    k = min(k, plan.dim)
    R = lanczos_run(plan, v, k, 1, 0.0, want=1, tiny=tiny * norm, store=False)
    j = int(R.j)
    alphas = np.asarray(R.a[:j], float)
    betas = np.asarray(R.b[:max(j - 1, 0)], float)
    converged = bool(R.stopped or j >= plan.dim)
    # q_n = u_n / |u_n|, the last vector the run produced
    ortho = 0.0
    last = R.nrm[j].item() if j < R.nrm.numel() else 0.0
    un = lanczos_vector(R, j)
    if j >= 1 and last > 0 and not R.stopped and un is not None:
        qn = un / last
        ortho = float(abs(torch.vdot(v.to(qn.dtype), qn).item()) / norm)
    return {'alphas': alphas, 'betas': betas, 'norm2': norm ** 2, 'n': j,
            'converged': converged, 'ortho_loss': ortho, 'reortho': False,
            'note': 'device three-term recurrence, no reorthogonalisation',
            'seconds': round(time.perf_counter() - t0, 3)}


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
