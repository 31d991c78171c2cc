"""``evolve_manybody`` on the HIP engine: the reference's ED time-evolution
entry point with its result contract (tomisnot/energy-level
src/core/manybody/engine_ed.py:153-372), on the full basis or a projected
plan's basis.

Propagation is the device Chebyshev series per piecewise-constant interval
(``propagate_segments_hip``; a static drive is one segment).  At every output
time the device reduces: ⟨n_i⟩ and the norm, the bipartite entropy at
``entropy_cut`` (singular values of the Schmidt matrix on the device;
engine_ed.py:259-268, observables.py:50-115), the spin observables
sx, sy, sz, ⟨σ+⟩, ⟨σ−⟩, phase from Pauli-string expectations
(observables.py:117-204: σ− = (X + iY)/2), ⟨σ+_a σ−_b⟩ of ``link_pairs``
and the Pauli strings; at the last time C_ij and the full ⟨σ+_i σ−_j⟩
matrix.  Initial states follow ``state_vector`` (state.py:112-160): None
(|g…g>), an index, a g/r pattern, a 2^N vector, ('sup', components), a
product dict or list of local states, or 'ground' (the device Lanczos ground
state of the drive at t[0]).

``loops`` are post-processed on the host from the measured bonds
(observables.py:206-300).  A non-Hermitian drive (gamma != 0, the loss
diagonal −(i/2)Γ_i n_i) is propagated by the device Taylor series
(``ryd_evolve_taylor``); the norm then decays and ``loss`` = 1 − norm is the
physical loss (engine_ed.py:311-318).
"""
from __future__ import annotations

import numpy as np

__all__ = ['evolve_manybody_hip', 'state_vector_np', 'product_label']

FULLPOP_MAX = 65536          # joint populations up to this dimension (engine_ed.py:47)
ENTROPY_MAX_ENTRIES = 1 << 20   # Schmidt matrix budget of a projected basis (observables.py:89)
FULL_ENTROPY_MAX_N = 26      # full-basis Schmidt SVD up to 2^26 amplitudes


def product_label(n, s):
    """'|g r g …>' for code s, site 0 leftmost (observables.py:83-86)."""
    bits = ''.join('r' if ((s >> (n - 1 - i)) & 1) else 'g' for i in range(n))
    return '|%s>' % bits if n else '|>'


def _scalar(x):
    if isinstance(x, (list, tuple)) and len(x) == 2:
        return complex(float(x[0]), float(x[1]))
    return complex(x)


def _local_vector(x):
    if isinstance(x, str):
        if x == 'g':
            return np.array([1., 0.], complex)
        if x == 'r':
            return np.array([0., 1.], complex)
        raise ValueError('a local two-level state string is g or r')
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        if int(x) not in (0, 1):
            raise ValueError('a local two-level state index is 0 or 1')
        return np.eye(2, dtype=complex)[:, int(x)]
    a = np.asarray([_scalar(v) for v in x], complex) if hasattr(x, '__len__') else None
    if a is None or a.shape != (2,):
        raise ValueError('a local state is [a_g, a_r]')
    nrm = float(np.linalg.norm(a))
    if not np.isfinite(nrm) or nrm <= 0.0:
        raise ValueError('a local state must be finite and non-zero')
    return a / nrm


def _product_vectors(spec, n):
    """Local vectors of a product spec, None for 'ground' (state.py:181-210)."""
    if spec is None:
        return [np.array([1., 0.], complex) for _ in range(n)]
    if isinstance(spec, (int, np.integer)) and not isinstance(spec, bool):
        if not 0 <= int(spec) < (1 << n):
            raise ValueError('initial state index out of range')
        return [_local_vector((int(spec) >> (n - 1 - i)) & 1) for i in range(n)]
    if isinstance(spec, str):
        if spec == 'ground':
            return None
        s = spec.replace('|', '').replace('>', '').replace(' ', '')
        if len(s) != n or any(c not in 'gr' for c in s):
            raise ValueError('a pattern is N characters g/r')
        return [_local_vector(c) for c in s]
    if isinstance(spec, dict):
        if spec.get('kind', 'product') != 'product':
            raise ValueError('unknown product state kind %r' % (spec.get('kind'),))
        raw = spec.get('amplitudes', spec.get('local_states'))
        if raw is None or len(raw) != n:
            raise ValueError('a product state needs N local states')
        return [_local_vector(x) for x in raw]
    if isinstance(spec, (list, tuple)) and len(spec) == n:
        return [_local_vector(x) for x in spec]
    raise ValueError('unsupported initial state spec %r' % (type(spec).__name__,))


def state_vector_np(spec, n, basis=None):
    """Normalised initial state on ``basis`` (codes in the caller's order; None:
    all 2^n codes), as ``state_vector`` (state.py:112-160): weight outside a
    projected basis is an error, never a silent projection."""
    n = int(n)
    projected = basis is not None
    codes = np.arange(1 << n, dtype=np.int64) if basis is None else np.asarray(basis, np.int64)
    if isinstance(spec, np.ndarray):
        full = np.asarray(spec, complex)
        if full.shape != (1 << n,):
            raise ValueError('an initial vector has shape (2**N,)')
        nrm = np.linalg.norm(full)
        vec = full[codes] / nrm if nrm > 0 else full[codes]
    elif (isinstance(spec, (list, tuple)) and spec and isinstance(spec[0], str)
          and spec[0] == 'sup'):
        vec = np.zeros(len(codes), complex)
        for comp in spec[1]:
            if (isinstance(comp, (list, tuple)) and len(comp) == 2
                    and not isinstance(comp[0], (list, tuple, dict))):
                pattern, amp = comp
                vec += _scalar(amp) * state_vector_np(pattern, n, basis)
            else:
                vec += state_vector_np(comp, n, basis)
        nrm = np.linalg.norm(vec)
        if not np.isfinite(nrm) or nrm <= 0:
            raise ValueError('initial state has zero norm or non-finite entries')
        return vec / nrm
    elif isinstance(spec, (int, np.integer)) and not isinstance(spec, bool):
        if not 0 <= int(spec) < (1 << n):
            raise ValueError('initial state index out of range')
        vec = (codes == int(spec)).astype(complex)
    elif (isinstance(spec, (list, tuple)) and len(spec) == (1 << n)
          and not (len(spec) == n and all(isinstance(x, (str, int, np.integer)) for x in spec))):
        return state_vector_np(np.asarray(spec, complex), n, basis)
    else:
        vectors = _product_vectors(spec, n)
        if vectors is None:
            raise ValueError("the 'ground' initial state is prepared by the solver")
        vec = np.ones(len(codes), complex)
        for i, local in enumerate(vectors):
            vec *= local[(codes >> (n - 1 - i)) & 1]
    nrm = float(np.linalg.norm(vec))
    if not np.isfinite(nrm) or nrm <= 0:
        raise ValueError('initial state has zero norm or lies outside the restricted basis')
    if projected and abs(nrm - 1.0) > 1e-10:
        raise ValueError('initial state has weight outside the restricted basis')
    return vec / nrm


def _entropy(plan, x, codes, n, cut, projected):
    """Bipartite entropy (nat) of the device state x at ``cut``; None when the
    Schmidt matrix exceeds the budget (entropy_at_cut / entropy_in_basis)."""
    import torch
    if n < 2:
        return 0.0
    v = x.to(torch.complex128)
    if not projected:
        if n > FULL_ENTROPY_MAX_N:
            return None
        M = v.reshape(1 << cut, 1 << (n - cut))
        w = torch.linalg.svdvals(M) ** 2
        w = w[w > 1e-15]
        return float(-(w * torch.log(w)).sum().item()) if w.numel() else 0.0
    left = codes >> (n - cut)
    right = codes & ((1 << (n - cut)) - 1)
    _, li = torch.unique(left, return_inverse=True)
    _, ri = torch.unique(right, return_inverse=True)
    shape = (int(li.max().item()) + 1, int(ri.max().item()) + 1)
    if shape[0] * shape[1] > ENTROPY_MAX_ENTRIES:
        return None
    M = torch.zeros(shape, dtype=torch.complex128, device=v.device)
    M[li, ri] = v
    w = torch.linalg.svdvals(M) ** 2
    tot = w.sum()
    if not tot.item() > 0:
        return None
    w = w / tot
    w = w[w > 0]
    return float(-(w * torch.log(w)).sum().item())


def _pair_strings(pairs):
    """XX, XY, YX, YY strings of each pair (a, b): <σ+_a σ−_b> =
    (XX + i·XY − i·YX + YY)/4."""
    out = []
    for a, b in pairs:
        for ax in ('x', 'y'):
            for bx in ('x', 'y'):
                out.append({'label': '%s@%d %s@%d' % (ax, a, bx, b), 'terms': [(a, ax), (b, bx)]})
    return out


def _pair_values(vals):
    v = np.asarray(vals, complex).reshape(-1, 4)
    return (v[:, 0] + 1j * v[:, 1] - 1j * v[:, 2] + v[:, 3]) / 4.0


def _stage_progress(arr, stage, frac):
    """The three-slot progress protocol of the reference's engines
    (src/core/coupling.py:24-36, used by engine_ed.py:283-307): arr[0] the
    fraction done within the stage, arr[1] the caller's cancel flag (read
    only), arr[2] the stage (0 preparation, 1 integration, 2 finish); a
    two-slot array gets the fraction only."""
    if arr is None:
        return
    arr[0] = min(max(float(frac), 0.0), 1.0)
    if len(arr) > 2:
        arr[2] = float(stage)


def evolve_manybody_hip(calib, psi0=None, t_grid=None, Om=None, delta=None, gamma=None,
                        terms=None, segments=None, progress_arr=None, log=None, want_corr=True,
                        want_entropy=True, use_sectors='auto', dense_max=None, fullpop_max=None,
                        max_dim=None, entropy_cut=None, record_psi=False, diag=False, plan=None,
                        want_observables=True, loops=None, want_xy_corr=True, link_pairs=None,
                        pauli_strings=None, memory_budget_mb=None, drive_interp='left',
                        dtype='complex128', device='cuda'):
    """``evolve_manybody`` (engine_ed.py:153-372) on the device; same arguments
    and result keys: t, pops, states, method, backend, pops_atom, S, corr,
    loss, norms, basis_probabilities, basis_states, basis, error_report,
    diag, observables[, psi_amp].  ``use_sectors``, ``dense_max`` and
    ``memory_budget_mb`` are accepted for the signature (the device series
    needs neither sectors nor a dense fallback)."""
    import torch
    from .evolve import propagate_segments_hip
    from .observables import normalize_pauli_strings, PauliPlanHip, corr_matrix_hip, \
        loop_observables, link_loop_observables
    from .model import active_terms
    from .basis import sector_codes
    log = log or (lambda *a: None)
    atoms = calib.get('atoms') or []
    n = len(atoms)
    if n == 0:
        raise ValueError('many-body evolution: calib has no atoms')
    if link_pairs is not None:
        link_pairs = [(int(a), int(b)) for a, b in link_pairs]
        if any(a < 0 or a >= n or b < 0 or b >= n for a, b in link_pairs):
            raise ValueError('link_pairs holds an atom index out of range')
    projected = bool(plan is not None and plan.get('projected'))
    states = edges = None
    if projected:
        states, edges = plan.get('states'), plan.get('blockade_edges')
        Om = plan['Om'] if Om is None else Om
        delta = plan['delta'] if delta is None else delta
        gamma = plan.get('gamma') if gamma is None else gamma
        terms = plan.get('terms') if terms is None else terms
    t = np.asarray(t_grid, float) if t_grid is not None else None
    if t is None or t.ndim != 1 or len(t) < 2 or not np.isfinite(t).all() or np.any(np.diff(t) < 0):
        raise ValueError('t_grid needs at least two finite, non-decreasing times')
    gamma = np.zeros(n) if gamma is None else np.asarray(gamma, float)
    if gamma.ndim == 0:
        gamma = np.full(n, float(gamma))
    if gamma.shape != (n,) or not np.all(np.isfinite(gamma)):
        raise ValueError('gamma needs one finite entry per atom')
    hermitian = not bool(np.any(np.abs(gamma) > 0.0))
    Om = np.zeros(n, complex) if Om is None else np.asarray(Om, complex)
    delta = np.zeros(n) if delta is None else np.asarray(delta, float)
    if Om.shape != (n,) or delta.shape != (n,):
        raise ValueError('Om and delta need one entry per atom')
    cut = int(entropy_cut if entropy_cut is not None else max(1, n // 2))
    if want_entropy and n >= 2 and not 1 <= cut <= n - 1:
        raise ValueError('entropy_cut=%d out of range (must be in [1, %d])' % (cut, n - 1))
    if pauli_strings and not want_observables:
        raise ValueError('pauli_strings need want_observables=True')
    specs = normalize_pauli_strings(pauli_strings, n) if pauli_strings else None
    fullpop_max = int(FULLPOP_MAX if fullpop_max is None else fullpop_max)
    if segments:
        segs = segments
    else:
        segs = [dict(t0=float(t[0]), Om=np.abs(Om), phase=np.angle(Om), delta=delta)]
    # the basis in the caller's order (host codes) for the initial state
    if projected:
        if states is not None and not (hasattr(states, 'cpu') and edges is not None):
            basis_host = np.asarray(states.cpu() if hasattr(states, 'cpu') else states, np.int64)
            kw = dict(basis_indices=basis_host)
        else:
            from .basis import independent_states_hip
            basis_host = independent_states_hip(n, edges, device=device).cpu().numpy()
            kw = dict(blockade_edges=edges)
        dim = len(basis_host)
    else:
        basis_host = None
        kw = {}
        dim = 1 << n
    if max_dim is not None and dim > int(max_dim):
        raise ValueError('many-body dimension %d exceeds max_dim %d' % (dim, int(max_dim)))
    if isinstance(psi0, str) and psi0 == 'ground':
        from .evolve import segment_drive
        from .lanczos import hip_lanczos_ground_state
        om0, de0 = segment_drive(float(t[0]), segs, n, interp='left')
        g = hip_lanczos_ground_state(calib, om0, de0, None, terms, device=device,
                                     return_state=True,
                                     basis_indices=basis_host if projected else None,
                                     blockade_edges=edges if projected else None)
        psi0v = np.asarray(g['vec'], complex)
    else:
        psi0v = state_vector_np(psi0, n, basis_host)
    n_t = len(t)
    S = np.full(n_t, np.nan) if want_entropy else None
    S_note = (None if (not want_entropy or n >= 2)
              else 'n_atoms=%d < 2: no bipartition, S = 0 (a physical fact, not missing data)' % n)
    sx = np.zeros((n_t, n))
    sy = np.zeros((n_t, n))
    sz = np.zeros((n_t, n))
    spv = np.zeros((n_t, n), complex)
    smv = np.zeros((n_t, n), complex)
    phase = np.full((n_t, n), np.nan)
    link_series = (np.full((n_t, len(link_pairs)), np.nan + 0j, complex)
                   if link_pairs is not None else None)
    state = {'xy': None, 'corr': None, 'codes': None, 'spin': None, 'link': None, 'xyp': None}
    spin_specs = [{'label': '%s@%d' % (ax, i), 'terms': [(i, ax)]} for i in range(n)
                  for ax in ('x', 'y')]
    upper = [(i, j) for i in range(n) for j in range(i + 1, n)]

    def on_state(k, P, x):
        if progress_arr is not None and float(progress_arr[1]) != 0:
            raise RuntimeError('cancelled')
        _stage_progress(progress_arr, 1, (k + 1) / n_t)
        if want_entropy:
            if projected and state['codes'] is None:
                state['codes'] = sector_codes(P)
            val = _entropy(P, x, state['codes'], n, cut, projected)
            S[k] = np.nan if val is None else val
        if want_observables:
            if state['spin'] is None:
                state['spin'] = PauliPlanHip(P, spin_specs)
            xy = np.asarray(state['spin'].expect(x), complex).reshape(n, 2)
            sm = (xy[:, 0] + 1j * xy[:, 1]) / 2.0           # <σ−> = (<X> + i<Y>)/2
            smv[k], spv[k] = sm, np.conj(sm)
            sx[k], sy[k] = 2.0 * sm.real, 2.0 * sm.imag
            ph = np.angle(sm)
            ph[np.abs(sm) < 1e-14] = np.nan
            phase[k] = ph
            if link_pairs is not None:
                if state['link'] is None:
                    state['link'] = PauliPlanHip(P, _pair_strings(link_pairs))
                link_series[k] = _pair_values(state['link'].expect(x))
            if k == n_t - 1 and want_xy_corr:
                xyc = np.zeros((n, n), complex)
                if upper:
                    vals = _pair_values(PauliPlanHip(P, _pair_strings(upper)).expect(x))
                    for (i, j), v in zip(upper, vals):
                        xyc[i, j] = v
                        xyc[j, i] = np.conj(v)
                state['xyp'] = xyc
        if k == n_t - 1 and want_corr:
            state['corr'] = corr_matrix_hip(P, x)[0]

    _stage_progress(progress_arr, 1, 0.0)
    out = propagate_segments_hip(calib, segs, t, psi0=psi0v, terms=terms, interp=drive_interp,
                                 dtype=dtype, device=device, return_states=False,
                                 basis_probabilities_max_dim=fullpop_max, pauli_strings=specs,
                                 on_state=on_state, gamma=None if hermitian else gamma, **kw)
    norms = out['norms'] ** 2                  # Σ|ψ|², as the reference's p.sum()
    pops_atom = out['pops_atom'] * norms[:, None]
    sz[:] = norms[:, None] - 2.0 * pops_atom
    if want_observables and state['xyp'] is not None:
        np.fill_diagonal(state['xyp'], pops_atom[-1])
    corr = state['corr'] if want_corr else None
    joint = out.get('basis_probabilities')
    loss = 1.0 - norms
    norm_dev = float(np.max(np.abs(norms - 1.0))) if hermitian else 0.0
    if want_entropy and np.isnan(S).any():
        S_note = 'entropy not computed at some output times (Schmidt matrix budget); NaN there'
    capped = joint is None or projected
    if not capped:
        pops = joint
        labels = [{'label': product_label(n, s), 'index': s} for s in range(dim)]
    else:
        pops = pops_atom
        labels = [{'label': 'atom%d@(%.1f,%.1f,%.1f)' % ((a.get('id', i),) + tuple(
            float(v) for v in (a.get('pos') or (0.0, 0.0, 0.0)))), 'atom_id': a.get('id', i)}
                  for i, a in enumerate(atoms)]
        log('[manybody] pops are the per-atom marginals; joint probabilities in basis_probabilities')
    kind = (plan.get('constraint') or {}).get('kind', 'blockade') if projected else None
    backend = ('hip-ed-plan-%s' % kind if projected else
               'hip-ed' if hermitian else 'hip-ed-nonhermitian')
    use = set(active_terms(calib) if terms is None else terms)
    obs = None
    if want_observables:
        xy_last = state['xyp']
        lp, lw = (loop_observables(xy_last, loops) if xy_last is not None and loops is not None
                  else ([], []))
        lps, lws = (link_loop_observables(link_series, link_pairs, loops)
                    if link_series is not None and loops is not None
                    else (np.empty((0, 0)), np.empty((0, 0))))
        obs = dict(sx=sx, sy=sy, sz=sz, sp=spv, sm=smv, phase=phase, xy_corr=xy_last,
                   loop_phase=lp, loop_winding=lw, loop_phase_series=lps,
                   loop_winding_series=lws, link_corr=link_series,
                   link_pairs=link_pairs, pauli_strings=out.get('pauli'),
                   pauli_string_labels=out.get('pauli_labels', []))
    result = dict(
        t=t, pops=pops, states=labels, method='manybody', backend=backend,
        pops_atom=pops_atom, S=S, corr=corr, loss=loss, norms=norms,
        basis_probabilities=joint,
        basis_states=(out.get('basis_states') if projected else None),
        basis=dict(kind='projected' if projected else 'full', bit_order='site0-msb',
                   n_atoms=n, dim=dim, full_dim=1 << n),
        error_report=dict(backend=backend, solver='chebyshev', dim=dim, n_atoms=n, n_t=n_t,
                          hermitian=hermitian, used_sectors=False, full_pops=not capped,
                          projected=projected, entropy_cut=cut if want_entropy else None,
                          entropy_note=S_note, max_norm_dev=norm_dev),
        diag=({'dim': dim, 'n_atoms': n, 'backend': backend, 'used_sectors': False,
               'hermitian': hermitian, 'terms': sorted(use),
               'n_pairs': len(calib.get('pairs') or []),
               'dropped_amp': calib.get('dropped_amp'),
               'entropy_cut': cut if want_entropy else None, 'full_pops': not capped}
              if diag else None),
        observables=obs)
    if record_psi:
        result['psi_amp'] = out['psi']
    if loops is not None:
        result['loops'] = [list(loop) for loop in loops]
    # data descriptors as time_result attaches them (results.py:273-287,
    # engine_ed.py:153): the loss is a physical loss under gamma, else the
    # numerical norm drift
    from .results import describe_result
    _stage_progress(progress_arr, 2, 1.0)
    return describe_result(result, calib, loss_kind='norm_drift' if hermitian else 'physical_loss')
