"""Observables of a device state on any plan's basis.

Mirrors tomisnot/energy-level src/core/manybody/observables.py:
``corr_matrix`` (:40-47) as ``corr_matrix_hip``, ``normalize_pauli_strings``
(:315-377), ``PauliPlan`` (:486-622, the full and basis paths) as
``PauliPlanHip``, ``measure_pauli_strings`` (:624-661, 1-D states) as
``measure_pauli_strings_hip`` and the loop post-processing of :206-300
(``loop_observables``, ``link_loop_observables``) on the host, over the
link correlations the device measured.  The sums run on the device
(``ryd_pauli_expect``, ``ryd_site_correlations``); no 2^n bit matrix or
position table is built.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat

__all__ = ['normalize_pauli_strings', 'PauliPlanHip', 'measure_pauli_strings_hip',
           'corr_matrix_hip', 'site_correlation_sums', 'loop_observables',
           'link_loop_observables']

# below this modulus a bond correlation has no phase (NaN), as observables.py:225
PHASE_FLOOR = 1e-14

_AXIS_ALIASES = {'x': 'x', 'y': 'y', 'z': 'z', 'sx': 'x', 'sy': 'y', 'sz': 'z',
                 'σx': 'x', 'σy': 'y', 'σz': 'z', 'sigmax': 'x', 'sigmay': 'y', 'sigmaz': 'z'}


def _axis_name(raw):
    axis = _AXIS_ALIASES.get(str(raw).strip().lower())
    if axis is None:
        raise ValueError('a Pauli axis must be x/y/z (or σx etc.), got %r' % (raw,))
    return axis


def normalize_pauli_strings(spec, n_atoms=None, err=None):
    """``[{'label', 'terms'}]`` from a list of term lists or dicts; the rules
    of observables.py:315-377 (non-empty, integer sites in range, no site
    twice, axes x/y/z; label defaults to 'x@0 y@1 …')."""
    if spec is None:
        return None
    err = err or ValueError
    if not isinstance(spec, (list, tuple)):
        raise err('pauli_strings must be a list (got %s)' % type(spec).__name__)
    n = None if n_atoms is None else int(n_atoms)
    out = []
    for k, item in enumerate(spec):
        tag = 'pauli_strings[%d]' % k
        if isinstance(item, dict):
            terms_raw = item.get('terms', item.get('sites'))
            label = item.get('label')
        else:
            terms_raw, label = item, None
        if terms_raw is None:
            raise err('%s has no terms' % tag)
        if not isinstance(terms_raw, (list, tuple)) or not terms_raw:
            raise err('%s.terms must be a non-empty sequence' % tag)
        terms, seen = [], set()
        for t in terms_raw:
            if not isinstance(t, (list, tuple)) or len(t) != 2:
                raise err('%s: every term is [site, axis], got %r' % (tag, t))
            site = t[0]
            if isinstance(site, bool) or not isinstance(site, (int, np.integer)):
                raise err('%s: a site must be an integer, got %r' % (tag, site))
            site = int(site)
            if n is not None and not 0 <= site < n:
                raise err('%s: site %d out of range (n = %d)' % (tag, site, n))
            if site in seen:
                raise err('%s: site %d appears twice' % (tag, site))
            seen.add(site)
            terms.append((site, _axis_name(t[1])))
        if label is None:
            label = ' '.join('%s@%d' % (ax, s) for s, ax in terms)
        out.append({'label': str(label), 'terms': terms})
    return out


def _masks(n, specs):
    """(flip = X|Y, sign = Y|Z, ny) per string, bit position of site i = n-1-i."""
    flip = np.zeros(len(specs), np.uint64)
    sign = np.zeros(len(specs), np.uint64)
    ny = np.zeros(len(specs), np.int32)
    for q, s in enumerate(specs):
        seen = set()
        for site, axis in s['terms']:
            site = int(site)
            if not 0 <= site < n:
                raise ValueError('Pauli string site %r out of range (n=%d)' % (site, n))
            if site in seen:
                raise ValueError('Pauli string has two operators on site %d' % site)
            seen.add(site)
            if axis not in ('x', 'y', 'z'):
                raise ValueError('Pauli axis must be x/y/z, got %r' % (axis,))
            bit = np.uint64(1) << np.uint64(n - 1 - site)
            if axis in ('x', 'y'):
                flip[q] |= bit
            if axis in ('y', 'z'):
                sign[q] |= bit
            if axis == 'y':
                ny[q] += 1
    return flip, sign, ny


class PauliPlanHip:
    """Pauli strings compiled to bit masks once, evaluated on the device at
    every call (``PauliPlan`` of observables.py:486-622 on a ``Plan``'s basis:
    full, chain sector, code list or blockade graph)."""

    def __init__(self, plan, specs):
        self.plan = plan
        self.specs = [dict(s) for s in (specs or [])]
        self.flip, self.sign, self.ny = _masks(plan.n, self.specs)

    def expect(self, x):
        """``list[complex]`` in the order of ``specs``; an empty string is <x|x>."""
        import torch
        P = self.plan
        code = P.check_vec(x)
        if not self.specs:
            return []
        out = torch.zeros(2 * len(self.specs), dtype=torch.float64, device=P.device)
        nat.check(nat.lib().ryd_pauli_expect(
            P.handle, code, nat.ptr(x), len(self.specs), self.flip.ctypes.data,
            self.sign.ctypes.data, self.ny.ctypes.data, nat.ptr(out), nat.ptr(P.workspace()),
            nat.stream_ptr()))
        o = out.cpu().numpy()
        return [complex(o[2 * q], o[2 * q + 1]) for q in range(len(self.specs))]


def measure_pauli_strings_hip(x, specs, plan):
    """``measure_pauli_strings`` for a device vector on ``plan``'s basis."""
    specs = list(specs or [])
    if not specs:
        return []
    return PauliPlanHip(plan, specs).expect(x)


def site_correlation_sums(plan, x):
    """(S, norm): S[i, j] = Σ_s |x_s|^2 b_i(s) b_j(s), norm = Σ_s |x_s|^2."""
    import torch
    n = plan.n
    code = plan.check_vec(x)
    width = n * (n + 1) // 2 + 1
    scratch = torch.empty(width * 513, dtype=torch.float64, device=plan.device)
    out = torch.empty(n * n + 1, dtype=torch.float64, device=plan.device)
    nat.check(nat.lib().ryd_site_correlations(plan.handle, code, nat.ptr(x), nat.ptr(out),
                                              nat.ptr(scratch), scratch.numel() * 8,
                                              nat.stream_ptr()))
    o = out.cpu().numpy()
    return o[:n * n].reshape(n, n), float(o[n * n])


def corr_matrix_hip(plan, x):
    """C_ij = <n_i n_j> − <n_i><n_j> with the weights |x_s|^2 as they are
    (corr_matrix, observables.py:40-47: a state of norm below one, as under
    loss, is not renormalised); also returns <n_i>."""
    S, norm = site_correlation_sums(plan, x)
    if not norm > 0:
        raise ValueError('zero state')
    ni = np.diag(S).copy()
    return S - np.outer(ni, ni), ni


def _loop_sites(raw, n=None):
    q = [int(x) for x in raw]
    if len(q) < 2 or (n is not None and any(not 0 <= x < n for x in q)):
        raise ValueError('a loop needs at least two valid atom indices, got %r' % (list(raw),))
    return q


def loop_observables(xy_corr, loops):
    """(phases, windings) of closed loops over a bond matrix
    ``xy_corr[a, b] = <σ+_a σ−_b>`` (observables.py:206-229).

    A loop lists atoms without repeating the first; its closing bond is
    implied.  The phase is the sum of arg(xy_corr[a, b]) over its bonds (not
    wrapped), the winding that sum over 2π; a bond of modulus at most
    ``PHASE_FLOOR`` makes both NaN."""
    if loops is None:
        return [], []
    xy = np.asarray(xy_corr, complex)
    phases, windings = [], []
    for raw in loops:
        q = _loop_sites(raw, xy.shape[0])
        z = xy[q, q[1:] + q[:1]]
        if np.all(np.abs(z) > PHASE_FLOOR):
            total = float(np.sum(np.angle(z)))
            phases.append(total)
            windings.append(total / (2.0 * np.pi))
        else:
            phases.append(np.nan)
            windings.append(np.nan)
    return phases, windings


def link_loop_observables(link_corr, link_pairs, loops):
    """(phase, winding) arrays of shape (times, loops) from a series of
    measured links (observables.py:232-279).

    ``link_corr[t, q] = <σ+_a σ−_b>`` for ``link_pairs[q] = (a, b)``; a bond
    walked as (b, a) uses the conjugate, the first listed pair wins, and a
    loop with a bond that no pair covers stays NaN.  At a time where any of
    its bonds is at most ``PHASE_FLOOR`` in modulus the loop is NaN."""
    if loops is None:
        return np.empty((0, 0), float), np.empty((0, 0), float)
    where = {}
    for q, (a, b) in enumerate((int(a), int(b)) for a, b in (link_pairs or [])):
        where.setdefault((a, b), (q, False))
        where.setdefault((b, a), (q, True))
    arr = np.asarray(link_corr, complex)
    if arr.ndim == 1:
        arr = arr[None, :]
    phase = np.full((arr.shape[0], len(loops)), np.nan)
    winding = np.full_like(phase, np.nan)
    for ell, raw in enumerate(loops):
        q = _loop_sites(raw)
        bonds = [where.get(e) for e in zip(q, q[1:] + q[:1])]
        if any(b is None for b in bonds):
            continue
        cols = arr[:, [c for c, _ in bonds]]
        flip = np.array([f for _, f in bonds])
        cols = np.where(flip[None, :], np.conj(cols), cols)
        ok = np.all(np.abs(cols) > PHASE_FLOOR, axis=1)
        total = np.angle(cols).sum(axis=1)
        phase[ok, ell] = total[ok]
        winding[ok, ell] = total[ok] / (2.0 * np.pi)
    return phase, winding
