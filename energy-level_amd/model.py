"""Host-side model description: calib + per-site drive → arrays for the C ABI.

Follows the reference's channel semantics (tomisnot/energy-level
src/core/manybody/channels.py:91-132, couplings.py:26-40): terms default to
the ones whose V/J/A values are non-zero; the diagonal carries +delta_i n_i;
the model is real exactly when Ω, J and A are real and Γ = 0
(channels.py:41-56).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ['RydModel', 'active_terms', 'from_calib', 'chain_positions', 'square_positions',
           'from_positions', 'c6_coupling']


def active_terms(calib):
    """``calib['terms']`` when given, else I1/I2/I3 by non-zero V/J/A
    (reference couplings.py:26-40)."""
    given = (calib or {}).get('terms')
    if given is not None:
        return list(given)
    pairs = (calib or {}).get('pairs') or []
    out = []
    for name, key in (('I1', 'V'), ('I2', 'J'), ('I3', 'A')):
        if any(abs(complex(p.get(key) or 0.0)) > 0.0 for p in pairs):
            out.append(name)
    return out


@dataclass
class RydModel:
    """Model on n sites: Ω_i (complex), delta_i, and per pair (i, j) the
    density interaction V, exchange J and pair creation A (complex; zero
    when the term is not active)."""
    n: int
    Om: np.ndarray
    delta: np.ndarray
    pair_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pair_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pair_v: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    pair_J: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex128))
    pair_A: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex128))

    def __post_init__(self):
        k = len(self.pair_i)
        if len(self.pair_J) != k:
            self.pair_J = np.zeros(k, np.complex128)
        if len(self.pair_A) != k:
            self.pair_A = np.zeros(k, np.complex128)

    @property
    def has_pair_flips(self):
        return bool(np.any(self.pair_J != 0) or np.any(self.pair_A != 0))

    @property
    def is_real(self):
        return not (np.any(np.asarray(self.Om).imag != 0) or np.any(self.pair_J.imag != 0)
                    or np.any(self.pair_A.imag != 0))

    def pairs(self):
        return [{'i': int(i), 'j': int(j), 'V': float(v), 'J': complex(J), 'A': complex(A)}
                for i, j, v, J, A in zip(self.pair_i, self.pair_j, self.pair_v, self.pair_J,
                                         self.pair_A)]


def from_calib(calib, Om=None, delta=None, gamma=None, terms=None):
    """``(calib, Om, delta, gamma, terms)`` → :class:`RydModel`.

    Γ≠0 is refused like the reference GPU path (matrix_free.py:142-143).
    Active terms (``terms`` or ``active_terms(calib)``) select V (I1),
    J (I2) and A (I3) per pair, as channel_spec (channels.py:91-132).
    """
    atoms = (calib or {}).get('atoms') or []
    n = len(atoms)
    if n == 0:
        raise ValueError('calib needs at least one atom')
    Om = np.zeros(n, complex) if Om is None else np.asarray(Om, complex).reshape(-1)
    delta = np.zeros(n) if delta is None else np.asarray(delta, float).reshape(-1)
    gamma = np.zeros(n) if gamma is None else np.asarray(gamma, float).reshape(-1)
    if any(x.shape != (n,) or not np.isfinite(x).all() for x in (Om, delta, gamma)):
        raise ValueError('Om/delta/gamma must be finite arrays of length N')
    if np.any(gamma):
        raise ValueError('the HIP engine needs a Hermitian Hamiltonian (gamma=0)')
    use = set(active_terms(calib) if terms is None else terms)
    pairs = list(calib.get('pairs') or [])
    pi, pj, pv, pJ, pA = [], [], [], [], []
    for p in pairs:
        v = float(np.real(p.get('V', 0.0) or 0.0)) if 'I1' in use else 0.0
        J = complex(p.get('J', 0.0) or 0.0) if 'I2' in use else 0j
        A = complex(p.get('A', 0.0) or 0.0) if 'I3' in use else 0j
        if v == 0.0 and J == 0 and A == 0:
            continue
        i, j = int(p['i']), int(p['j'])
        if not (0 <= i < n and 0 <= j < n) or i == j:
            raise ValueError('pair (%d, %d) out of range' % (i, j))
        if not (np.isfinite(v) and np.isfinite(J) and np.isfinite(A)):
            raise ValueError('pair couplings must be finite')
        pi.append(i)
        pj.append(j)
        pv.append(v)
        pJ.append(J)
        pA.append(A)
    return RydModel(n, Om, delta, np.asarray(pi, np.int32), np.asarray(pj, np.int32),
                    np.asarray(pv, np.float64), np.asarray(pJ, np.complex128),
                    np.asarray(pA, np.complex128))


# ---------------------------------------------------------------------------
# geometry helpers (reference src/core/manybody/model.py:101-120 and
# calib_build.py:121-166, restated for the synthetic workloads)
# ---------------------------------------------------------------------------
def chain_positions(n, a=1.0, jitter=0.0, seed=0):
    """Site i at (i·a, 0, 0); optional Gaussian jitter on all three coordinates."""
    pos = np.zeros((n, 3))
    pos[:, 0] = np.arange(n) * a
    if jitter:
        pos += np.random.default_rng(seed).normal(0.0, jitter, size=pos.shape)
    return pos


def square_positions(n, a=1.0):
    """Row-major square array: site i at ((i % side)·a, (i // side)·a, 0)."""
    side = int(np.ceil(np.sqrt(n)))
    return np.array([[(i % side) * a, (i // side) * a, 0.0] for i in range(n)])


def c6_coupling(c6):
    """coupling(r) = {'V': C6 / r^6}."""
    return lambda r: {'V': c6 / r ** 6, 'J': 0.0, 'A': 0.0}


def from_positions(positions, coupling, terms=('I1',)):
    """Point cloud + coupling kernel → calib dict with pairs i<j."""
    pts = np.asarray(positions, float)
    pairs = []
    for i in range(len(pts)):
        for j in range(i + 1, len(pts)):
            r = float(np.linalg.norm(pts[j] - pts[i]))
            if r <= 0.0:
                continue
            c = coupling(r)
            pairs.append({'i': i, 'j': j, 'V': float(c.get('V', 0.0)),
                          'J': float(c.get('J', 0.0)), 'A': float(c.get('A', 0.0))})
    return {'atoms': [{'id': i, 'pos': [float(x) for x in p]} for i, p in enumerate(pts)],
            'pairs': pairs, 'terms': list(terms)}
