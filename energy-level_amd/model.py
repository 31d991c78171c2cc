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


def _coupling_triplet(value):
    """(V, J, A) from a coupling kernel's return value: a dict, a (V, J, A)
    tuple, or a scalar (pure exchange J), as calib_build.py:32-41."""
    if isinstance(value, dict):
        return (float(value.get('V') or 0.0), float(value.get('J') or 0.0),
                float(value.get('A') or 0.0))
    if isinstance(value, (tuple, list)):
        if len(value) == 3:
            return tuple(float(x) for x in value)
        raise ValueError('a coupling tuple is (V, J, A), got %r' % (value,))
    return (0.0, float(value), 0.0)


def from_positions(positions, coupling, *, terms=('I2',), cutoff_um=None, periods_um=None,
                   images=0, source='calib_build.from_positions'):
    """Point cloud (N, 3) µm + coupling kernel → calib dict, as
    calib_build.from_positions (calib_build.py:121-166): pairs i < j with
    V, J, A summed over the periodic images (ix·Lx, iy·Ly), |ix|, |iy| <=
    ``images`` when ``periods_um`` = (Lx, Ly) is given, images beyond
    ``cutoff_um`` dropped (and counted), zero-distance self images skipped."""
    pts = np.asarray(positions, float)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError('positions must be an (N, 3) point cloud in µm, got %r' % (pts.shape,))
    if pts.shape[0] == 0:
        raise ValueError('positions is empty: a calib needs at least one atom')
    shifts = [(0.0, 0.0)]
    if periods_um is not None:
        periods = np.asarray(periods_um, float)
        if periods.shape != (2,) or np.any(periods <= 0):
            raise ValueError('periods_um must be two positive numbers (Lx, Ly)')
        k = int(images)
        if k < 0:
            raise ValueError('images must be non-negative')
        shifts = [(ix * periods[0], iy * periods[1]) for ix in range(-k, k + 1)
                  for iy in range(-k, k + 1)]
    cutoff = None if cutoff_um is None else float(cutoff_um)
    n = pts.shape[0]
    pairs, dropped = [], 0
    for i in range(n):
        for j in range(i + 1, n):
            acc = np.zeros(3)
            touched = False
            for sx, sy in shifts:
                r = float(np.linalg.norm(pts[j] - pts[i] + np.array([sx, sy, 0.0])))
                if r <= 0.0:
                    continue
                if cutoff is not None and r > cutoff:
                    dropped += 1
                    continue
                acc += np.asarray(_coupling_triplet(coupling(r)), float)
                touched = True
            if touched:
                pairs.append({'i': i, 'j': j, 'V': float(acc[0]), 'J': float(acc[1]),
                              'A': float(acc[2])})
    return {'atoms': [{'id': i, 'pos': [float(x) for x in p]} for i, p in enumerate(pts)],
            'pairs': pairs, 'terms': list(terms), 'source': str(source),
            'dropped': [], 'dropped_amp': {},
            'notes': {'n_dropped_by_cutoff': dropped,
                      'periods_um': None if periods_um is None else list(map(float, periods_um)),
                      'images': int(0 if periods_um is None else images)}}
