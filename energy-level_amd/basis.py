"""Basis identities and the chain-blockade sector on the device.

Bit order follows the reference (tomisnot/energy-level
src/core/manybody/basis.py:6-9): site i is bit (n-1-i), bit 1 = |r>.  The
chain nearest-neighbour blockade sector is the reference's
``independent_states(n, [(i, i+1)])`` (basis.py:24-47): codes without two
adjacent excitations, ascending, F(n+2) of them; rank(code) = Σ_{site i
excited} F(n-i+1) — computed by ``ryd_sector_rank`` / ``ryd_sector_codes``.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat

__all__ = ['fibonacci', 'chain_blockade_dim', 'is_chain_blockade_basis', 'sector_codes',
           'sector_rank', 'chain_edges', 'number_sector_basis', 'check_number_conserving',
           'binomial_table', 'validate_edges', 'is_chain_edges', 'graph_plan',
           'independent_states_hip', 'independent_set_count_hip', 'blockade_edges',
           'build_plan_hip']


def fibonacci(k):
    """F(1) = F(2) = 1."""
    a, b = 0, 1
    for _ in range(int(k)):
        a, b = b, a + b
    return a


def chain_blockade_dim(n):
    """Dimension F(n+2) of the chain blockade sector."""
    return fibonacci(int(n) + 2)


def chain_edges(n):
    return [(i, i + 1) for i in range(int(n) - 1)]


def is_chain_blockade_basis(n, codes):
    """True when ``codes`` is exactly the chain blockade sector in ascending order."""
    c = np.asarray(codes, np.int64).reshape(-1)
    n = int(n)
    if c.size != chain_blockade_dim(n):
        return False
    if c.size and (c[0] < 0 or c[-1] >= (1 << n)):
        return False
    if np.any(np.diff(c) <= 0):
        return False
    return not np.any(c & (c >> 1))


def sector_codes(plan, first=0, count=None, ranks=None):
    """Codes (device int64) of ranks ``first .. first+count-1`` or of ``ranks``."""
    import torch
    if ranks is not None:
        ranks = torch.as_tensor(ranks, dtype=torch.int64, device=plan.device).contiguous()
        count = ranks.numel()
    elif count is None:
        count = plan.dim - int(first)
    out = torch.empty(int(count), dtype=torch.int64, device=plan.device)
    nat.check(nat.lib().ryd_sector_codes(plan.handle, nat.ptr(ranks), int(first), int(count),
                                         nat.ptr(out), nat.stream_ptr()))
    return out


def sector_rank(plan, codes):
    """Sector rank of each code (device int64), -1 outside the sector."""
    import torch
    codes = torch.as_tensor(codes, dtype=torch.int64, device=plan.device).contiguous()
    out = torch.empty(codes.numel(), dtype=torch.int64, device=plan.device)
    nat.check(nat.lib().ryd_sector_rank(plan.handle, nat.ptr(codes), int(codes.numel()),
                                        nat.ptr(out), nat.stream_ptr()))
    return out


# ---------------------------------------------------------------------------
# number sectors (reference src/core/manybody/matrix_free.py:12-48)
# ---------------------------------------------------------------------------
def binomial_table(n):
    """C[a, b] = binomial(a, b) for 0 <= a, b <= n (int64)."""
    C = np.zeros((n + 1, n + 1), np.int64)
    for a in range(n + 1):
        C[a, 0] = 1
        for b in range(1, a + 1):
            C[a, b] = C[a - 1, b - 1] + (C[a - 1, b] if b <= a - 1 else 0)
    return C


def number_sector_basis(n, number):
    """Ascending int64 codes of n bits with ``number`` bits set.

    Codes come from unranking r = 0 .. C(n,k)-1 in the combinatorial number
    system (r = Σ_i C(p_i, i) over the set bit positions p_1 < … < p_k,
    i = 1..k), whose order is ascending code order — the order of the
    reference's ``number_sector_basis``.  Errors as there: a non-integer
    number or one outside [0, n], n outside [0, 63], raise ValueError."""
    if isinstance(number, (bool, np.bool_)) or int(number) != number:
        raise ValueError('the charge sector must be an integer')
    n, k = int(n), int(number)
    if not 0 <= n <= 63 or not 0 <= k <= n:
        raise ValueError('charge sector out of range or n outside [0, 63]')
    C = binomial_table(n)
    total = int(C[n, k])
    r = np.arange(total, dtype=np.int64)
    codes = np.zeros(total, np.int64)
    for i in range(k, 0, -1):
        col = C[:, i]                          # C(p, i), non-decreasing in p for p >= i-1
        # largest p with C(p, i) <= r
        p = np.searchsorted(col, r, side='right') - 1
        codes |= np.left_shift(np.int64(1), p.astype(np.int64))
        r -= col[p]
    return codes


def check_number_conserving(calib, Om=None, terms=None):
    """ValueError when Ω or active pair creation (I3) would leave a
    fixed-number sector (reference matrix_free.py:41-48)."""
    use = set((calib or {}).get('terms', ()) if terms is None else terms)
    if Om is not None and np.any(np.asarray(Om, complex) != 0):
        raise ValueError('a fixed charge sector needs Omega = 0')
    if 'I3' in use and any(complex(p.get('A', 0) or 0) != 0 for p in (calib or {}).get('pairs', ())):
        raise ValueError('a fixed charge sector does not support I3 pair creation')


# ---------------------------------------------------------------------------
# blockade graphs (reference src/core/manybody/basis.py:12-76 and
# hamiltonian_plan.py:15-72); enumeration and ranks on the device
# ---------------------------------------------------------------------------
def validate_edges(n, edges):
    """Sorted unique (i, j) with i < j; ValueError for a non-integer index, an
    index outside [0, n) or a self loop (as basis.py:12-21)."""
    out = set()
    for edge in edges or []:
        if len(edge) != 2 or any(isinstance(i, (bool, np.bool_)) or int(i) != i for i in edge):
            raise ValueError('a blockade edge must be two integer site indices')
        i, j = int(edge[0]), int(edge[1])
        if not 0 <= i < n or not 0 <= j < n or i == j:
            raise ValueError('blockade edge out of range or a self loop: %r' % (tuple(edge),))
        out.add((min(i, j), max(i, j)))
    return sorted(out)


def is_chain_edges(n, edges):
    """True when the validated edges are exactly the chain (i, i+1)."""
    return list(edges) == chain_edges(n)


def graph_plan(n, edges, model=None, device='cuda'):
    """A ``Plan`` on the independent sets of the blockade graph (the device
    enumerates them; ``plan.dim`` is their number)."""
    from .model import RydModel
    from .plan import Plan
    n = int(n)
    e = validate_edges(n, edges)
    if model is None:
        model = RydModel(n=n, Om=np.zeros(n, complex), delta=np.zeros(n))
    return Plan(model, basis='graph', edges=e, device=device)


def independent_states_hip(n, edges, max_states=None, device='cuda'):
    """Ascending int64 codes (a device tensor) of the independent sets of the
    graph: the list of ``independent_states`` (basis.py:24-47) without its
    2^20 default cap.  ``max_states``: ValueError when the basis is larger
    (the reference's explicit output-size limit).  n in [0, 60]."""
    import torch
    n = int(n)
    if not 0 <= n <= 60:
        raise ValueError('blockade-graph bases need 0 <= N <= 60')
    if n == 0:
        validate_edges(0, edges)
        return torch.zeros(1, dtype=torch.int64, device=device)
    P = graph_plan(n, edges, device=device)
    try:
        if max_states is not None and P.dim > int(max_states):
            raise ValueError('restricted basis has %d states, more than max_states=%d'
                             % (P.dim, int(max_states)))
        return sector_codes(P)
    finally:
        P.close()


def independent_set_count_hip(n, edges, device='cuda'):
    """Number of independent sets (``independent_set_count``, basis.py:50-76),
    counted by the device enumeration (no node budget)."""
    n = int(n)
    if n == 0:
        validate_edges(0, edges)
        return 1
    P = graph_plan(n, edges, device=device)
    try:
        return int(P.dim)
    finally:
        P.close()


def _constraint_dict(constraint):
    """Normalised constraint (hamiltonian_plan.py:15-33)."""
    if constraint is None:
        return dict(kind='none', mode='full', threshold_mhz=0.0, edges=None)
    if isinstance(constraint, str):
        constraint = dict(kind=constraint)
    if not isinstance(constraint, dict):
        raise ValueError('manybody.constraint must be a dict or a string')
    kind = str(constraint.get('kind', 'none')).lower()
    mode = str(constraint.get('mode', 'project' if kind in ('pxp', 'blockade') else 'full')).lower()
    if kind not in ('none', 'blockade', 'pxp'):
        raise ValueError('constraint.kind must be none/blockade/pxp')
    if mode not in ('full', 'project'):
        raise ValueError('constraint.mode must be full/project')
    threshold = float(constraint.get('threshold_mhz', 0.0) or 0.0)
    if not np.isfinite(threshold) or threshold < 0:
        raise ValueError('constraint.threshold_mhz must be finite and non-negative')
    return dict(kind=kind, mode=mode, threshold_mhz=threshold, edges=constraint.get('edges'))


def blockade_edges(calib, constraint):
    """Blockade edges of a projected constraint (hamiltonian_plan.py:36-50):
    explicit edges, or pairs with |V| >= threshold_mhz."""
    c = _constraint_dict(constraint)
    n = len(calib.get('atoms') or [])
    explicit = c.get('edges')
    if explicit is not None and not isinstance(explicit, str):
        return validate_edges(n, explicit)
    if explicit not in (None, 'auto'):
        raise ValueError('constraint.edges must be auto or a list of integer edges')
    threshold = float(c.get('threshold_mhz', 0.0))
    if threshold <= 0:
        raise ValueError('blockade/PXP projection needs threshold_mhz > 0 or explicit edges')
    out = set()
    for p in calib.get('pairs') or []:
        if abs(complex(p.get('V', 0.0))) >= threshold:
            i, j = int(p['i']), int(p['j'])
            out.add((min(i, j), max(i, j)))
    return sorted(out)


def build_plan_hip(calib, Om=None, delta=None, gamma=None, terms=None, constraint=None,
                   max_states=None, device='cuda'):
    """``build_plan`` (hamiltonian_plan.py:53-72) with the projected basis
    enumerated on the device: the same keys, ``states`` a device int64 tensor
    (ascending, the reference's order), no 2^20 default cap
    (``max_states`` still limits when given)."""
    from .model import active_terms
    atoms = calib.get('atoms') or []
    n = len(atoms)
    Om = np.zeros(n, complex) if Om is None else np.asarray(Om, complex)
    delta = np.zeros(n) if delta is None else np.asarray(delta, float)
    gamma = np.zeros(n) if gamma is None else np.asarray(gamma, float)
    if any(x.shape != (n,) or not np.isfinite(x).all() for x in (Om, delta, gamma)):
        raise ValueError('Om/delta/gamma must be finite arrays with one entry per atom')
    c = _constraint_dict(constraint)
    projected = c['kind'] in ('blockade', 'pxp') and c['mode'] == 'project'
    edges = blockade_edges(calib, c) if projected else []
    states = independent_states_hip(n, edges, max_states, device=device) if projected else None
    return dict(version=1, kind='HamiltonianPlan', constraint=c, n_atoms=n, full_dim=1 << n,
                dim=(int(states.numel()) if states is not None else 1 << n),
                projected=projected, blockade_edges=[list(e) for e in edges], states=states,
                Om=Om.tolist(), delta=delta.tolist(), gamma=gamma.tolist(),
                terms=list(active_terms(calib) if terms is None else terms), calib=calib)
