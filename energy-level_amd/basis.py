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
           'binomial_table']


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
