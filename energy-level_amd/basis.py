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
           'sector_rank', 'chain_edges']


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
