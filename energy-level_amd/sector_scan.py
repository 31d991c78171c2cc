"""Ground state per charge sector, shaped like the reference's
``scan_ground_sectors`` (tomisnot/energy-level
src/core/manybody/sector_scan.py:61-132): for k = 0..N (or the given sectors)
the number-sector basis is built (``number_sector_basis``) and solved by the
HIP Lanczos on that code list (closed-form combinatorial ranks on the
device); the lowest sector wins.  Number conservation is checked first
(Ω = 0, no pair creation), as there.
"""
from __future__ import annotations

import time

import numpy as np

from .basis import check_number_conserving, number_sector_basis
from .lanczos import hip_lanczos_ground_state

__all__ = ['scan_ground_sectors_hip']


def scan_ground_sectors_hip(calib, delta=None, Om=None, gamma=None, terms=None, *, sectors=None,
                            m=48, tol=None, dtype='float64', device='cuda',
                            memory_budget_mb=None, seed=0, progress=None):
    """→ {E0, sector, per_sector [{sector, dim, E0, gap, residual, converged,
    backend, seconds}], total_seconds, warmup_seconds, sectors_run, converged,
    dimension_sum, full_dimension} (keys of sector_scan.py:121-132)."""
    atoms = (calib or {}).get('atoms') or []
    n = len(atoms)
    if n <= 0:
        raise ValueError('calib needs atoms')
    delta = np.zeros(n) if delta is None else np.asarray(delta, float)
    Om = np.zeros(n, complex) if Om is None else np.asarray(Om, complex)
    gamma = np.zeros(n) if gamma is None else np.asarray(gamma, float)
    if np.any(gamma):
        raise ValueError('the sector scan needs a Hermitian Hamiltonian (gamma = 0)')
    check_number_conserving(calib, Om, terms)
    ks = list(range(n + 1)) if sectors is None else [int(k) for k in sectors]
    for k in ks:
        if not 0 <= k <= n:
            raise ValueError('sector must lie in [0, N]')
    import torch
    tw = time.perf_counter()
    torch.zeros(1, device=device)
    torch.cuda.synchronize()
    warmup = time.perf_counter() - tw
    t_start = time.perf_counter()
    rows, best = [], None
    for k in ks:
        t0 = time.perf_counter()
        idx = number_sector_basis(n, k)
        r = hip_lanczos_ground_state(calib, Om, delta, gamma, terms, basis_indices=idx, m=m,
                                     seed=seed, dtype=dtype, device=device, tol=tol,
                                     memory_budget_mb=memory_budget_mb)
        res = np.asarray(r.get('residuals', [0.0])).reshape(-1)
        row = dict(sector=k, dim=len(idx), E0=float(r['E0']),
                   gap=float(r.get('gap', float('nan'))),
                   residual=float(res[0]) if res.size else 0.0,
                   converged=bool(r.get('converged', True)), backend=str(r.get('backend', '')),
                   seconds=time.perf_counter() - t0)
        rows.append(row)
        if best is None or row['E0'] < best['E0']:
            best = dict(row)
        if progress is not None:
            progress(row)
    return dict(E0=best['E0'], sector=best['sector'], per_sector=rows,
                total_seconds=time.perf_counter() - t_start, warmup_seconds=warmup,
                sectors_run=ks, converged=all(r['converged'] for r in rows),
                dimension_sum=int(sum(r['dim'] for r in rows)), full_dimension=int(1 << n))
