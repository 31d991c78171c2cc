"""Matrix-free KPM Chebyshev moments on the HIP engine.

Restates the reference's Hutchinson estimator (tomisnot/energy-level
src/core/manybody/kpm.py:115-151): mu_n = (1/D) Tr T_n(H~), H~ = (H − c)/r,
probes drawn exactly like the reference CPU path,
``default_rng(seed).choice([-1., 1.], size=D)`` sequentially per probe
(kpm.py:137-140), so equal seeds give equal probes.  The probe is shipped
bit-packed (D/8 bytes) and expanded on the device.

Two moment estimators over the same recurrence (``ryd_cheb_moments``):

* ``doubling=False`` — the reference's direct form acc[n] += <z, t_n>
  (kpm.py:143-148), ``num`` Chebyshev steps per probe;
* ``doubling=True`` (default) — mu_{2k} = 2<t_k,t_k> − mu_0,
  mu_{2k+1} = 2<t_{k+1},t_k> − mu_1, ceil(num/2) steps per probe (the same
  trace estimator: z^T T_{2k} z = 2|T_k z|^2 − z^T z for Hermitian H~).

Bounds: explicit (c, r), or emin/emax from a device Lanczos run with
r = (emax − emin)/2·(1 + margin) + 1e-12 (reference margin 1e-3, kpm.py:96).
Post-processing (Jackson kernel, thermal, DOS) follows kpm.py:100-205.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .lanczos import lanczos_run, _start_vector
from .model import from_calib
from .plan import Plan

__all__ = ['kpm_moments_hip', 'spectral_bounds', 'jackson_kernel', 'cheb_coeffs',
           'thermal_from_moments', 'dos_from_moments', 'probe_bits', 'device_probes',
           'kpm_moments_gpu_hip']


def probe_bits(D, nvec, seed):
    """Yield the reference probe stream bit-packed as uint32 words (bit set = +1)."""
    rng = np.random.default_rng(int(seed))
    words = (int(D) + 31) // 32
    for _ in range(int(nvec)):
        z = rng.choice(np.array([-1.0, 1.0]), size=int(D))
        b = np.packbits(z > 0, bitorder='little')
        pad = np.zeros(words * 4, np.uint8)
        pad[:b.size] = b
        yield pad.view(np.uint32)


def spectral_bounds(plan, m=120, seed=0, tol=1e-8, margin=1e-3):
    """(emin, emax, c, r) from one FP64 device Lanczos run in which both the
    lowest and the highest Ritz value pass the bound test (as the
    reference's eigsh SA and LA runs, kpm.py:92-93)."""
    import torch
    v0 = _start_vector(plan.dim, seed, not plan.is_real, plan.device)
    R = lanczos_run(plan, v0, m, 8, tol, want=1, store=False, top=True)   # no Ritz vectors needed
    # Ritz values lie inside the spectrum; widen each end by its residual
    # bound |beta_last · y[-1]| so that [emin, emax] covers the spectrum
    lo, hi = int(np.argmin(R.theta)), int(np.argmax(R.theta))
    blast = 0.0 if R.stopped else float(R.nrm[R.j].item())
    emin = float(R.theta[lo]) - abs(blast * R.Y[-1, lo])
    emax = float(R.theta[hi]) + abs(blast * R.Y[-1, hi])
    c = 0.5 * (emin + emax)
    r = 0.5 * (emax - emin) * (1.0 + margin) + 1e-12
    return emin, emax, c, r


def kpm_moments_gpu_hip(H, num=300, nvec=48, seed=0, device='cuda', chunk=32,
                        memory_budget_mb=None, dtype='float64'):
    """kpm_moments_gpu (tomisnot/energy-level src/core/manybody/kpm.py:26-75)
    with its signature: H is an assembled Hermitian matrix (scipy sparse or
    dense), copied once into a device CSR plan; → (mu[num+1], c, r).  The
    probes are that function's stream: blocks of min(chunk, nvec) columns of
    torch.randint(0, 2, (D, b)) from a device generator seeded with ``seed``
    (:60-63), 1 → +1 and 0 → −1; c and r use its 1e-3 bound margin (:82-97),
    here from the device Lanczos."""
    import torch
    from .plan import CsrPlan
    nat.require_gpu()
    if str(device).startswith('cpu'):
        raise RuntimeError('HIP device unavailable (device=%s)' % device)
    if dtype not in ('float32', 'float64'):
        raise ValueError('dtype must be float32/float64')
    if int(num) < 1 or int(nvec) < 1:
        raise ValueError('num and nvec must be positive integers')
    P = CsrPlan(H, device=device)
    try:
        D, nvec = P.dim, int(nvec)
        words = (D + 31) // 32
        if memory_budget_mb is not None:
            # the reference's rule and accounting (kpm.py:41-47 with
            # resources.budget_bytes / require_workspace): refuse (ValueError)
            # only when the operator and one probe do not fit, else shrink
            # the probe block to the budget.  The block size also fixes the
            # torch.randint stream, so the same budget draws the same probes.
            # This path holds less than that estimate: the CSR operator
            # (nnz·(4 + 8 or 16) + 8·(D + 1) bytes), three recurrence
            # vectors, the packed probes (4 bytes per 32 entries), a block's
            # int8 draw and 4·D + D/4 bytes while one column is packed.
            budget = float(memory_budget_mb)
            if not np.isfinite(budget) or budget <= 0:
                raise ValueError('memory_budget_mb must be finite and positive')
            budget = int(budget * 1024 ** 2)
            itemsize = 8 if dtype == 'float32' else 16
            base = (P.nnz + D) * (itemsize + 16) * 3
            one = base + D * itemsize * 6
            if one > budget:
                raise ValueError('GPU KPM operator and one probe need at least %.2f MiB, over '
                                 'the budget of %.2f MiB' % (one / 1024 ** 2, budget / 1024 ** 2))
            chunk = min(int(chunk), max(1, (budget - base) // (D * itemsize * 6)))
        gen = torch.Generator(device=P.device)
        gen.manual_seed(int(seed))
        chunk = max(1, min(int(chunk), nvec))
        # bit k of word i = entry 32·i + k (two's-complement int32 words)
        shifts = torch.arange(32, device=P.device, dtype=torch.int32)
        probes, done = [], 0
        while done < nvec:
            b = min(chunk, nvec - done)
            z = torch.randint(0, 2, (D, b), device=P.device, generator=gen, dtype=torch.int8)
            for j in range(b):
                bits = torch.zeros(words * 32, dtype=torch.int32, device=P.device)
                bits[:D] = z[:, j]
                # the int64 sum of the int32 terms (bit 31 weighs -2^31) is the
                # word's signed value, exactly representable in int32
                w = torch.bitwise_left_shift(bits.view(words, 32), shifts).sum(dim=1)
                probes.append(w.to(torch.int32).contiguous())
                del bits, w
            done += b
            del z
        _, _, c, r = spectral_bounds(P, margin=1e-3)
        return kpm_moments_hip(plan=P, num=num, nvec=nvec, seed=seed, dtype=dtype, c=c, r=r,
                               probes=probes)
    finally:
        P.close()


def device_probes(plan, nvec, seed):
    """The reference probe stream, bit-packed and uploaded (list of device tensors)."""
    import torch
    return [torch.from_numpy(b).to(plan.device) for b in probe_bits(plan.dim, nvec, seed)]


def kpm_moments_hip(calib=None, Om=None, delta=None, terms=None, *, plan=None, num=300, nvec=48,
                    seed=0, dtype='float64', c=None, r=None, margin=1e-3, doubling=True,
                    device='cuda', return_sums=False, probes=None):
    """→ (mu[num+1], c, r) matrix-free on the GPU (no CSR is assembled).

    ``probes``: pre-uploaded probes from :func:`device_probes` (else drawn here)."""
    import torch
    nat.require_gpu()
    if dtype not in ('float32', 'float64'):
        raise ValueError('dtype must be float32/float64')
    if int(num) < 1 or int(nvec) < 1:
        raise ValueError('num and nvec must be positive integers')
    P = plan if plan is not None else Plan(from_calib(calib, Om, delta, None, terms), device=device)
    if c is None or r is None:
        _, _, c, r = spectral_bounds(P, seed=0, margin=margin)
    D = P.dim
    complex_kind = not P.is_real
    if complex_kind:
        tdt = torch.complex64 if dtype == 'float32' else torch.complex128
    else:
        tdt = torch.float32 if dtype == 'float32' else torch.float64
    num = int(num)
    K = (num + 1) // 2 if doubling else num
    K = max(K, 1)
    sums = torch.zeros((K + 1) * 4, dtype=torch.float64, device=P.device)
    bufs = [torch.empty(D, dtype=tdt, device=P.device) for _ in range(3)]
    L = nat.lib()
    code = nat.dtype_code(tdt)
    ws = P.workspace()
    if probes is None:
        probes = (torch.from_numpy(b).to(P.device) for b in probe_bits(D, nvec, seed))
    for pb in probes:
        nat.check(L.ryd_cheb_moments(P.handle, code, nat.ptr(pb), nat.ptr(bufs[0]),
                                     nat.ptr(bufs[1]), nat.ptr(bufs[2]), K, float(c), float(r),
                                     nat.ptr(sums), nat.ptr(ws), nat.stream_ptr()))
    S = sums.cpu().numpy().reshape(K + 1, 4) / (float(nvec) * D)
    mu = np.zeros(num + 1)
    mu[0] = 1.0
    mu[1] = S[1, 3]
    if doubling:
        for k in range(1, K + 1):
            if 2 * k <= num:
                mu[2 * k] = 2.0 * S[k, 1] - mu[0]
            if 2 * k - 1 <= num and k >= 1:
                # <t_k, t_{k-1}> sits in row k, column 0
                mu[2 * k - 1] = 2.0 * S[k, 0] - mu[1] if k >= 2 else mu[1]
    else:
        for k in range(1, num + 1):
            mu[k] = S[k, 3]
    # with every eigenvalue of H~ in [-1, 1], |z^T T_n(H~) z| <= |z|^2 = D for
    # each ±1 probe, so |mu_n| <= 1; an eigenvalue outside makes T_n grow
    # exponentially, which this catches before the moments are used
    if not np.all(np.isfinite(mu)) or np.max(np.abs(mu)) > 1.0 + 1e-3:
        raise RuntimeError('KPM moments exceed 1 in magnitude (max %.6g): the bounds '
                           '[c - r, c + r] do not enclose the spectrum'
                           % float(np.max(np.abs(np.nan_to_num(mu, nan=np.inf)))))
    if return_sums:
        return mu, c, r, S
    return mu, c, r


# ---------------------------------------------------------------------------
# post-processing (reference kpm.py:100-205)
# ---------------------------------------------------------------------------
def jackson_kernel(num):
    """Jackson damping g_0..g_num (reference kpm.py:109-112)."""
    n = np.arange(num + 1)
    a = np.pi / (num + 1)
    return ((num - n + 1) * np.cos(a * n) + np.sin(a * n) / np.tan(a)) / (num + 1)


def cheb_coeffs(func, num):
    """f(x) = Σ a_n T_n(x) coefficients by Chebyshev interpolation (kpm.py:100-106)."""
    from numpy.polynomial.chebyshev import chebinterpolate
    a = np.asarray(chebinterpolate(func, num), float)
    if a.size < num + 1:
        a = np.concatenate([a, np.zeros(num + 1 - a.size)])
    return a[:num + 1]


def thermal_from_moments(D, betas, mu, c, r, num):
    """E, C, F, Z, lnZ per β, shifted by the lower bound c − r (kpm.py:159-179)."""
    betas = np.atleast_1d(np.asarray(betas, float))
    E, C, F, Z, lnZ = (np.zeros_like(betas) for _ in range(5))
    for i, b in enumerate(betas):
        shift = c - r
        w = lambda x: np.exp(-b * (r * x + c - shift))
        z = float(np.dot(cheb_coeffs(w, num), mu)) * D
        e1 = float(np.dot(cheb_coeffs(lambda x: (r * x + c) * w(x), num), mu)) * D
        e2 = float(np.dot(cheb_coeffs(lambda x: (r * x + c) ** 2 * w(x), num), mu)) * D
        if not np.isfinite(z) or z <= 0:
            raise ValueError('KPM partition function is not positive; raise moments/probes or lower beta')
        E[i] = e1 / z
        C[i] = b * b * max(e2 / z - E[i] ** 2, 0.0)
        F[i] = shift - np.log(z) / b if b > 0 else np.nan
        lnZ[i] = np.log(z) - b * shift
        with np.errstate(over='ignore', under='ignore'):
            Z[i] = np.exp(lnZ[i])
    return dict(betas=betas, E=E, C=C, F=F, Z=Z, lnZ=lnZ)


def dos_from_moments(mu, c, r, energies, num, use_jackson=True):
    """ρ(E) normalised to ∫ρ dE = 1, clipped at 0 (kpm.py:182-189)."""
    x = np.clip((np.asarray(energies, float) - c) / r, -1 + 1e-9, 1 - 1e-9)
    g = jackson_kernel(num) if use_jackson else np.ones(num + 1)
    s = mu[0] * np.ones_like(x)
    if num >= 1:
        t0, t1 = np.ones_like(x), x.copy()
        s += 2 * g[1] * mu[1] * t1
        for k in range(2, num + 1):
            t0, t1 = t1, 2 * x * t1 - t0
            s += 2 * g[k] * mu[k] * t1
    return np.maximum(s / (np.pi * np.sqrt(1 - x * x)) / r, 0.0)
