"""CPU oracle for the Rydberg H·v path — TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The HIP library never calls into it and has no CPU fallback through it.

It restates, in plain numpy/scipy, the reference algorithms of
``tomisnot/energy-level`` that the GPU path replaces.  Every function cites
the reference file:line it follows (paths relative to the reference root).

Pinning (what keeps each function honest, besides retyping the formula):

* channel semantics / H·v / CSR  — checked against an independent textbook
  ``np.kron`` construction (``kron_hamiltonian``), the hand-written N=2 4×4
  matrix, the XX dimer spectrum and the blockade 3-level block
  (reference ``checks/equibench_engine_ed_selfcheck.py:56-79, 196-264``,
  ``checks/manybody_m1_ed_selfcheck.py:56-75``);
* bases — brute force over 2^N, Fibonacci counts F(N+2)
  (``checks/equibench_engine_ed_selfcheck.py:226-238``), Zeckendorf closed form;
* KPM — exact-moment path vs ``diag(linspace(-2,3,32))`` known answer
  (``checks/equibench_gpu_backends_selfcheck.py:27-32``), Jackson closed form,
  Chebyshev identities;
* Lanczos / ground state — dense ``eigh``;
* propagation — Rabi closed form ``sin²(CONV·Ω·t/2)``
  (``checks/manybody_m1_ed_selfcheck.py:43-54``) and ``scipy.linalg.expm``.

Conventions (reference ``src/core/manybody/channels.py:12-13``): site 0 is the
most significant bit, bit=1 is |r>, H[t, s] = <t|H|s>, diagonal +delta_i n_i.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

CONV = 2.0 * np.pi * 1e-3        # reference src/core/evolve_util.py:12


# ---------------------------------------------------------------------------
# channel semantics  (reference src/core/manybody/channels.py, couplings.py)
# ---------------------------------------------------------------------------
def active_terms(pairs):
    """V≠0 → I1, J≠0 → I2, A≠0 → I3 (reference couplings.py:26-40)."""
    out = []
    for name, key in (('I1', 'V'), ('I2', 'J'), ('I3', 'A')):
        if any(abs(complex(p.get(key) or 0.0)) > 0.0 for p in pairs):
            out.append(name)
    return out


def channel_list(n, Om, pairs, terms):
    """Flip channels ``(selector, xor_mask, amp)`` (reference channels.py:91-132).

    selector is a tuple of (site, required bit of the *source* state).
    Ω: g→r carries Ω/2, r→g carries conj(Ω)/2; J: r_i g_j → g_i r_j carries
    conj(J), reverse J; A: g g → r r carries A, reverse conj(A).
    """
    Om = np.asarray(Om, complex)
    use = set(terms)
    ch = []
    for i in range(n):
        if Om[i] == 0:
            continue
        m = 1 << (n - 1 - i)
        ch.append((((i, 0),), m, Om[i] / 2.0))
        ch.append((((i, 1),), m, np.conj(Om[i]) / 2.0))
    for p in pairs:
        i, j = int(p['i']), int(p['j'])
        m = (1 << (n - 1 - i)) ^ (1 << (n - 1 - j))
        J = complex(p.get('J', 0.0) or 0.0)
        A = complex(p.get('A', 0.0) or 0.0)
        if 'I2' in use and J != 0:
            ch.append((((i, 1), (j, 0)), m, np.conj(J)))
            ch.append((((i, 0), (j, 1)), m, J))
        if 'I3' in use and A != 0:
            ch.append((((i, 0), (j, 0)), m, A))
            ch.append((((i, 1), (j, 1)), m, np.conj(A)))
    return ch


def model_dtype(Om, pairs, terms, gamma=None):
    """float64 unless Γ≠0 or any channel amplitude is complex (channels.py:41-56)."""
    if gamma is not None and np.any(np.asarray(gamma) != 0):
        return np.dtype(np.complex128)
    for _sel, _m, amp in channel_list(len(Om), Om, pairs, terms):
        if complex(amp).imag != 0.0:
            return np.dtype(np.complex128)
    return np.dtype(np.float64)


def bits_of(n, codes):
    """bit(s, i) = (s >> (n-1-i)) & 1 (reference basis.py:6-9)."""
    codes = np.asarray(codes, np.int64)
    return np.asarray([(codes >> (n - 1 - i)) & 1 for i in range(n)], np.int8)


def diagonal(n, delta, pairs, terms, codes, gamma=None, dtype=np.float64):
    """diag = Σ δ_i b_i + Σ_pairs V b_i b_j (− iΓ/2 term) (channels.py:58-77)."""
    b = bits_of(n, codes)
    d = np.zeros(len(codes), dtype)
    delta = np.asarray(delta, float)
    for i in range(n):
        if np.issubdtype(np.dtype(dtype), np.complexfloating) and gamma is not None:
            d += (delta[i] - 0.5j * float(gamma[i]) * 1e-3 / CONV) * b[i]
        else:
            d += delta[i] * b[i]
    if 'I1' in set(terms):
        for p in pairs:
            d += float(np.real(p.get('V', 0.0) or 0.0)) * (b[int(p['i'])] * b[int(p['j'])])
    return d


def channel_matvec(n, Om, delta, pairs, terms, v, codes=None):
    """Per-channel gather/scatter H·v (reference channels.py:135-221, matvec :210-221).

    ``codes`` (optional) is the basis in any order; targets that leave the
    basis are dropped, as in ``ChannelKernel._indices`` (:177-192).
    """
    full = codes is None
    codes = np.arange(1 << n, dtype=np.int64) if full else np.asarray(codes, np.int64)
    v = np.asarray(v)
    cdt = np.result_type(v.dtype, model_dtype(Om, pairs, terms), np.float64)
    y = diagonal(n, delta, pairs, terms, codes, dtype=cdt) * v
    b = bits_of(n, codes)
    pos = None
    if not full:
        order = np.argsort(codes)
        sorted_codes = codes[order]
    for sel, mask, amp in channel_list(n, Om, pairs, terms):
        ok = np.ones(len(codes), bool)
        for site, bit in sel:
            ok &= b[site] == bit
        cols = np.flatnonzero(ok)
        tgt = codes[cols] ^ mask
        if full:
            loc = tgt
        else:
            k = np.searchsorted(sorted_codes, tgt)
            k = np.minimum(k, len(codes) - 1)
            hit = sorted_codes[k] == tgt
            cols, loc = cols[hit], order[k[hit]]
        a = amp if np.iscomplexobj(y) else float(np.real(amp))
        y[loc] += a * v[cols]
    return y


def build_H_sparse(n, Om, delta, pairs, terms, codes=None, gamma=None):
    """Two-pass CSR assembly (reference assemble.py:277-393).

    Pass 1 counts one diagonal entry plus one entry per valid channel hit per
    row; pass 2 fills through a per-row cursor, diagonal first; indices sorted.
    """
    full = codes is None
    s = np.arange(1 << n, dtype=np.int64) if full else np.asarray(codes, np.int64)
    dim = len(s)
    dt = model_dtype(Om, pairs, terms, gamma)
    d = diagonal(n, delta, pairs, terms, s, gamma=gamma, dtype=dt)
    b = bits_of(n, s)
    if not full:
        perm = np.argsort(s)
        ordered = s[perm]

    def rowmap(t):
        if full:
            return t, np.ones(t.size, bool)
        k = np.searchsorted(ordered, t)
        ok = k < dim
        ok[ok] &= ordered[k[ok]] == t[ok]
        rows = np.zeros(t.size, np.int64)
        rows[ok] = perm[k[ok]]
        return rows, ok

    chans = channel_list(n, Om, pairs, terms)
    hits = []
    counts = np.ones(dim, np.int64)
    for sel, mask, amp in chans:
        ok = np.ones(dim, bool)
        for site, bit in sel:
            ok &= b[site] == bit
        cols = np.flatnonzero(ok)
        rows, valid = rowmap(s[cols] ^ mask)
        hits.append((rows[valid], cols[valid], amp))
        counts += np.bincount(rows[valid], minlength=dim)
    indptr = np.zeros(dim + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, np.int64)
    data = np.zeros(nnz, dt)
    cur = indptr[:-1].copy()
    rr = np.arange(dim)
    data[cur] = d
    indices[cur] = rr
    cur += 1
    for rows, cols, amp in hits:
        pos = cur[rows]
        data[pos] += amp if np.issubdtype(dt, np.complexfloating) else float(np.real(amp))
        indices[pos] = cols
        cur[rows] += 1
    H = sp.csr_matrix((data, indices, indptr), shape=(dim, dim))
    H.sort_indices()
    return H


# ---------------------------------------------------------------------------
# independent textbook construction (reference checks/equibench_engine_ed_selfcheck.py:56-79)
# ---------------------------------------------------------------------------
_I2 = np.eye(2, dtype=complex)
_NOP = np.array([[0, 0], [0, 1]], complex)
_SX = np.array([[0, 1], [1, 0]], complex)
_SP = np.array([[0, 0], [1, 0]], complex)      # |r><g|
_SM = _SP.conj().T


def _embed(n, ops):
    out = np.array([[1.0 + 0j]])
    for i in range(n):
        out = np.kron(out, ops.get(i, _I2))
    return out


def kron_hamiltonian(n, Om, delta, pairs, terms):
    """Dense H from Kronecker products, site 0 leftmost (independent of channels)."""
    H = np.zeros((1 << n, 1 << n), complex)
    Om = np.asarray(Om, complex)
    for i in range(n):
        if delta[i] != 0:
            H += delta[i] * _embed(n, {i: _NOP})
        if Om[i] != 0:
            # Ω σ+ /2 + conj(Ω) σ- /2 ; equals Ω/2 σx for real Ω
            H += (Om[i] / 2.0) * _embed(n, {i: _SP}) + (np.conj(Om[i]) / 2.0) * _embed(n, {i: _SM})
    use = set(terms)
    for p in pairs:
        i, j = int(p['i']), int(p['j'])
        V = float(np.real(p.get('V', 0.0) or 0.0))
        J = complex(p.get('J', 0.0) or 0.0)
        A = complex(p.get('A', 0.0) or 0.0)
        if 'I1' in use and V != 0:
            H += V * _embed(n, {i: _NOP, j: _NOP})
        if 'I2' in use and J != 0:
            H += J * _embed(n, {i: _SP, j: _SM}) + np.conj(J) * _embed(n, {i: _SM, j: _SP})
        if 'I3' in use and A != 0:
            H += A * _embed(n, {i: _SP, j: _SP}) + np.conj(A) * _embed(n, {i: _SM, j: _SM})
    return H


# ---------------------------------------------------------------------------
# bases  (reference basis.py:24-76, matrix_free.py:12-38)
# ---------------------------------------------------------------------------
def number_sector_basis(n, k):
    """Sorted int64 codes with popcount k (reference matrix_free.py:12-38)."""
    n, k = int(n), int(k)
    if not 0 <= n <= 63 or not 0 <= k <= n:
        raise ValueError('charge_sector out of range')
    codes = np.zeros(1, np.int64)
    cnt = np.zeros(1, np.int16)
    for i in range(n):
        nc = codes | np.int64(1 << i)
        ncnt = cnt + 1
        keep = ncnt <= k
        codes = np.concatenate((codes, nc[keep]))
        cnt = np.concatenate((cnt, ncnt[keep]))
    out = codes[cnt == k]
    out.sort()
    return out


def independent_states(n, edges, max_states=1 << 20):
    """Independent sets in ascending code order (reference basis.py:24-47).

    Depth-first from site 0 (the MSB), "empty" before "occupied"; because
    site 0 is the most significant bit this yields ascending integers.
    """
    n = int(n)
    blocked = [0] * n
    for i, j in sorted({tuple(sorted(map(int, e))) for e in edges}):
        blocked[i] |= 1 << j
        blocked[j] |= 1 << i
    out = []
    stack = [(0, 0, 0)]
    while stack:
        site, occ, code = stack.pop()
        if site == n:
            if len(out) >= max_states:
                raise ValueError('restricted basis exceeds max_states=%d' % max_states)
            out.append(code)
            continue
        # push "occupied" first so that "empty" is expanded first
        if not blocked[site] & occ:
            stack.append((site + 1, occ | (1 << site), code | (1 << (n - 1 - site))))
        stack.append((site + 1, occ, code))
    return out


def independent_states_bruteforce(n, edges):
    """All codes in [0, 2^n) with no edge fully occupied (brute force)."""
    s = np.arange(1 << n, dtype=np.int64)
    ok = np.ones(s.size, bool)
    for i, j in edges:
        ok &= ~((((s >> (n - 1 - i)) & 1) == 1) & (((s >> (n - 1 - j)) & 1) == 1))
    return s[ok]


def fibonacci(k):
    """F(1) = F(2) = 1."""
    a, b = 0, 1
    for _ in range(int(k)):
        a, b = b, a + b
    return a


def chain_fib_rank(n, codes):
    """Rank of a chain-blockade code: Σ_{b_i=1} F(n-i+1) (Zeckendorf)."""
    codes = np.asarray(codes, np.int64)
    r = np.zeros(codes.shape, np.int64)
    for i in range(n):
        r += ((codes >> (n - 1 - i)) & 1) * fibonacci(n - i + 1)
    return r


def chain_fib_unrank(n, ranks):
    """Inverse of chain_fib_rank by the greedy Zeckendorf decomposition."""
    ranks = np.asarray(ranks, np.int64).copy()
    codes = np.zeros(ranks.shape, np.int64)
    prev = np.zeros(ranks.shape, bool)
    for i in range(n):
        f = fibonacci(n - i + 1)
        take = (ranks >= f) & ~prev
        codes |= take.astype(np.int64) << (n - 1 - i)
        ranks -= np.where(take, f, 0)
        prev = take
    return codes


# ---------------------------------------------------------------------------
# KPM  (reference src/core/manybody/kpm.py)
# ---------------------------------------------------------------------------
def kpm_bounds(H, margin=1e-3):
    """emin/emax (eigvalsh for D<=16, else eigsh SA/LA from default_rng(0)) and
    c=(emin+emax)/2, r=(emax-emin)/2·(1+margin)+1e-12 (reference kpm.py:82-97)."""
    from scipy.sparse.linalg import eigsh
    D = H.shape[0]
    if D <= 16:
        w = np.linalg.eigvalsh(H.toarray() if sp.issparse(H) else np.asarray(H))
        emin, emax = float(w[0]), float(w[-1])
    else:
        Hc = H if sp.issparse(H) else sp.csr_matrix(H)
        v0 = np.random.default_rng(0).standard_normal(D)
        emin = float(eigsh(Hc, k=1, which='SA', v0=v0.copy(), return_eigenvectors=False)[0])
        emax = float(eigsh(Hc, k=1, which='LA', v0=v0.copy(), return_eigenvectors=False)[0])
    c = 0.5 * (emin + emax)
    r = 0.5 * (emax - emin) * (1.0 + margin) + 1e-12
    return emin, emax, c, r


def jackson_kernel(num):
    """g_n = [(M-n+1)cos(πn/(M+1)) + sin(πn/(M+1))cot(π/(M+1))]/(M+1) (kpm.py:109-112)."""
    n = np.arange(num + 1)
    a = np.pi / (num + 1)
    return ((num - n + 1) * np.cos(a * n) + np.sin(a * n) / np.tan(a)) / (num + 1)


def rademacher_probes(D, nvec, seed):
    """The reference's probe stream: default_rng(seed).choice([-1,1], D) per probe
    drawn sequentially (reference kpm.py:137-140)."""
    rng = np.random.default_rng(int(seed))
    for _ in range(int(nvec)):
        yield rng.choice(np.array([-1.0, 1.0]), size=int(D))


def kpm_moments(H, num, nvec, seed=0, exact=False, c=None, r=None, margin=1e-3, probes=None):
    """μ_n = (1/D)Tr T_n(H̃) by Hutchinson Z2 probes (reference kpm.py:115-151).

    ``c``/``r`` may be given explicitly (the eigsh bounds are the one part
    that is iteration-path dependent); ``probes`` (nvec arrays of ±1) replaces
    the CPU probe stream, e.g. by the stream of the GPU path (kpm.py:60-63)."""
    D = H.shape[0]
    if c is None or r is None:
        _, _, c, r = kpm_bounds(H, margin)
    mu = np.zeros(num + 1)
    if exact:
        Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
        lam = np.linalg.eigvalsh((Hd - c * np.eye(D)) / r)
        t0, t1 = np.ones_like(lam), lam.copy()
        mu[0] = t0.mean()
        if num >= 1:
            mu[1] = t1.mean()
        for k in range(2, num + 1):
            t0, t1 = t1, 2 * lam * t1 - t0
            mu[k] = t1.mean()
        return mu, c, r
    Hc = H if sp.issparse(H) else sp.csr_matrix(H)
    Ht = (Hc - c * sp.eye(D, format='csr')) * (1.0 / r)
    acc = np.zeros(num + 1)
    for z in (rademacher_probes(D, nvec, seed) if probes is None else probes):
        z = np.asarray(z, float)
        tp, tc = z.copy(), Ht @ z
        acc[0] += (tp @ z).real
        if num >= 1:
            acc[1] += (tc @ z).real
        for k in range(2, num + 1):
            tp, tc = tc, 2 * (Ht @ tc) - tp
            acc[k] += (tc @ z).real
    return acc / (nvec * D), c, r


def cheb_coeffs(func, num):
    """Chebyshev interpolation coefficients a_0..a_num (reference kpm.py:100-106)."""
    from numpy.polynomial.chebyshev import chebinterpolate
    a = np.asarray(chebinterpolate(func, num), float)
    if a.size < num + 1:
        a = np.concatenate([a, np.zeros(num + 1 - a.size)])
    return a[:num + 1]


def thermal_from_moments(D, betas, mu, c, r, num):
    """E, C, F, Z, lnZ per β with the lower-bound shift (reference kpm.py:159-179)."""
    betas = np.atleast_1d(np.asarray(betas, float))
    out = {k: np.zeros_like(betas) for k in ('E', 'C', 'F', 'Z', 'lnZ')}
    for i, b in enumerate(betas):
        shift = c - r
        z = float(np.dot(cheb_coeffs(lambda x: np.exp(-b * (r * x + c - shift)), num), mu)) * D
        e1 = float(np.dot(cheb_coeffs(lambda x: (r * x + c) * np.exp(-b * (r * x + c - shift)), num), mu)) * D
        e2 = float(np.dot(cheb_coeffs(lambda x: (r * x + c) ** 2 * np.exp(-b * (r * x + c - shift)), num), mu)) * D
        if not np.isfinite(z) or z <= 0:
            raise ValueError('KPM partition function is not positive')
        out['E'][i] = e1 / z
        out['C'][i] = b * b * max(e2 / z - out['E'][i] ** 2, 0.0)
        out['F'][i] = shift - np.log(z) / b if b > 0 else np.nan
        out['lnZ'][i] = np.log(z) - b * shift
        with np.errstate(over='ignore', under='ignore'):
            out['Z'][i] = np.exp(out['lnZ'][i])
    out['betas'] = betas
    return out


def dos_from_moments(mu, c, r, energies, num, use_jackson=True):
    """Jackson-damped DOS series, clipped at 0 (reference kpm.py:182-189)."""
    x = (np.asarray(energies, float) - c) / r
    x = np.clip(x, -1 + 1e-9, 1 - 1e-9)
    g = jackson_kernel(num) if use_jackson else np.ones(num + 1)
    s = mu[0] * np.ones_like(x)
    if num >= 1:
        t0, t1 = np.ones_like(x), x.copy()
        s += 2 * g[1] * mu[1] * t1
        for k in range(2, num + 1):
            t0, t1 = t1, 2 * x * t1 - t0
            s += 2 * g[k] * mu[k] * t1
    return np.maximum(s / (np.pi * np.sqrt(1 - x * x)) / r, 0.0)


# ---------------------------------------------------------------------------
# Lanczos / ground state  (reference spectrum.py:34-103, equilibrium.py:35-96)
# ---------------------------------------------------------------------------
def lanczos(matvec, phi, k):
    """Lanczos with two full reorthogonalisation sweeps (reference spectrum.py:34-103)."""
    v = np.asarray(phi, complex).reshape(-1)
    dim = v.size
    nrm = float(np.linalg.norm(v))
    k = min(int(k), dim)
    q, qp = v / nrm, np.zeros_like(v)
    alphas, betas, basis = [], [], []
    beta = 0.0
    for j in range(k):
        basis.append(q)
        w = np.asarray(matvec(q), complex)
        a = float(np.vdot(q, w).real)
        w = w - a * q - (beta * qp if j else 0.0)
        for _ in range(2):
            for bvec in basis:
                w = w - np.vdot(bvec, w) * bvec
        beta = float(np.linalg.norm(w))
        alphas.append(a)
        if beta <= 1e-12 or j + 1 >= dim:
            break
        if j + 1 < k:
            betas.append(beta)
        qp, q = q, w / beta
    return np.asarray(alphas), np.asarray(betas), nrm ** 2


def ground_state_dense(H):
    """Dense eigh ground state (reference equilibrium.py:65-75)."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
    w, V = np.linalg.eigh(Hd)
    return dict(E0=float(w[0]), gap=float(w[1] - w[0]) if len(w) > 1 else 0.0,
                vec=V[:, 0], evals=w)


def ground_state_eigsh(H, k=2, seed=0):
    """eigsh(which='SA', tol=0) from default_rng(seed) v0 (reference equilibrium.py:76-96)."""
    from scipy.sparse.linalg import eigsh
    dim = H.shape[0]
    rng = np.random.default_rng(seed)
    v0 = rng.standard_normal(dim)
    if np.issubdtype(H.dtype, np.complexfloating):
        v0 = v0 + 1j * rng.standard_normal(dim)
    w, V = eigsh(H, k=min(k, dim - 1), which='SA', v0=v0, maxiter=dim * 20, tol=0)
    o = np.argsort(w)
    w, V = w[o], V[:, o]
    return dict(E0=float(w[0]), gap=float(w[1] - w[0]) if len(w) > 1 else 0.0,
                vec=V[:, 0], evals=w)


def site_occupations(n, vec, codes=None):
    """<n_i> = Σ_s |ψ_s|² b_i(s) (reference eq_route.py:108-122)."""
    codes = np.arange(1 << n) if codes is None else np.asarray(codes)
    p = np.abs(np.asarray(vec)) ** 2
    return bits_of(n, codes).astype(float) @ p


def corr_matrix(n, vec, codes=None):
    """Connected occupation correlation C_ij = <n_i n_j> - <n_i><n_j>."""
    codes = np.arange(1 << n) if codes is None else np.asarray(codes)
    p = np.abs(np.asarray(vec)) ** 2
    b = bits_of(n, codes).astype(float)
    nn = (b * p) @ b.T
    ni = b @ p
    return nn - np.outer(ni, ni)


# ---------------------------------------------------------------------------
# piecewise-constant propagation  (reference engine_ed.py:130-147, runtime.py:56-114)
# ---------------------------------------------------------------------------
def segment_intervals(a, b, segments):
    """Split [a, b] at every interior segment start (reference runtime.py:56-60)."""
    cuts = sorted({float(s['t0']) for s in segments or [] if a < float(s['t0']) < b})
    edges = [float(a)] + cuts + [float(b)]
    return list(zip(edges[:-1], edges[1:]))


def segment_drive(t, segments, n, interp='left'):
    """Latest-t0 (left) or midpoint rule (reference runtime.py:90-114)."""
    ordered = sorted(segments, key=lambda s: float(s['t0']))
    idx = 0
    for k, s in enumerate(ordered):
        if float(s['t0']) <= t:
            idx = k

    def arrays(s):
        om = np.asarray(s.get('Om', np.zeros(n)), float).reshape(-1)
        ph = np.asarray(s.get('phase', np.zeros(n)), float).reshape(-1)
        de = np.asarray(s.get('delta', np.zeros(n)), float).reshape(-1)
        return om * np.exp(1j * ph), de
    om, de = arrays(ordered[idx])
    if interp == 'midpoint' and idx < len(ordered) - 1:
        om2, de2 = arrays(ordered[idx + 1])
        om, de = 0.5 * (om + om2), 0.5 * (de + de2)
    return om, de


def propagate_segments(n, pairs, terms, segments, psi0, t, codes=None, interp='left',
                       method='expm', gamma=None):
    """ψ(t_k) by exact exponentials per piece, exp(-i·CONV·H·Δt)
    (reference engine_ed.py:130-147).  ``method='expm_multiply'`` is the
    reference's own sparse stepping (engine_ed.py:145); ``'expm'`` uses a
    dense exponential (small N).  ``gamma`` adds the loss diagonal
    −(i/2)Γ_i n_i of build_H_sparse (non-Hermitian H)."""
    from scipy.sparse.linalg import expm_multiply
    psi = np.asarray(psi0, complex).copy()
    out = [psi.copy()]
    cache = {}
    for k in range(1, len(t)):
        for a, b in segment_intervals(t[k - 1], t[k], segments):
            om, de = segment_drive(0.5 * (a + b), segments, n, interp)
            key = (tuple(om), tuple(de))
            if key not in cache:
                H = build_H_sparse(n, om, de, pairs, terms, codes, gamma=gamma)
                cache[key] = H.toarray() if method == 'expm' else H.tocsr()
            if method == 'expm':
                psi = sla.expm(-1j * CONV * (b - a) * cache[key]) @ psi
            else:
                psi = expm_multiply((-1j * CONV * (b - a)) * cache[key], psi)
        out.append(psi.copy())
    return np.asarray(out)


# ---------------------------------------------------------------------------
# pair tables from positions  (reference calib_build.py:121-166)
# ---------------------------------------------------------------------------
def from_positions(positions, coupling):
    """Pairs i<j with coupling(r) → {'V','J','A'}; no periodic images."""
    pts = np.asarray(positions, float)
    if pts.ndim == 2 and pts.shape[1] == 2:
        pts = np.concatenate([pts, np.zeros((len(pts), 1))], axis=1)
    pairs = []
    for i in range(len(pts)):
        for j in range(i + 1, len(pts)):
            r = float(np.linalg.norm(pts[j] - pts[i]))
            if r <= 0:
                continue
            c = coupling(r)
            if isinstance(c, dict):
                V, J, A = c.get('V', 0.0), c.get('J', 0.0), c.get('A', 0.0)
            elif isinstance(c, (tuple, list)):
                V, J, A = c
            else:
                V, J, A = 0.0, c, 0.0
            pairs.append({'i': i, 'j': j, 'V': float(V), 'J': float(J), 'A': float(A)})
    return pairs


# ---------------------------------------------------------------------------
# Pauli strings  (reference observables.py:461-622, PauliPlan)
# ---------------------------------------------------------------------------
PAULI_1 = {'x': np.array([[0, 1], [1, 0]], complex),
           'y': np.array([[0, -1j], [1j, 0]], complex),      # rows/cols (g, r): <r|y|g> = +i
           'z': np.array([[1, 0], [0, -1]], complex)}        # |r> has -1


def pauli_expect(n, vec, terms, codes=None):
    """<v|P|v> for P = Π σ^{axis}_{site} on the full basis or a code list:
    the basis path of PauliPlan.expect (observables.py:564-622) — every
    source state s maps to t = s ^ (X|Y) with the source-bit phase
    (σz: −1 on |r>, σy: −i from |r>, +i from |g>); targets outside the basis
    contribute nothing."""
    v = np.asarray(vec, complex).reshape(-1)
    codes = np.arange(1 << n, dtype=np.int64) if codes is None else np.asarray(codes, np.int64)
    pos = {int(c): k for k, c in enumerate(codes)}
    total = 0j
    for k, c in enumerate(codes):
        c = int(c)
        t, ph = c, 1 + 0j
        for site, axis in terms:
            b = (c >> (n - 1 - site)) & 1
            if axis == 'z':
                ph *= -1 if b else 1
            elif axis == 'y':
                ph *= -1j if b else 1j
            if axis != 'z':
                t ^= 1 << (n - 1 - site)
        r = pos.get(t)
        if r is not None:
            total += np.conj(v[r]) * ph * v[k]
    return total


def pauli_dense(n, terms):
    """The same operator as a dense 2^n matrix by Kronecker products
    (site 0 = most significant factor) — an independent construction."""
    ops = [np.eye(2, dtype=complex)] * n
    ops = list(ops)
    for site, axis in terms:
        ops[site] = PAULI_1[axis]
    M = np.ones((1, 1), complex)
    for o in ops:
        M = np.kron(M, o)
    return M


# ---------------------------------------------------------------------------
# spin observables and entanglement  (reference observables.py:50-204)
# ---------------------------------------------------------------------------
def spin_observables(n, vec, codes=None):
    """sx, sy, sz, sp = <σ+>, sm = <σ->, xy_corr[i,j] = <σ+_i σ-_j> (diagonal
    <n_i>), phase = arg(<σ->) (NaN where |<σ->| < 1e-14); σ+ = |r><g|, on the
    full basis or a code list (observables.py:117-204)."""
    v = np.asarray(vec, complex).reshape(-1)
    codes = np.arange(1 << n, dtype=np.int64) if codes is None else np.asarray(codes, np.int64)
    pos = {int(c): k for k, c in enumerate(codes)}
    sz = np.zeros(n)
    sm = np.zeros(n, complex)
    xy = np.zeros((n, n), complex)
    for i in range(n):
        mi = 1 << (n - 1 - i)
        for k, c in enumerate(codes):
            c = int(c)
            if c & mi:
                sz[i] -= abs(v[k]) ** 2
                kg = pos.get(c ^ mi)
                if kg is not None:
                    sm[i] += np.conj(v[kg]) * v[k]
            else:
                sz[i] += abs(v[k]) ** 2
    for i in range(n):
        mi = 1 << (n - 1 - i)
        for j in range(n):
            mj = 1 << (n - 1 - j)
            for k, c in enumerate(codes):
                c = int(c)
                if i == j:
                    if c & mi:
                        xy[i, i] += abs(v[k]) ** 2
                elif not (c & mi) and (c & mj):
                    kd = pos.get((c | mi) ^ mj)
                    if kd is not None:
                        xy[i, j] += np.conj(v[kd]) * v[k]
    phase = np.angle(sm)
    phase[np.abs(sm) < 1e-14] = np.nan
    return dict(sx=2 * sm.real, sy=2 * sm.imag, sz=sz, sp=np.conj(sm), sm=sm, xy_corr=xy,
                phase=phase)


def sigma_pm_dense(n, i, j):
    """σ+_i σ-_j as a dense 2^n matrix by Kronecker products (site 0 most
    significant; σ+ = |r><g| has its 1 in row r, column g)."""
    sp = np.array([[0, 0], [1, 0]], complex)
    sm = np.array([[0, 1], [0, 0]], complex)
    ops = [np.eye(2, dtype=complex) for _ in range(n)]
    if i == j:
        ops[i] = sp @ sm
    else:
        ops[i] = sp
        ops[j] = sm
    M = np.ones((1, 1), complex)
    for o in ops:
        M = np.kron(M, o)
    return M


def entropy_at_cut(n, vec, cut, codes=None):
    """Von Neumann entropy (nat) of sites [0, cut) from the singular values of
    the Schmidt matrix; on a code list the matrix rows/columns are the
    occupied left/right parts (observables.py:50-115)."""
    v = np.asarray(vec, complex).reshape(-1)
    if n < 2:
        return 0.0
    if codes is None:
        w = np.linalg.svd(v.reshape(1 << cut, 1 << (n - cut)), compute_uv=False) ** 2
        w = w[w > 1e-15]
        return float(-np.sum(w * np.log(w))) if w.size else 0.0
    codes = np.asarray(codes, np.int64)
    _, li = np.unique(codes >> (n - cut), return_inverse=True)
    _, ri = np.unique(codes & ((1 << (n - cut)) - 1), return_inverse=True)
    M = np.zeros((li.max() + 1, ri.max() + 1), complex)
    M[li, ri] = v
    w = np.linalg.svd(M, compute_uv=False) ** 2
    w = w / w.sum()
    w = w[w > 0]
    return float(-np.sum(w * np.log(w)))
