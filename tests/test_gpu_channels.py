"""Generic channel kernel: J/A channels on the full basis, code bases
(number sectors with closed-form ranks, blockade graphs by binary search),
against the oracle's channel matvec and dense diagonalisation."""
import json
import os

import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _pairs(n, rng, J=True, A=True, cplx=True):
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            p = {'i': i, 'j': j, 'V': float(rng.uniform(0, 2) / (j - i) ** 3)}
            if J:
                p['J'] = complex(rng.normal(), rng.normal() if cplx else 0.0) / (j - i)
            if A:
                p['A'] = complex(rng.normal(), rng.normal() if cplx else 0.0) / (j - i) ** 2
            out.append(p)
    return out


def _plan(pkg, dev, n, Om, delta, pairs, terms, codes=None):
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': terms}
    return pkg.Plan(pkg.from_calib(calib, Om, delta, None, terms),
                    basis='codes' if codes is not None else 'full', codes=codes, device=dev)


@pytest.mark.parametrize('n', [2, 5, 9, 12, 15])
def test_full_basis_with_exchange_and_pair_creation(pkg, dev, n):
    """From n = 11/12 the full basis runs the tiled channel kernel (2^11
    complex128 and 2^12 complex64 codes per tile): n = 15 has pairs with both
    bits in the tile, one bit above it and both above it."""
    import torch
    rng = np.random.default_rng(n)
    terms = ['I1', 'I2', 'I3']
    pairs = _pairs(n, rng)
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.uniform(-2, 2, n)
    P = _plan(pkg, dev, n, Om, delta, pairs, terms)
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    ref = orc.channel_matvec(n, Om, delta, pairs, terms, x)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y64 = P.hv(torch.as_tensor(x.astype(np.complex64), device=dev)).cpu().numpy()
    assert np.max(np.abs(y64 - ref)) <= 2e-6 * np.max(np.abs(ref))
    with pytest.raises(ValueError):                     # complex H refuses real vectors
        P.hv(torch.as_tensor(x.real, device=dev))


@pytest.mark.parametrize('n', [10, 17])
def test_real_exchange_model_real_vectors(pkg, dev, n):
    import torch
    rng = np.random.default_rng(3)
    pairs = _pairs(n, rng, A=False, cplx=False)
    P = _plan(pkg, dev, n, rng.uniform(0, 1, n), rng.uniform(-1, 1, n), pairs, ['I1', 'I2'])
    x = rng.normal(size=1 << n)
    ref = orc.channel_matvec(n, P.model.Om, P.model.delta, pairs, ['I1', 'I2'], x).real
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        y = P.hv(torch.as_tensor(x, device=dev).to(dt)).double().cpu().numpy()
        assert np.max(np.abs(y - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize('n,k,A', [(10, 4, False), (12, 6, False), (9, 0, False), (9, 9, False),
                                   (11, 5, True)])
def test_number_sector_closed_form_ranks(pkg, dev, n, k, A):
    """With A pairs the pair-creation targets leave the sector and drop out."""
    import torch
    rng = np.random.default_rng(100 + n + k)
    terms = ['I1', 'I2', 'I3'] if A else ['I1', 'I2']
    pairs = _pairs(n, rng, A=A)
    codes = pkg.number_sector_basis(n, k)
    delta = rng.uniform(-2, 2, n)
    P = _plan(pkg, dev, n, np.zeros(n), delta, pairs, terms, codes=codes)
    assert P.dim == len(codes)
    x = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, np.zeros(n), delta, pairs, terms, x, codes)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    # ranks and codes round trip; codes outside the sector rank -1
    got = pkg.sector_codes(P).cpu().numpy()
    assert np.array_equal(got, codes)
    probe = np.concatenate([codes, [0 if k else 1, (1 << n) - 1 if k != n else 0]])
    r = pkg.sector_rank(P, torch.as_tensor(probe, device=dev)).cpu().numpy()
    assert np.array_equal(r[:len(codes)], np.arange(len(codes)))
    assert np.all(r[len(codes):] == -1)


def test_blockade_graph_binary_search_ranks(pkg, dev):
    """4x4 square array with nearest-neighbour blockade: independent sets."""
    import torch
    side, n = 4, 16
    edges = []
    for i in range(n):
        if (i % side) + 1 < side:
            edges.append((i, i + 1))
        if i + side < n:
            edges.append((i, i + side))
    codes = np.asarray(orc.independent_states(n, edges), np.int64)
    rng = np.random.default_rng(7)
    pairs = [{'i': i, 'j': j, 'V': float(rng.uniform(0.1, 1))} for i in range(n)
             for j in range(i + 1, n) if (i, j) not in edges]
    Om = rng.uniform(0.5, 1.5, n)
    delta = rng.uniform(-1, 1, n)
    P = _plan(pkg, dev, n, Om, delta, pairs, ['I1'], codes=codes)
    x = rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, Om, delta, pairs, ['I1'], x, codes).real
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    # occupations over the code basis
    occ = pkg.site_occupations(P, torch.as_tensor(x, device=dev))
    assert np.allclose(occ, orc.site_occupations(n, x, codes) / np.sum(x * x), atol=1e-12)


def test_unsorted_basis_indices_ground_state(pkg, dev):
    """Lanczos on a shuffled restricted basis: E0 and vec in the caller's order."""
    n = 10
    rng = np.random.default_rng(5)
    edges = [(i, i + 2) for i in range(n - 2)]           # a next-nearest blockade graph
    codes = np.asarray(orc.independent_states(n, edges), np.int64)
    perm = rng.permutation(len(codes))
    idx = codes[perm]
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(5.0), terms=('I1',))
    Om, delta = np.ones(n), -1.0 * np.ones(n)
    out = pkg.hip_lanczos_ground_state(calib, Om, delta, basis_indices=idx, return_state=True,
                                       device=dev)
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], idx)
    ref = orc.ground_state_dense(H)
    assert abs(out['E0'] - ref['E0']) < 1e-10 * abs(ref['E0'])
    assert np.allclose(np.abs(out['vec']) ** 2, np.abs(ref['vec']) ** 2, atol=1e-9)
    with pytest.raises(ValueError):
        pkg.hip_lanczos_ground_state(calib, Om, delta, basis_indices=np.r_[idx, idx[:1]],
                                     device=dev)


def test_sector_scan_kagome_xy(pkg, dev):
    """Reference kagome N=12 XY calib: every sector's E0 against dense eigh."""
    calib = json.load(open(os.path.join(HERE, 'golden', 'kagome_n12.json')))
    n = len(calib['atoms'])
    delta = np.array([a.get('delta_mhz', 0.0) for a in calib['atoms']])
    out = pkg.scan_ground_sectors_hip(calib, delta=delta, device=dev)
    assert out['sectors_run'] == list(range(n + 1)) and out['converged']
    assert out['dimension_sum'] == out['full_dimension'] == 1 << n
    best = None
    for row in out['per_sector']:
        codes = orc.number_sector_basis(n, row['sector'])
        H = orc.build_H_sparse(n, np.zeros(n), delta, calib['pairs'], ['I2'], codes)
        e0 = orc.ground_state_dense(H)['E0']
        assert abs(row['E0'] - e0) < 1e-9 * max(1.0, abs(e0)), row
        best = e0 if best is None else min(best, e0)
    assert abs(out['E0'] - best) < 1e-9 * max(1.0, abs(best))
    with pytest.raises(ValueError):
        pkg.scan_ground_sectors_hip(calib, delta=delta, Om=np.ones(n), device=dev)


def test_full_basis_pair_channels_ground_state(pkg, dev):
    """Lanczos on the full basis with J and A (the tiled channel kernel in
    the fused three-term step) against eigsh of the oracle's CSR."""
    from scipy.sparse.linalg import eigsh
    n = 14
    rng = np.random.default_rng(14)
    terms = ['I1', 'I2', 'I3']
    pairs = _pairs(n, rng, cplx=False)
    Om, delta = rng.uniform(0.5, 1.5, n), rng.uniform(-2, 2, n)
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': terms}
    out = pkg.hip_lanczos_ground_state(calib, Om, delta, terms=terms, device=dev)
    H = orc.build_H_sparse(n, Om, delta, pairs, terms)
    E0 = float(eigsh(H.real, k=1, which='SA', tol=1e-12)[0][0])
    assert out['converged']
    assert abs(out['E0'] - E0) <= 1e-10 * abs(E0)


def test_full_basis_pair_channels_kpm(pkg, dev):
    """KPM moments (the probe-dot reductions of the tiled channel kernel) on
    the full basis with J and A against the oracle's estimator with the
    same Rademacher probes."""
    n = 13
    rng = np.random.default_rng(13)
    terms = ['I1', 'I2', 'I3']
    pairs = _pairs(n, rng, cplx=False)
    Om, delta = rng.uniform(0.5, 1.5, n), rng.uniform(-2, 2, n)
    P = _plan(pkg, dev, n, Om, delta, pairs, terms)
    H = orc.build_H_sparse(n, Om, delta, pairs, terms)
    emin, emax, c, r = orc.kpm_bounds(H, margin=0.01)
    mref, _, _ = orc.kpm_moments(H, 30, 2, seed=5, c=c, r=r)
    mu, _, _ = pkg.kpm_moments_hip(plan=P, num=30, nvec=2, seed=5, c=c, r=r)
    assert np.max(np.abs(mu - mref)) < 1e-10
    mu32, _, _ = pkg.kpm_moments_hip(plan=P, num=30, nvec=2, seed=5, c=c, r=r, dtype='float32')
    assert np.max(np.abs(mu32 - mref)) < 2e-4


@pytest.mark.parametrize('dtype,tol', [('complex128', 1e-10), ('complex64', 2e-5)])
def test_full_basis_pair_channels_segments(pkg, dev, dtype, tol):
    """Piecewise-constant evolution under Ω, Δ, V, J and A on the full basis
    (the Chebyshev recurrence through the tiled channel kernel) against the
    oracle's sparse exponential steps."""
    n = 13
    rng = np.random.default_rng(31)
    terms = ['I1', 'I2', 'I3']
    pairs = _pairs(n, rng, cplx=False)
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': terms}
    segs = [dict(t0=t0, Om=rng.uniform(0.5, 3.0, n), phase=rng.uniform(0, 2 * np.pi, n),
                 delta=rng.uniform(-2, 2, n)) for t0 in (0.0, 30.0)]
    t = np.array([0.0, 20.0, 50.0])
    out = pkg.propagate_segments_hip(calib, segs, t, terms=terms, dtype=dtype, device=dev,
                                     return_states=True)
    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1.0
    ref = orc.propagate_segments(n, pairs, terms, segs, psi0, t, method='expm_multiply')
    assert np.max(np.abs(out['states'] - ref)) < tol


def test_kagome_n24_sector_is_the_restricted_full_hv(pkg, dev):
    """Reference kagome N=24 XY calib (J channels conserve the excitation
    number): the full-basis H·v (D = 2^24) of a vector supported on the
    half-filled sector, restricted to that sector, equals the code-list H·v
    of the sector (D = 2,704,156) — two different rank modes of the channel
    kernel checked against each other at the full size."""
    import torch
    calib = json.load(open(os.path.join(HERE, 'golden', 'kagome_n24.json')))
    n = len(calib['atoms'])
    delta = np.array([a.get('delta_mhz', 0.0) for a in calib['atoms']])
    model = pkg.from_calib(calib, np.zeros(n), delta)
    codes = pkg.number_sector_basis(n, n // 2)
    rng = np.random.default_rng(24)
    xs = rng.normal(size=len(codes))
    Ps = pkg.Plan(model, basis='codes', codes=codes, device=dev)
    ys = Ps.hv(torch.as_tensor(xs, device=dev)).cpu().numpy()
    Ps.close()
    Pf = pkg.Plan(model, device=dev)
    xf = torch.zeros(1 << n, dtype=torch.float64, device=dev)
    ci = torch.as_tensor(codes, device=dev)
    xf[ci] = torch.as_tensor(xs, device=dev)
    yf = Pf.hv(xf)
    Pf.close()
    assert float(torch.linalg.vector_norm(yf).item() ** 2) == pytest.approx(
        float(np.sum(ys ** 2)), rel=1e-12)                 # nothing leaves the sector
    assert np.max(np.abs(yf[ci].cpu().numpy() - ys)) <= 1e-12 * np.max(np.abs(ys))
