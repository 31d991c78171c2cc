"""Chain-blockade (Fibonacci) sector on the GPU: bit-exact enumeration/ranking,
sector H·v and sector ground states against the oracle."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc
from oracle import hv_ref

pytestmark = pytest.mark.gpu


def _pairs(n, v1=30.0):
    return [{'i': i, 'j': j, 'V': v1 / (j - i) ** 6} for i in range(n) for j in range(i + 1, n)]


def _plan(pkg, dev, n, Om, delta, pairs):
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': ['I1']}
    return pkg.Plan(pkg.from_calib(calib, Om, delta), basis='chain_fib', device=dev)


@pytest.mark.parametrize('n', [1, 2, 3, 7, 12, 18, 19, 20])
def test_enumeration_matches_reference_order(pkg, dev, n):
    P = _plan(pkg, dev, n, np.ones(n), np.zeros(n), [])
    ref = np.asarray(orc.independent_states(n, [(i, i + 1) for i in range(n - 1)]))
    assert P.dim == len(ref) == orc.fibonacci(n + 2)
    codes = pkg.sector_codes(P).cpu().numpy()
    assert np.array_equal(codes, ref)
    ranks = pkg.sector_rank(P, torch_codes(ref, dev)).cpu().numpy()
    assert np.array_equal(ranks, np.arange(len(ref)))
    # codes outside the sector rank to -1
    if n >= 2:
        bad = np.array([3 << (n - 2), (1 << n) - 1, -5], np.int64)
        assert np.all(pkg.sector_rank(P, torch_codes(bad, dev)).cpu().numpy() == -1)


def torch_codes(a, dev):
    import torch
    return torch.as_tensor(np.asarray(a, np.int64), device=dev)


def test_enumeration_n40_properties(pkg, dev):
    """N=40 (D = F(42) = 267,914,296): sampled ranks round-trip, codes ascend,
    no adjacent excitations, closed-form rank."""
    import torch
    n = 40
    P = _plan(pkg, dev, n, np.ones(n), np.zeros(n), [])
    assert P.dim == 267914296
    rng = np.random.default_rng(0)
    starts = rng.integers(0, P.dim - 4096, size=16)
    for s in list(starts) + [0, P.dim - 4096]:
        c = pkg.sector_codes(P, first=int(s), count=4096).cpu().numpy()
        assert np.all(np.diff(c) > 0) and not np.any(c & (c >> 1)) and np.all(c < (1 << n))
        assert np.array_equal(orc.chain_fib_rank(n, c), np.arange(s, s + 4096))
        assert np.array_equal(orc.chain_fib_unrank(n, np.arange(s, s + 4096)), c)
    r = pkg.sector_rank(P, torch.as_tensor(c, device=dev)).cpu().numpy()
    assert np.array_equal(r, np.arange(P.dim - 4096, P.dim))


@pytest.mark.parametrize('n', [1, 2, 3, 5, 8, 11, 14])
def test_sector_hv_vs_channel_oracle(pkg, dev, n):
    import torch
    rng = np.random.default_rng(n)
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    pairs = _pairs(n)
    codes = np.asarray(orc.independent_states(n, [(i, i + 1) for i in range(n - 1)]))
    x = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, Om, delta, pairs, ['I1'], x, codes)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    # real drive, real vectors, both precisions
    Pr = _plan(pkg, dev, n, np.abs(Om), delta, pairs)
    xr = x.real.copy()
    refr = orc.channel_matvec(n, np.abs(Om), delta, pairs, ['I1'], xr, codes).real
    yr = Pr.hv(torch.as_tensor(xr, device=dev)).cpu().numpy()
    assert np.max(np.abs(yr - refr)) <= 1e-13 * max(1.0, np.max(np.abs(refr)))
    y32 = Pr.hv(torch.as_tensor(xr.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - refr)) <= 1e-5 * max(1.0, np.max(np.abs(refr)))


@pytest.mark.parametrize('n', [20, 24, 27, 30, 34, 36, 38])
def test_sector_hv_vs_c_oracle(pkg, dev, n):
    import torch
    Om, delta = np.ones(n), -1.5 * np.ones(n)
    pairs = _pairs(n)
    codes = hv_ref.fib_codes(n)
    d = hv_ref.diag(n, delta, pairs, codes)
    x = np.random.default_rng(n).standard_normal(len(codes))
    ref = hv_ref.hv_fib(n, Om, d, codes, x)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y32 = P.hv(torch.as_tensor(x.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_sector_ground_state_vs_projected_dense(pkg, dev):
    n = 12
    calib = {'atoms': [{}] * n, 'pairs': _pairs(n), 'terms': ['I1']}
    Om, delta = np.ones(n), -1.5 * np.ones(n)
    codes = np.asarray(orc.independent_states(n, [(i, i + 1) for i in range(n - 1)]))
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes)
    ref = orc.ground_state_dense(H)
    r = pkg.hip_lanczos_ground_state(calib, Om, delta, basis_indices=codes, return_state=True)
    assert r['converged'] and r['dim'] == len(codes)
    assert abs(r['E0'] - ref['E0']) <= 1e-10 * abs(ref['E0'])
    assert abs(abs(np.vdot(ref['vec'], r['vec'])) - 1) < 1e-8
    nref = orc.site_occupations(n, ref['vec'], codes)
    assert np.max(np.abs(r['n_atom'] - nref)) < 1e-8
    # any other code list runs the generic kernel (binary-search ranks)
    sub = codes[:-1]
    r2 = pkg.hip_lanczos_ground_state(calib, Om, delta, basis_indices=sub)
    H2 = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], sub)
    assert abs(r2['E0'] - orc.ground_state_dense(H2)['E0']) <= 1e-10 * abs(ref['E0'])


def test_sector_ground_state_fp32_refined_n22(pkg, dev):
    """FP32 bulk + FP64 refinement on the sector, against scipy eigsh on the C oracle."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    n = 22
    calib = {'atoms': [{}] * n, 'pairs': _pairs(n), 'terms': ['I1']}
    Om, delta = np.ones(n), -1.5 * np.ones(n)
    codes = hv_ref.fib_codes(n)
    d = hv_ref.diag(n, delta, calib['pairs'], codes)
    op = LinearOperator((len(codes), len(codes)), matvec=lambda v: hv_ref.hv_fib(n, Om, d, codes, np.ascontiguousarray(v.ravel(), np.float64)), dtype=np.float64)
    w = eigsh(op, k=2, which='SA', tol=1e-13, v0=np.random.default_rng(0).standard_normal(len(codes)))[0]
    e0 = float(np.min(w))
    r = pkg.hip_lanczos_ground_state(calib, Om, delta, basis='chain_blockade', dtype='float32')
    assert r['converged']
    assert abs(r['E0'] - e0) <= 1e-10 * abs(e0)


@pytest.mark.parametrize('n,dtype', [(36, 'complex128'), (36, 'complex64'), (38, 'complex128'),
                                     (40, 'float64'), (40, 'float32')])
def test_sector_hv_hermitian_large(pkg, dev, n, dtype):
    """<z, H x> = conj(<x, H z>) at sizes beyond the oracle (FP64 and complex
    use the two-pass H·v from n = 38 on, FP32 the single pass)."""
    import torch
    rng = np.random.default_rng(n)
    cplx = dtype.startswith('complex')
    Om = rng.uniform(0.5, 1.5, n) * (np.exp(1j * rng.uniform(0, 2 * np.pi, n)) if cplx else 1.0)
    delta = rng.uniform(-2, 2, n)
    P = _plan(pkg, dev, n, Om, delta, _pairs(n))
    tdt = getattr(torch, dtype)
    g = torch.Generator(device=dev)
    g.manual_seed(n)
    x = torch.randn(P.dim, dtype=tdt, device=dev, generator=g)
    z = torch.randn(P.dim, dtype=tdt, device=dev, generator=g)
    hx, hz = P.hv(x), P.hv(z)
    a = torch.vdot(z.to(torch.complex128 if cplx else torch.float64), hx.to(torch.complex128 if cplx else torch.float64)).item()
    b = torch.vdot(hz.to(torch.complex128 if cplx else torch.float64), x.to(torch.complex128 if cplx else torch.float64)).item()
    scale = (torch.linalg.vector_norm(hx.double() if not cplx else hx.to(torch.complex128)) *
             torch.linalg.vector_norm(z.double() if not cplx else z.to(torch.complex128))).item()
    tol = 1e-5 if dtype in ('complex64', 'float32') else 1e-12
    assert abs(a - b) <= tol * scale


@pytest.mark.parametrize('n', [18, 19, 20, 23])
def test_sector_hv_random_complex_vs_channel_oracle(pkg, dev, n):
    """Blocks with all 18 low bits: every site's Ω (complex) and Δ distinct,
    so a flip or diagonal term on the wrong bit changes the result."""
    import torch
    rng = np.random.default_rng(100 + n)
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    pairs = [{'i': i, 'j': j, 'V': float(rng.uniform(0.1, 5.0)) / (j - i) ** 3}
             for i in range(n) for j in range(i + 1, n)]
    codes = hv_ref.fib_codes(n)
    x = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, Om, delta, pairs, ['I1'], x, codes)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y64 = P.hv(torch.as_tensor(x.astype(np.complex64), device=dev)).cpu().numpy()
    assert np.max(np.abs(y64 - ref)) <= 2e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize('n', [21, 30, 38])
def test_sector_hv_random_real_vs_c_oracle(pkg, dev, n):
    """Real drive with distinct per-site Ω and Δ against the C oracle, FP64
    and FP32 (n = 38 runs the FP64 two-pass H·v)."""
    import torch
    rng = np.random.default_rng(200 + n)
    Om, delta = rng.uniform(0.5, 1.5, n), rng.normal(size=n)
    pairs = _pairs(n)
    codes = hv_ref.fib_codes(n)
    d = hv_ref.diag(n, delta, pairs, codes)
    x = rng.standard_normal(len(codes))
    ref = hv_ref.hv_fib(n, Om, d, codes, x)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y32 = P.hv(torch.as_tensor(x.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.max(np.abs(ref))
