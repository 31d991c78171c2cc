"""GPU parity of the full-basis fused H·v kernels against the oracle."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc
from oracle import hv_ref

pytestmark = pytest.mark.gpu


def _chain(n, c6=10.0):
    return [{'i': i, 'j': j, 'V': c6 / (j - i) ** 6} for i in range(n) for j in range(i + 1, n)]


def _rand_model(n, seed, complex_om=False):
    rng = np.random.default_rng(seed)
    Om = rng.uniform(0.3, 1.7, n)
    if complex_om:
        Om = Om * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    pairs = [{'i': i, 'j': j, 'V': float(rng.normal() * 3)} for i in range(n)
             for j in range(i + 1, n) if rng.random() < 0.7]
    return Om, delta, pairs


def _plan(pkg, dev, n, Om, delta, pairs):
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': ['I1']}
    return pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)


@pytest.mark.parametrize('n', list(range(1, 15)))
def test_hv_fp64_real_vs_csr(pkg, dev, n):
    import torch
    Om, delta, pairs = _rand_model(n, n)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    H = orc.build_H_sparse(n, Om, delta, pairs, ['I1'])
    x = np.random.default_rng(n + 50).normal(size=1 << n)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    ref = H @ x
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize('n', [1, 3, 6, 10, 11, 12, 13])
def test_hv_complex_vs_kron(pkg, dev, n):
    import torch
    Om, delta, pairs = _rand_model(n, 100 + n, complex_om=True)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    H = orc.build_H_sparse(n, Om, delta, pairs, ['I1'])
    if n <= 10:
        Hk = orc.kron_hamiltonian(n, Om, delta, pairs, ['I1'])
        assert np.max(np.abs(H.toarray() - Hk)) < 1e-12
    rng = np.random.default_rng(n)
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    ref = H @ x
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.linalg.norm(ref)
    y32 = P.hv(torch.as_tensor(x.astype(np.complex64), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 2e-6 * np.linalg.norm(ref)
    with pytest.raises(ValueError):
        P.hv(torch.as_tensor(x.real, device=dev))     # complex phases need complex vectors


@pytest.mark.parametrize('n', [12, 13, 16, 18, 20, 22, 23, 24])
def test_hv_fp64_fp32_vs_c_oracle(pkg, dev, n):
    import torch
    Om = np.ones(n)
    delta = -1.2 * np.ones(n)
    pairs = _chain(n)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    d = hv_ref.diag(n, delta, pairs)
    x = np.random.default_rng(0).standard_normal(1 << n)
    ref = hv_ref.hv_full(n, Om, d, x)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y32 = P.hv(torch.as_tensor(x.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_hv_n26_hermitian_and_linear(pkg, dev):
    """Size-independent properties at a large size: <a,Hb> = <Ha,b>, linearity,
    and a random 4096-element window against the C oracle formula."""
    import torch
    n = 26
    Om = np.ones(n)
    delta = -1.2 * np.ones(n)
    P = _plan(pkg, dev, n, Om, delta, _chain(n))
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    a = torch.randn(1 << n, dtype=torch.float64, device=dev, generator=g)
    b = torch.randn(1 << n, dtype=torch.float64, device=dev, generator=g)
    Ha, Hb = P.hv(a), P.hv(b)
    lhs = float(torch.dot(a, Hb))
    rhs = float(torch.dot(Ha, b))
    assert abs(lhs - rhs) <= 1e-12 * float(torch.linalg.vector_norm(Ha) * torch.linalg.vector_norm(b))
    Hab = P.hv(a + 2.0 * b)
    assert float(torch.max(torch.abs(Hab - (Ha + 2.0 * Hb)))) <= 1e-12 * float(torch.max(torch.abs(Hab)))
    # spot check: element-wise formula on a window of codes with high bits set
    idx = np.arange(3 << 24, (3 << 24) + 4096)
    aa = a.cpu().numpy()
    d = hv_ref.diag(n, delta, _chain(n), idx)
    ref = d * aa[idx]
    for i in range(n):
        ref = ref + 0.5 * Om[i] * aa[idx ^ (1 << (n - 1 - i))]
    got = Ha.cpu().numpy()[idx]
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize('n', [5, 14])
def test_apply_three_term_and_sums(pkg, dev, n):
    import torch
    Om, delta, pairs = _rand_model(n, 7 + n)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    H = orc.build_H_sparse(n, Om, delta, pairs, ['I1'])
    rng = np.random.default_rng(1)
    x, z = rng.normal(size=1 << n), rng.normal(size=1 << n)
    s, ch, cx, cz = 0.7, 1.3, -0.4, -1.0
    ref = ch * (H @ (s * x)) + cx * s * x + cz * z
    sums = torch.zeros(4, dtype=torch.float64, device=dev)
    xt, zt = torch.as_tensor(x, device=dev), torch.as_tensor(z, device=dev)
    yt = torch.empty_like(xt)
    words = ((1 << n) + 31) // 32
    pbits = rng.integers(0, 2 ** 32, size=words, dtype=np.uint64).astype(np.uint32)
    p = np.unpackbits(pbits.view(np.uint8), bitorder='little')[:1 << n].astype(float) * 2 - 1
    P.apply(xt, yt, s=s, ch=ch, cx=cx, cz=cz, z=zt, probe=torch.as_tensor(pbits, device=dev),
            sums=sums)
    y = yt.cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    S = sums.cpu().numpy()
    sx = s * x
    assert np.allclose(S, [sx @ ref, ref @ ref, sx @ sx, p @ ref], rtol=1e-12, atol=0)
    # accumulate mode adds
    P.apply(xt, yt, s=s, ch=ch, cx=cx, cz=cz, z=zt, sums=sums, accumulate=True)
    assert np.allclose(sums.cpu().numpy()[:3], 2 * S[:3], rtol=1e-12)


def test_site_occupations(pkg, dev):
    import torch
    for n in (3, 9, 15):
        P = _plan(pkg, dev, n, np.ones(n), np.zeros(n), [])
        x = np.random.default_rng(n).normal(size=1 << n)
        got = pkg.site_occupations(P, torch.as_tensor(x, device=dev))
        ref = orc.site_occupations(n, x / np.linalg.norm(x))
        assert np.max(np.abs(got - ref)) < 1e-13


@pytest.mark.parametrize('n', [23, 24])
def test_hv_split_complex_and_fused_sums(pkg, dev, n):
    """Sizes that run the two-pass (contiguous + top-bit) H·v: complex drive,
    three-term update and reductions against the C oracle."""
    import torch
    rng = np.random.default_rng(n)
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    pairs = _chain(n)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    d = hv_ref.diag(n, delta, pairs)
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    ref = hv_ref.hv_full(n, Om, d, x)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y64 = P.hv(torch.as_tensor(x.astype(np.complex64), device=dev)).cpu().numpy()
    assert np.max(np.abs(y64 - ref)) <= 2e-6 * np.max(np.abs(ref))
    # real model, fused three-term update with sums
    Omr = np.abs(Om)
    Pr = _plan(pkg, dev, n, Omr, delta, pairs)
    xr, zr = rng.normal(size=1 << n), rng.normal(size=1 << n)
    s, ch, cx, cz = 0.5, 2.0, -0.3, -1.0
    ref = ch * hv_ref.hv_full(n, Omr, d, s * xr) + cx * s * xr + cz * zr
    sums = torch.zeros(4, dtype=torch.float64, device=dev)
    xt, zt = torch.as_tensor(xr, device=dev), torch.as_tensor(zr, device=dev)
    yt = torch.empty_like(xt)
    Pr.apply(xt, yt, s=s, ch=ch, cx=cx, cz=cz, z=zt, sums=sums)
    yr = yt.cpu().numpy()
    assert np.max(np.abs(yr - ref)) <= 1e-13 * np.max(np.abs(ref))
    S = sums.cpu().numpy()
    assert np.allclose(S[:3], [(s * xr) @ ref, ref @ ref, (s * xr) @ (s * xr)], rtol=1e-11)


@pytest.mark.parametrize('n', [23, 25])
def test_hv_split_random_model_vs_c_oracle(pkg, dev, n):
    """Two-pass H·v with every site's Ω, Δ and random couplings distinct, so
    a flip or diagonal term on the wrong bit changes the result; n = 25 runs
    the FP32 split as well."""
    import torch
    Om, delta, pairs = _rand_model(n, 40 + n)
    P = _plan(pkg, dev, n, Om, delta, pairs)
    d = hv_ref.diag(n, delta, pairs)
    x = np.random.default_rng(n).standard_normal(1 << n)
    ref = hv_ref.hv_full(n, Om, d, x)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y32 = P.hv(torch.as_tensor(x.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.max(np.abs(ref))
