"""A plan over an assembled sparse matrix (ryd_plan_create_csr): H·v against
scipy, the kpm_moments_gpu mirror (kpm.py:26-75) against the oracle's
recurrence on the same device-drawn probes, and the Lanczos ground state
against eigsh."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu


def _hermitian(D, density, rng, cplx):
    A = sp.random(D, D, density=density, random_state=rng, format='csr')
    if cplx:
        B = sp.random(D, D, density=density, random_state=rng, format='csr')
        A = A + 1j * B
    return ((A + A.conj().T) * 0.5 + sp.diags(rng.normal(size=D))).tocsr()


@pytest.mark.parametrize('cplx', [False, True])
def test_csr_hv(pkg, dev, cplx):
    import torch
    rng = np.random.default_rng(3)
    D = 3001
    H = _hermitian(D, 0.004, rng, cplx)
    P = pkg.CsrPlan(H, device=dev)
    assert P.dim == D and P.is_real == (not cplx)
    x = rng.normal(size=D) + (1j * rng.normal(size=D) if cplx else 0)
    ref = H @ x
    dt = torch.complex128 if cplx else torch.float64
    y = P.hv(torch.as_tensor(x, device=dev).to(dt)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    dt32 = torch.complex64 if cplx else torch.float32
    y32 = P.hv(torch.as_tensor(x, device=dev).to(dt32)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.max(np.abs(ref))
    with pytest.raises(NotImplementedError):              # no site codes
        pkg.site_occupations(P, torch.as_tensor(x, device=dev).to(dt))
    P.close()


def test_kpm_moments_gpu_mirror(pkg, dev):
    import torch
    rng = np.random.default_rng(5)
    D, num, nvec, seed, chunk = 2000, 40, 5, 11, 2
    H = _hermitian(D, 0.005, rng, True)
    mu, c, r = pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, seed=seed, device=dev,
                                       chunk=chunk)
    # the reference's probe draw (kpm.py:60-63), repeated here on the device
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    probes, done = [], 0
    while done < nvec:
        b = min(chunk, nvec - done)
        z = torch.randint(0, 2, (D, b), device=dev, generator=gen, dtype=torch.int8)
        z = z.to(torch.float64).mul_(2).sub_(1).cpu().numpy()
        probes += [z[:, j] for j in range(b)]
        done += b
    w = np.linalg.eigvalsh(H.toarray())
    assert c - r <= w[0] and c + r >= w[-1]
    assert abs(r - (w[-1] - w[0]) / 2 * (1 + 1e-3) - 1e-12) < 1e-8 * r
    ref, _, _ = orc.kpm_moments(H, num, nvec, c=c, r=r, probes=probes)
    assert mu.shape == (num + 1,)
    assert np.max(np.abs(mu - ref)) < 1e-10
    mu32, _, _ = pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, seed=seed, device=dev,
                                         chunk=chunk, dtype='float32')
    assert np.max(np.abs(mu32 - ref)) < 2e-4


def test_csr_lanczos_ground_state(pkg, dev):
    from scipy.sparse.linalg import eigsh
    rng = np.random.default_rng(8)
    H = _hermitian(4000, 0.003, rng, False)
    P = pkg.CsrPlan(H, device=dev)
    out = pkg.hip_lanczos_ground_state(None, plan=P, device=dev)
    E0 = float(eigsh(H, k=1, which='SA', tol=1e-12)[0][0])
    assert out['converged'] and abs(out['E0'] - E0) <= 1e-9 * abs(E0)
    P.close()


def test_batched_ground_state_dense_signature(pkg, dev):
    """batched_ground_state (equilibrium.py:251-311) on a dense Hermitian
    stack: E0, gap and <n_i> (site 0 = most significant bit) against eigh."""
    rng = np.random.default_rng(12)
    B, n = 3, 7
    dim = 1 << n
    Hs = []
    for _ in range(B):
        A = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
        Hs.append((A + A.conj().T) / 2)
    out = pkg.batched_ground_state_dense_hip(np.array(Hs), device=dev)
    bits = np.array([[(s >> (n - 1 - i)) & 1 for s in range(dim)] for i in range(n)], float)
    for b, H in enumerate(Hs):
        w, V = np.linalg.eigh(H)
        assert abs(out['E0'][b] - w[0]) <= 1e-9 * abs(w[0])
        assert abs(out['gap'][b] - (w[1] - w[0])) <= 1e-7
        assert np.allclose(out['n_atom'][b], bits @ np.abs(V[:, 0]) ** 2, atol=1e-7)
    with pytest.raises(ValueError):
        pkg.batched_ground_state_dense_hip(np.zeros((2, 6, 6)), device=dev)


def test_kpm_gpu_mirror_budget_rule(pkg, dev):
    """kpm.py:41-47: a budget that holds the operator and k probes shrinks
    the probe block to k (and so draws the stream of chunk=k); one below a
    single probe, or a non-positive budget, is a ValueError."""
    rng = np.random.default_rng(7)
    D, num, nvec, seed = 1500, 20, 6, 3
    H = _hermitian(D, 0.006, rng, False)
    itemsize = 16
    base = (H.nnz + D) * (itemsize + 16) * 3
    per = D * itemsize * 6
    budget_mb = (base + 2 * per + per // 2) / 1024 ** 2        # room for 2 probes
    mu_b, c, r = pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, seed=seed, device=dev, chunk=32,
                                         memory_budget_mb=budget_mb)
    mu_2, c2, r2 = pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, seed=seed, device=dev, chunk=2)
    assert (c, r) == (c2, r2)
    assert np.array_equal(mu_b, mu_2)
    mu_32, _, _ = pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, seed=seed, device=dev, chunk=32)
    assert not np.array_equal(mu_32, mu_2)        # the block size shapes the probe draw
    with pytest.raises(ValueError):
        pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, device=dev,
                                memory_budget_mb=(base + per // 2) / 1024 ** 2)
    with pytest.raises(ValueError):
        pkg.kpm_moments_gpu_hip(H, num=num, nvec=nvec, device=dev, memory_budget_mb=0)


def test_kpm_bounds_enclose_both_ends_and_moment_guard(pkg, dev):
    """The device bounds converge the highest Ritz value as well as the
    lowest (eigsh SA and LA, kpm.py:92-93): [emin, emax] covers eigvalsh.
    Bounds that leave eigenvalues outside make |mu_n| exceed 1, which is
    refused."""
    rng = np.random.default_rng(9)
    D = 2500
    H = _hermitian(D, 0.004, rng, True)
    w = np.linalg.eigvalsh(H.toarray())
    P = pkg.CsrPlan(H, device=dev)
    emin, emax, c, r = pkg.spectral_bounds(P, margin=1e-3)
    assert emin <= w[0] + 1e-9 and emax >= w[-1] - 1e-9
    assert c - r <= w[0] and c + r >= w[-1]
    with pytest.raises(RuntimeError):
        pkg.kpm_moments_hip(plan=P, num=30, nvec=2, seed=1, c=c, r=r / 3.0)
    P.close()
