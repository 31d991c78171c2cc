"""Batched Lanczos over arbitrary Hermitian instances (ryd_batch_create_csr):
uneven dims, real and complex, sparse and dense rows, against numpy's dense
eigensolver; the locked second run against eigh's w[1] with multiplicity;
graph replays against the eager sequence; the dense-stack mirror of
batched_ground_state (equilibrium.py:251-311) against eigh."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _rand_sparse_herm(d, rng, cplx, density=0.05):
    A = sp.random(d, d, density=density, random_state=rng, format='csr')
    if cplx:
        A = A + 1j * sp.random(d, d, density=density, random_state=rng, format='csr')
    H = (A + A.conj().T) * 0.5 + sp.diags(rng.normal(size=d))
    return sp.csr_matrix(H)


def _rand_dense_herm(d, rng, cplx):
    A = rng.normal(size=(d, d))
    if cplx:
        A = A + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _check(out, mats, tol_e=1e-9, tol_gap=1e-8):
    for b, H in enumerate(mats):
        Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
        w = np.linalg.eigvalsh(Hd)
        assert abs(out['E0'][b] - w[0]) <= tol_e * max(1.0, abs(w[0])), (b, out['E0'][b], w[0])
        assert abs(out['gap'][b] - (w[1] - w[0])) <= tol_gap * max(1.0, abs(w[1])), \
            (b, out['gap'][b], w[1] - w[0])


@pytest.mark.parametrize('cplx', [False, True])
def test_uneven_sparse_instances_vs_eigvalsh(pkg, dev, cplx):
    rng = np.random.default_rng(3 + cplx)
    dims = [17, 50, 128, 300, 1000, 64, 2]
    mats = [_rand_sparse_herm(d, rng, cplx) for d in dims]
    out = pkg.batched_ground_state_csr_hip(mats, device=dev)
    assert out['converged'].all()
    _check(out, mats)


def test_dense_rows_run_one_row_per_wavefront(pkg, dev):
    """Dense instances (rows of 256 entries) take the wavefront-per-row apply."""
    rng = np.random.default_rng(11)
    mats = [_rand_dense_herm(256, rng, True) for _ in range(3)] + [_rand_dense_herm(96, rng, False)]
    B = pkg.BatchCsrLanczos(mats, steps=20, dtype='float64', device=dev)
    import torch
    assert B.tdt == torch.complex128
    B.close()
    out = pkg.batched_ground_state_csr_hip(mats, device=dev)
    _check(out, mats)


def test_fp32_bulk_then_fp64_refinement(pkg, dev):
    rng = np.random.default_rng(5)
    mats = [_rand_sparse_herm(d, rng, True, density=0.02) for d in (400, 700, 256)]
    out = pkg.batched_ground_state_csr_hip(mats, dtype='float32', device=dev)
    assert out['converged'].all()
    _check(out, mats)


def test_degenerate_levels_counted_with_multiplicity(pkg, dev):
    """H ⊗ I_2 doubles every level and a diagonal whose minimum repeats has a
    double ground level: gap 0 there, as eigh gives, and the true gap for the
    simple instances of the same batch."""
    rng = np.random.default_rng(21)
    D = 64
    d = rng.uniform(0.0, 5.0, D)
    d[5] = d[40] = -1.0
    Hdiag = np.diag(d)
    rest = [i for i in range(D) if i not in (5, 40)]
    for i, j in zip(rest[:-1], rest[1:]):
        Hdiag[i, j] = Hdiag[j, i] = 0.05
    mats = [sp.kron(_rand_sparse_herm(60, rng, True, 0.1), sp.eye(2), format='csr'),
            sp.csr_matrix(Hdiag),
            _rand_sparse_herm(120, rng, True, 0.1),
            sp.kron(sp.csr_matrix(_rand_dense_herm(32, rng, False)), sp.eye(2), format='csr')]
    out = pkg.batched_ground_state_csr_hip(mats, device=dev)
    _check(out, mats)
    assert abs(out['gap'][0]) <= 1e-8 and abs(out['gap'][1]) <= 1e-8 and abs(out['gap'][3]) <= 1e-8
    assert out['gap'][2] > 1e-6


def test_occupations_and_graph_replays(pkg, dev):
    """n_atom of 2^n instances against the dense ground vector (site 0 the
    most significant bit); captured replays reproduce the eager run."""
    import torch
    rng = np.random.default_rng(9)
    mats = [_rand_sparse_herm(1 << n, rng, False, 0.1) for n in (6, 8, 7)]
    out = pkg.batched_ground_state_csr_hip(mats, device=dev)
    for b, H in enumerate(mats):
        w, v = np.linalg.eigh(H.toarray())
        if w[1] - w[0] < 1e-3:
            continue
        n = int(np.log2(H.shape[0]))
        p = np.abs(v[:, 0]) ** 2
        idx = np.arange(H.shape[0])
        ref = np.array([p[((idx >> (n - 1 - i)) & 1) == 1].sum() for i in range(n)])
        assert np.max(np.abs(out['n_atom'][b] - ref)) < 1e-7
    B = pkg.BatchCsrLanczos(mats, steps=60, dtype='float64', device=dev, store_basis=False)
    v0 = B.start_vectors(4)
    B.run(v0, use_graph=False)
    a0, n0, _ = B.coefficients()
    for rep in range(3):
        B.run(v0, use_graph=True)
        a, nr, _ = B.coefficients()
        assert np.array_equal(a, a0) and np.array_equal(nr, n0), rep
        torch.randn(1000, device=dev).sum().item()
    B.close()


def test_dense_stack_mirror_matches_eigh(pkg, dev):
    rng = np.random.default_rng(2)
    Hs = np.stack([_rand_dense_herm(64, rng, True) for _ in range(5)])
    out = pkg.batched_ground_state_dense_hip(Hs, device=dev)
    assert out['backend'] == 'hip-batched-csr-lanczos'
    assert out['n_atom'].shape == (5, 6)
    for b in range(5):
        w = np.linalg.eigvalsh(Hs[b])
        assert abs(out['E0'][b] - w[0]) <= 1e-9 * max(1.0, abs(w[0]))
        assert abs(out['gap'][b] - (w[1] - w[0])) <= 1e-8


def test_dense_stack_workload_small(pkg, dev):
    """The dense-stack workload at a small size: its builder agrees with the
    device H·v and the batched results with eigvalsh."""
    from energy_level_amd import workloads as W
    r = W.run_d1(dev, count=8, n=8, sequential=2)
    assert r['builder_rel_err'] <= 1e-12
    assert r['E0_err_vs_eigvalsh'] <= 1e-9 * 10
    assert r['gap_err_vs_eigvalsh'] <= 1e-8


@pytest.mark.parametrize('cplx,dims', [
    (False, [2, 5, 100, 1023, 1024, 1025, 3000, 8192]),   # one launch per run (<= 8192 rows)
    (False, [300, 9000]),                                 # 9000 rows: the multi-launch form
    (True, [3, 64, 1000, 4096]),                          # complex128: one launch up to 4096
])
@pytest.mark.parametrize('store', [True, False])
def test_batch_coefficients_match_the_oracle_recurrence(pkg, dev, cplx, dims, store):
    """Each instance's alpha/beta of one batched run against the oracle's
    Lanczos from the same start vector (first min(20, dim-1) steps, where the
    plain recurrence stays within rounding of the reorthogonalised one)."""
    from oracle import rydberg_oracle as orc
    rng = np.random.default_rng(11 + len(dims) + cplx)
    mats = [_rand_sparse_herm(d, rng, cplx, density=min(1.0, 6.0 / d)) for d in dims]
    B = pkg.BatchCsrLanczos(mats, steps=20, dtype='float64', store_basis=store)
    try:
        v0 = B.start_vectors(seed=4)
        B.run(v0, use_graph=False)
        alpha, nrm, stop = B.coefficients()
        v = v0.cpu().numpy()
        for b, H in enumerate(mats):
            d = dims[b]
            K = 2 if d == 2 else min(20, d - 1)
            x0 = v[B.offsets[b]:B.offsets[b + 1]].astype(complex)
            ra, rb, n0sq = orc.lanczos(lambda q: H @ q, x0, K)
            hn = abs(H).sum(axis=1).max()
            assert abs(nrm[b, 0] ** 2 - n0sq) <= 1e-12 * n0sq
            k = len(ra)
            assert np.max(np.abs(alpha[b, :k] - ra)) <= 1e-9 * hn, (b, d)
            assert np.max(np.abs(nrm[b, 1:1 + len(rb)] - rb), initial=0.0) <= 1e-9 * hn, (b, d)
            if d == 2:
                assert stop[b] == 1
    finally:
        B.close()
