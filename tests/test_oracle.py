"""Pins the CPU oracle against closed forms, independent constructions and the
reference's own known answers (CPU only)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import rydberg_oracle as orc
from oracle import hv_ref


def _random_model(n, seed, complex_terms=True):
    rng = np.random.default_rng(seed)
    Om = rng.uniform(0.3, 1.5, n) * (np.exp(1j * rng.uniform(0, 2 * np.pi, n)) if complex_terms else 1)
    delta = rng.normal(size=n)
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            p = {'i': i, 'j': j, 'V': float(rng.normal())}
            if complex_terms:
                p['J'] = complex(rng.normal(), rng.normal())
                p['A'] = complex(rng.normal(), rng.normal())
            pairs.append(p)
    return Om, delta, pairs


@pytest.mark.parametrize('n', [1, 2, 3, 4, 5])
def test_channels_equal_kron_construction(n):
    # reference checks/equibench_engine_ed_selfcheck.py:128-163 (E1): engine ≡ textbook kron
    Om, delta, pairs = _random_model(n, 100 + n)
    terms = ['I1', 'I2', 'I3']
    Hk = orc.kron_hamiltonian(n, Om, delta, pairs, terms)
    Hs = orc.build_H_sparse(n, Om, delta, pairs, terms).toarray()
    assert np.max(np.abs(Hk - Hs)) < 1e-12
    v = np.random.default_rng(n).normal(size=1 << n) + 1j * np.random.default_rng(n + 9).normal(size=1 << n)
    assert np.max(np.abs(orc.channel_matvec(n, Om, delta, pairs, terms, v) - Hk @ v)) < 1e-12


def test_hand_written_two_site_matrix():
    # reference checks/equibench_engine_ed_selfcheck.py:245-264 (A3)
    O0, O1, d0, d1 = 3.0, 5.0, 1.0, -2.0
    V, J, A = 30.0, -1.2 + 0.7j, 0.9 - 1.3j
    M = np.zeros((4, 4), complex)
    M[1, 1], M[2, 2], M[3, 3] = d1, d0, d0 + d1 + V
    M[0, 1] = M[1, 0] = M[2, 3] = M[3, 2] = O1 / 2
    M[0, 2] = M[2, 0] = M[1, 3] = M[3, 1] = O0 / 2
    M[2, 1], M[1, 2] = J, np.conj(J)
    M[3, 0], M[0, 3] = A, np.conj(A)
    pairs = [{'i': 0, 'j': 1, 'V': V, 'J': J, 'A': A}]
    H = orc.build_H_sparse(2, [O0, O1], [d0, d1], pairs, ['I1', 'I2', 'I3']).toarray()
    assert np.max(np.abs(H - M)) < 1e-12


def test_xx_dimer_spectrum():
    # reference checks/equibench_engine_ed_selfcheck.py:196-208 (A1)
    J = 1.7
    H = orc.build_H_sparse(2, [0, 0], [0, 0], [{'i': 0, 'j': 1, 'V': 0.0, 'J': J, 'A': 0.0}],
                           ['I1', 'I2', 'I3']).toarray()
    assert np.allclose(np.sort(np.linalg.eigvalsh(H)), [-J, 0, 0, J], atol=1e-10)


def test_dtype_derivation():
    # reference channels.py:41-56
    assert orc.model_dtype([1.0, 2.0], [{'i': 0, 'j': 1, 'V': 1.0}], ['I1']) == np.float64
    assert orc.model_dtype([1.0j, 2.0], [], ['I1']) == np.complex128
    assert orc.model_dtype([1.0, 2.0], [{'i': 0, 'j': 1, 'J': 1j}], ['I2']) == np.complex128
    assert orc.model_dtype([1.0, 2.0], [], [], gamma=[0.1, 0]) == np.complex128


def test_restricted_basis_matvec_unsorted():
    # reference checks/manybody_channel_selfcheck.py:26-123: restricted and unsorted bases
    n = 3
    Om, delta, pairs = _random_model(n, 7)
    terms = ['I1', 'I2', 'I3']
    H = orc.build_H_sparse(n, Om, delta, pairs, terms).toarray()
    for codes in ([1, 2, 4, 7], [7, 1, 4, 2]):
        codes = np.asarray(codes)
        v = np.random.default_rng(3).normal(size=len(codes)) + 0j
        ref = H[np.ix_(codes, codes)] @ v
        assert np.max(np.abs(orc.channel_matvec(n, Om, delta, pairs, terms, v, codes) - ref)) < 1e-12
        Hs = orc.build_H_sparse(n, Om, delta, pairs, terms, codes).toarray()
        assert np.max(np.abs(Hs - H[np.ix_(codes, codes)])) < 1e-12


@pytest.mark.parametrize('n', [3, 5, 8, 12])
def test_chain_blockade_basis(n):
    edges = [(i, i + 1) for i in range(n - 1)]
    st = orc.independent_states(n, edges)
    bf = orc.independent_states_bruteforce(n, edges)
    assert np.array_equal(np.asarray(st), bf)                   # ascending order
    assert len(st) == orc.fibonacci(n + 2)                      # F(N+2)
    assert np.array_equal(orc.chain_fib_rank(n, st), np.arange(len(st)))
    assert np.array_equal(orc.chain_fib_unrank(n, np.arange(len(st))), np.asarray(st))


def test_fibonacci_counts_reference_values():
    # reference checks/equibench_engine_ed_selfcheck.py:226-238
    assert [orc.fibonacci(k) for k in (5, 6, 7)] == [5, 8, 13]
    assert orc.fibonacci(42) == 267914296


@pytest.mark.parametrize('n,k', [(6, 0), (6, 3), (10, 4), (12, 12)])
def test_number_sector_basis(n, k):
    s = np.arange(1 << n)
    pc = np.array([bin(x).count('1') for x in s])
    assert np.array_equal(orc.number_sector_basis(n, k), s[pc == k])


def test_kpm_exact_diagonal_known_answer():
    # reference checks/equibench_gpu_backends_selfcheck.py:27-32
    lam = np.linspace(-2.0, 3.0, 32)
    H = np.diag(lam)
    mu, c, r = orc.kpm_moments(H, 24, 1, exact=True)
    assert abs(c - 0.5) < 1e-12 and abs(r - (2.5 * 1.001 + 1e-12)) < 1e-12
    x = (lam - c) / r
    ref = np.array([np.mean(np.cos(k * np.arccos(x))) for k in range(25)])
    assert np.max(np.abs(mu - ref)) < 1e-12
    # Hutchinson is exact on a diagonal matrix (z_s^2 = 1) for any probe
    mu2, _, _ = orc.kpm_moments(H, 24, 3, seed=5, c=c, r=r)
    assert np.max(np.abs(mu2 - ref)) < 1e-12


def test_jackson_kernel_closed_form():
    g = orc.jackson_kernel(10)
    assert abs(g[0] - 1.0) < 1e-14
    M = 10
    n = 3
    a = np.pi / (M + 1)
    assert abs(g[n] - ((M - n + 1) * np.cos(a * n) + np.sin(a * n) / np.tan(a)) / (M + 1)) < 1e-14
    assert np.all(np.diff(g) < 0)


def _kpm_selfcheck_model():
    # the random complex N=8 model with J and A of reference
    # checks/equibench_kpm_selfcheck.py:39-49 (nearest and next-nearest pairs)
    rng = np.random.default_rng(7)
    n = 8
    pairs = [{'i': i, 'j': i + 1, 'V': float(rng.uniform(0, 2)),
              'J': complex(rng.normal() + 1j * rng.normal()),
              'A': complex(rng.normal() + 1j * rng.normal())} for i in range(n - 1)]
    pairs += [{'i': i, 'j': i + 2, 'V': float(rng.uniform(0, 1)),
               'J': complex(rng.normal(0, .5)), 'A': complex(rng.normal(0, .5))} for i in range(n - 2)]
    Om = rng.uniform(0, 1, n)
    delta = rng.uniform(-1, 1, n)
    return n, orc.build_H_sparse(n, Om, delta, pairs, ['I1', 'I2', 'I3'])


def test_thermal_from_moments_vs_exact_spectrum():
    # reference checks/equibench_kpm_selfcheck.py:52-63: KPM thermodynamics from
    # exact moments equals the spectral thermodynamics, E < 1e-5, C < 1e-4
    n, H = _kpm_selfcheck_model()
    D = 1 << n
    evals = np.linalg.eigvalsh(H.toarray())
    betas = np.array([0.1, 0.5, 2.0])
    mu, c, r = orc.kpm_moments(H, 512, 1, exact=True)
    out = orc.thermal_from_moments(D, betas, mu, c, r, 512)
    for i, b in enumerate(betas):
        w = np.exp(-b * (evals - evals[0]))
        Z = w.sum()
        E = float(evals @ w / Z)
        C = float(b * b * ((evals ** 2) @ w / Z - E * E))
        lnZ = float(np.log(Z) - b * evals[0])
        assert abs(out['E'][i] - E) < 1e-5, (b, out['E'][i], E)
        assert abs(out['C'][i] - C) < 1e-4, (b, out['C'][i], C)
        assert abs(out['lnZ'][i] - lnZ) < 1e-6 * max(1.0, abs(lnZ)), (b, out['lnZ'][i], lnZ)
        assert abs(out['F'][i] + lnZ / b) < 1e-6 * max(1.0, abs(lnZ / b)), (b, out['F'][i], -lnZ / b)
    # the estimator from a few probes is deterministic for a seed
    m1, _, _ = orc.kpm_moments(H, 64, 4, seed=3, c=c, r=r)
    m2, _, _ = orc.kpm_moments(H, 64, 4, seed=3, c=c, r=r)
    assert np.array_equal(m1, m2)


def test_thermal_high_temperature_limit():
    # β → 0: E → Tr(H)/D and C/β² → Var of the spectrum (free of any interpolation)
    n, H = _kpm_selfcheck_model()
    D = 1 << n
    evals = np.linalg.eigvalsh(H.toarray())
    mu, c, r = orc.kpm_moments(H, 128, 1, exact=True)
    b = 1e-4
    out = orc.thermal_from_moments(D, [b], mu, c, r, 128)
    assert abs(out['E'][0] - evals.mean()) < 1e-3 * max(1.0, np.ptp(evals))
    assert abs(out['C'][0] / b ** 2 - evals.var()) < 1e-2 * evals.var()


def test_dos_normalisation():
    lam = np.linspace(-2.0, 3.0, 64)
    mu, c, r = orc.kpm_moments(np.diag(lam), 200, 1, exact=True)
    E = np.linspace(c - 0.99999 * r, c + 0.99999 * r, 20001)
    rho = orc.dos_from_moments(mu, c, r, E, 200)
    assert abs(np.trapezoid(rho, E) - 1.0) < 1e-2
    # uniform spectrum on [-2, 3]: density 1/5 in the bulk
    assert abs(orc.dos_from_moments(mu, c, r, [0.5], 200)[0] - 0.2) < 5e-3


def test_rabi_closed_form():
    # reference checks/manybody_m1_ed_selfcheck.py:43-54
    Om = 2 * np.pi * 5.0
    t = np.linspace(0.0, 200.0, 21)
    seg = [{'t0': 0.0, 'Om': [Om], 'delta': [0.0]}]
    psi0 = np.array([1.0, 0.0], complex)
    psi = orc.propagate_segments(1, [], ['I1'], seg, psi0, t)
    pred = np.sin(orc.CONV * Om * t / 2) ** 2
    assert np.max(np.abs(np.abs(psi[:, 1]) ** 2 - pred)) < 1e-10


def test_segment_jump_exact():
    # reference checks/manybody_contract_selfcheck.py:201-206: a drive jump inside
    # an output interval is integrated exactly at the jump
    n = 2
    pairs = [{'i': 0, 'j': 1, 'V': 3.0}]
    segs = [{'t0': 0.0, 'Om': [1.0, 2.0], 'delta': [0.5, -0.3]},
            {'t0': 0.7, 'Om': [2.0, 0.5], 'delta': [-1.0, 0.2]}]
    psi0 = np.zeros(4, complex)
    psi0[0] = 1.0
    out = orc.propagate_segments(n, pairs, ['I1'], segs, psi0, np.array([0.0, 1.0]))
    H1 = orc.kron_hamiltonian(n, [1.0, 2.0], [0.5, -0.3], pairs, ['I1'])
    H2 = orc.kron_hamiltonian(n, [2.0, 0.5], [-1.0, 0.2], pairs, ['I1'])
    ref = sla.expm(-1j * orc.CONV * 0.3 * H2) @ sla.expm(-1j * orc.CONV * 0.7 * H1) @ psi0
    assert np.max(np.abs(out[-1] - ref)) < 1e-12


def test_lanczos_reference_form_tridiagonal():
    # reference spectrum.py:34-103: with full reorthogonalisation the
    # tridiagonal's extreme eigenvalue is the matrix's
    n = 6
    Om, delta, pairs = _random_model(n, 11, complex_terms=False)
    H = orc.build_H_sparse(n, Om, delta, pairs, ['I1']).toarray()
    a, b, _ = orc.lanczos(lambda v: H @ v, np.random.default_rng(0).normal(size=1 << n), 40)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert abs(np.linalg.eigvalsh(T)[0] - np.linalg.eigvalsh(H)[0]) < 1e-10


@pytest.mark.parametrize('n', [4, 9, 13])
def test_c_oracle_matches_numpy(n):
    rng = np.random.default_rng(n)
    Om = rng.uniform(0.5, 1.5, n)
    delta = rng.normal(size=n)
    pairs = [{'i': i, 'j': j, 'V': 10.0 / (j - i) ** 6} for i in range(n) for j in range(i + 1, n)]
    d = hv_ref.diag(n, delta, pairs)
    x = rng.normal(size=1 << n)
    ref = orc.channel_matvec(n, Om, delta, pairs, ['I1'], x)
    assert np.max(np.abs(hv_ref.hv_full(n, Om, d, x) - ref)) < 1e-12 * np.max(np.abs(ref))
    Omc = Om * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    xc = x + 1j * rng.normal(size=1 << n)
    refc = orc.channel_matvec(n, Omc, delta, pairs, ['I1'], xc)
    assert np.max(np.abs(hv_ref.hv_full(n, Omc, d, xc) - refc)) < 1e-12 * np.max(np.abs(refc))
    # chain sector
    codes = hv_ref.fib_codes(n)
    assert np.array_equal(codes, np.asarray(orc.independent_states(n, [(i, i + 1) for i in range(n - 1)])))
    ds = hv_ref.diag(n, delta, pairs, codes)
    xs = rng.normal(size=len(codes))
    refs = orc.channel_matvec(n, Om, delta, pairs, ['I1'], xs, codes)
    assert np.max(np.abs(hv_ref.hv_fib(n, Om, ds, codes, xs) - refs)) < 1e-12 * np.max(np.abs(refs))


def test_pauli_expect_vs_kronecker_and_projection():
    """The oracle's Pauli-string expectation (PauliPlan basis path) equals the
    dense Kronecker operator on the full basis, and on a code list equals the
    dense operator applied to the state embedded with zeros outside the list."""
    rng = np.random.default_rng(4)
    for n in (1, 3, 5, 7):
        v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        codes = np.sort(rng.choice(1 << n, max(1, (1 << n) // 3), replace=False))
        w = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
        emb = np.zeros(1 << n, complex)
        emb[codes] = w
        for _ in range(6):
            k = int(rng.integers(1, n + 1))
            sites = rng.choice(n, k, replace=False)
            terms = [(int(s), str(rng.choice(['x', 'y', 'z']))) for s in sites]
            M = orc.pauli_dense(n, terms)
            assert abs(orc.pauli_expect(n, v, terms) - np.vdot(v, M @ v)) < 1e-10 * np.vdot(v, v).real
            assert abs(orc.pauli_expect(n, w, terms, codes) - np.vdot(emb, M @ emb)) < 1e-10 * np.vdot(w, w).real
    # single-qubit known answers: <g|z|g> = 1, <r|z|r> = -1, <+|x|+> = 1, y on (|g> + i|r>)/sqrt2 = 1
    s2 = 1 / np.sqrt(2)
    assert orc.pauli_expect(1, [1, 0], [(0, 'z')]) == 1
    assert orc.pauli_expect(1, [0, 1], [(0, 'z')]) == -1
    assert abs(orc.pauli_expect(1, [s2, s2], [(0, 'x')]) - 1) < 1e-15
    assert abs(orc.pauli_expect(1, [s2, 1j * s2], [(0, 'y')]) - 1) < 1e-15


def test_corr_matrix_oracle_known_answers():
    """C_ij of a product state is zero; of (|gg> + |rr>)/sqrt2 it is 1/4."""
    n = 3
    v = np.zeros(8, complex)
    v[0b101] = 1.0
    assert np.allclose(orc.corr_matrix(n, v), 0.0)
    v = np.zeros(4, complex)
    v[0] = v[3] = 1 / np.sqrt(2)
    C = orc.corr_matrix(2, v)
    assert np.allclose(C, [[0.25, 0.25], [0.25, 0.25]])


def test_spin_observables_vs_dense_operators():
    """The oracle's spin observables against dense Kronecker operators
    (σ+_i σ-_j, X, Y, Z) on random states, full basis and a code list
    (state embedded with zeros outside)."""
    rng = np.random.default_rng(11)
    for n in (1, 3, 5):
        v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        v /= np.linalg.norm(v)
        codes = np.sort(rng.choice(1 << n, max(2, (1 << n) // 2), replace=False))
        w = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
        w /= np.linalg.norm(w)
        emb = np.zeros(1 << n, complex)
        emb[codes] = w
        for vec, cd, full in ((v, None, v), (w, codes, emb)):
            ob = orc.spin_observables(n, vec, cd)
            for i in range(n):
                assert abs(ob['sx'][i] - np.vdot(full, orc.pauli_dense(n, [(i, 'x')]) @ full).real) < 1e-12
                assert abs(ob['sy'][i] - np.vdot(full, orc.pauli_dense(n, [(i, 'y')]) @ full).real) < 1e-12
                assert abs(ob['sz'][i] - np.vdot(full, orc.pauli_dense(n, [(i, 'z')]) @ full).real) < 1e-12
                for j in range(n):
                    ref = np.vdot(full, orc.sigma_pm_dense(n, i, j) @ full)
                    assert abs(ob['xy_corr'][i, j] - ref) < 1e-12


def test_entropy_vs_reduced_density_matrix():
    """Entropy from Schmidt values equals -Tr ρ_A ln ρ_A from the eigenvalues
    of ρ_A = M M†; Bell pair ln 2, product state 0."""
    rng = np.random.default_rng(12)
    for n, cut in ((4, 1), (5, 2), (6, 3)):
        v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        v /= np.linalg.norm(v)
        M = v.reshape(1 << cut, 1 << (n - cut))
        lam = np.linalg.eigvalsh(M @ M.conj().T)
        lam = lam[lam > 1e-15]
        assert abs(orc.entropy_at_cut(n, v, cut) + np.sum(lam * np.log(lam))) < 1e-12
        codes = np.arange(1 << n)
        assert abs(orc.entropy_at_cut(n, v, cut, codes) - orc.entropy_at_cut(n, v, cut)) < 1e-12
    bell = np.zeros(4, complex)
    bell[0] = bell[3] = 1 / np.sqrt(2)
    assert abs(orc.entropy_at_cut(2, bell, 1) - np.log(2)) < 1e-14
    prod = np.kron([0.6, 0.8], [1.0, 0.0]).astype(complex)
    assert abs(orc.entropy_at_cut(2, prod, 1)) < 1e-14


def test_oracle_loss_diagonal_closed_form():
    """γ convention of the oracle's non-Hermitian H (channels.py:58-77:
    −(i/2)·γ·1e-3/CONV per excited atom): with no drive and no couplings,
    |ψ|² of |r r g⟩ decays as exp(−(γ0 + γ1)·t·1e-3) (t in ns, γ in 1/µs),
    and a ground-state atom does not decay."""
    n = 3
    gamma = np.array([2.0, 5.0, 11.0])
    seg = [dict(t0=0.0, Om=np.zeros(n), phase=np.zeros(n), delta=np.zeros(n))]
    t = np.array([0.0, 100.0, 350.0])
    psi0 = np.zeros(1 << n, complex)
    psi0[0b110] = 1.0
    out = orc.propagate_segments(n, [], ['I1'], seg, psi0, t, gamma=gamma)
    norms = np.sum(np.abs(out) ** 2, axis=1)
    assert np.allclose(norms, np.exp(-(gamma[0] + gamma[1]) * t * 1e-3), rtol=1e-12, atol=0)
    g0 = np.zeros(1 << n, complex)
    g0[0] = 1.0
    outg = orc.propagate_segments(n, [], ['I1'], seg, g0, t, gamma=gamma)
    assert np.allclose(np.sum(np.abs(outg) ** 2, axis=1), 1.0, rtol=1e-14)
