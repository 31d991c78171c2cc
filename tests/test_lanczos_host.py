"""Host-side Lanczos post-processing (no GPU): the genuine-Ritz-value filter
of the single-system solver and the native selection ``ryd_ritz_lowest`` of
the batched solver."""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _cw_case():
    d = json.load(open(os.path.join(HERE, 'golden', 'lanczos_cw_case.json')))
    return np.array(d['alpha']), np.array(d['beta']), d


def _native_lowest(alphas, offds, blast, tol):
    from energy_level_amd import _native as nat
    cnt = len(alphas)
    ld = max(len(a) for a in alphas)
    A = np.zeros((cnt, ld))
    O = np.zeros((cnt, ld))
    st = np.zeros(cnt, np.int32)
    for i, (a, o) in enumerate(zip(alphas, offds)):
        A[i, :len(a)] = a
        O[i, :len(o)] = o
        st[i] = len(a)
    E = np.zeros((cnt, 2))
    Bd = np.zeros((cnt, 2))
    Y = np.zeros((cnt, 2, ld))
    bl = np.ascontiguousarray(blast, float)
    nat.check(nat.lib().ryd_ritz_lowest(cnt, st.ctypes.data, A.ctypes.data, ld, O.ctypes.data, ld,
                                        bl.ctypes.data, tol, 1e-10, E.ctypes.data,
                                        Bd.ctypes.data, Y.ctypes.data, ld))
    return E, Bd, Y


def test_filter_keeps_lowest_value_that_cullum_willoughby_flags(pkg):
    """The recorded device run: its lowest Ritz value (the true E0) is within
    the Cullum-Willoughby tolerance of an eigenvalue of T-hat; both filters
    keep it and return (E0, E1) of the oracle's eigsh."""
    from energy_level_amd.lanczos import _tridiag, _genuine_lowest, _ritz_bounds
    from scipy.linalg import eigvalsh_tridiagonal
    a, b, d = _cw_case()
    th, Y = _tridiag(a, b)
    i0 = int(np.argmin(th))
    # the case is what it says: the lowest value would be labelled spurious
    that = eigvalsh_tridiagonal(a[1:], b[1:])
    tol = max(1e-10 * np.abs(th).max(), 100 * len(a) * np.finfo(float).eps * np.abs(th).max())
    assert np.min(np.abs(that - th[i0])) <= tol
    for bounds in (None, _ritz_bounds(Y, d['beta_last'])):
        picks = _genuine_lowest(a, b, th, 2, bounds, 1e-9)
        assert picks[0] == i0
        assert abs(th[picks[0]] - d['eigsh_E0']) < 1e-9
        assert abs(th[picks[1]] - d['eigsh_E1']) < 1e-9
    E, Bd, Yn = _native_lowest([a], [b], [d['beta_last']], 1e-9)
    assert abs(E[0, 0] - d['eigsh_E0']) < 1e-9 and abs(E[0, 1] - d['eigsh_E1']) < 1e-9
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    for q in range(2):
        assert abs(np.linalg.norm(Yn[0, q]) - 1) < 1e-12
        assert np.linalg.norm(T @ Yn[0, q] - E[0, q] * Yn[0, q]) < 1e-9 * np.abs(th).max()


def _recurrence(lam, u, steps):
    up = np.zeros_like(u)
    beta, A, B = 0.0, [], []
    u = u / np.linalg.norm(u)
    for _ in range(steps):
        w = lam * u - beta * up
        a = u @ w
        w -= a * u
        beta = np.linalg.norm(w)
        A.append(a)
        B.append(beta)
        up, u = u, w / beta
    return np.array(A), np.array(B)


def test_filters_merge_ghosts_and_agree(pkg):
    """Plain FP64 recurrences run past convergence carry ghost copies of the
    lowest eigenvalue.  Both filters return the two lowest distinct
    eigenvalues of the diagonal operator (a close pair and a separated one),
    and the native selection equals the Python one on recurrences of 40, 150
    and 300 steps, with unit eigenvectors of T and the bound |beta·y[k-1]|."""
    from energy_level_amd.lanczos import _tridiag, _genuine_lowest, _ritz_bounds
    alphas, offds, blast, refs, lams = [], [], [], [], []
    for seed in range(12):
        rng = np.random.default_rng(seed)
        lam = np.sort(rng.uniform(-1.0, 1.0, 1500))
        lam[:2] = [-1.5, -1.5 + (1e-6 if seed % 2 else 0.2)]
        a, B = _recurrence(lam, rng.standard_normal(lam.size), [40, 150, 300][seed % 3])
        th, Y = _tridiag(a, B[:-1])
        picks = _genuine_lowest(a, B[:-1], th, 2, _ritz_bounds(Y, B[-1]), 1e-9)
        alphas.append(a)
        offds.append(B[:-1])
        blast.append(B[-1])
        refs.append(th[picks])
        lams.append(lam)
    E, Bd, Yn = _native_lowest(alphas, offds, blast, 1e-9)
    for i in range(len(alphas)):
        assert np.allclose(E[i], refs[i], rtol=0, atol=1e-12)
        if len(alphas[i]) >= 150:
            assert np.allclose(E[i], lams[i][:2], rtol=0, atol=1e-10)
            assert np.sum(np.abs(_tridiag(alphas[i], offds[i])[0] - lams[i][0]) < 1e-8) >= 1
        k = len(alphas[i])
        for q in range(2):
            assert abs(Bd[i, q] - abs(blast[i] * Yn[i, q, k - 1])) <= 1e-14 * max(1.0, Bd[i, q])


def test_native_selection_with_batch_layout(pkg):
    """The layout BatchLanczos hands over (alpha [count][steps], the norms
    [count][steps+1] with the off-diagonal taken from column 1, uneven valid
    steps) gives each instance the selection of its own recurrence."""
    from energy_level_amd import _native as nat
    from energy_level_amd.lanczos import _tridiag, _genuine_lowest, _ritz_bounds
    steps, count = 60, 5
    alpha = np.zeros((count, steps))
    nrm = np.zeros((count, steps + 1))
    ks, refs = [], []
    for b in range(count):
        rng = np.random.default_rng(30 + b)
        lam = np.sort(rng.uniform(-2, 3, 400 + 100 * b))
        k = steps - 7 * b
        a, B = _recurrence(lam, rng.standard_normal(lam.size), k)
        alpha[b, :k] = a
        nrm[b, 0] = 1.0
        nrm[b, 1:k + 1] = B
        th, Y = _tridiag(a, B[:-1])
        refs.append(th[_genuine_lowest(a, B[:-1], th, 2, _ritz_bounds(Y, B[-1]), 1e-9)])
        ks.append(k)
    kk = np.asarray(ks, np.int32)
    o = np.ascontiguousarray(nrm[:, 1:])
    bl = np.ascontiguousarray(nrm[np.arange(count), kk])
    E = np.zeros((count, 2))
    Bd = np.zeros((count, 2))
    nat.check(nat.lib().ryd_ritz_lowest(count, kk.ctypes.data, alpha.ctypes.data, steps,
                                        o.ctypes.data, steps, bl.ctypes.data, 1e-9, 1e-10,
                                        E.ctypes.data, Bd.ctypes.data, None, 0))
    for b in range(count):
        assert np.allclose(E[b], refs[b], rtol=0, atol=1e-12), (b, E[b], refs[b])


def test_native_bound_is_the_best_copy_of_a_ghost_cluster(pkg):
    """Run past convergence, the lowest eigenvalue has a cluster of ghost
    copies in T whose bounds |beta·y[k-1]| differ by orders of magnitude (the
    newest copy is still forming).  The native selection reports the copy
    with the smallest bound: the value is certified (bound <= 1e-9) and the
    returned vector is a unit eigenvector of T for it, with that bound."""
    from energy_level_amd.lanczos import _tridiag
    for seed in range(6):
        rng = np.random.default_rng(100 + seed)
        lam = np.sort(rng.uniform(-1.0, 1.0, 800))
        lam[0] = -1.6
        a, B = _recurrence(lam, rng.standard_normal(lam.size), 300)
        th, Y = _tridiag(a, B[:-1])
        scale = max(1.0, np.max(np.abs(a)) + 2 * np.max(np.abs(B[:-1])))
        cl = np.nonzero(np.abs(th - th.min()) <= 1e-10 * scale)[0]
        assert cl.size > 1                            # ghost copies are present
        bounds = np.abs(B[-1] * Y[-1, cl])
        assert bounds.max() > 1e3 * bounds.min()      # and their bounds differ
        E, Bd, Yn = _native_lowest([a], [B[:-1]], [B[-1]], 1e-9)
        k = len(a)
        y = Yn[0, 0, :k]
        T = np.diag(a) + np.diag(B[:-1], 1) + np.diag(B[:-1], -1)
        assert abs(E[0, 0] - (-1.6)) <= 1e-10
        assert Bd[0, 0] <= 1e-9
        assert abs(np.linalg.norm(y) - 1.0) <= 1e-10
        assert np.linalg.norm(T @ y - E[0, 0] * y) <= 1e-9 * scale
        assert abs(Bd[0, 0] - abs(B[-1] * y[-1])) <= 1e-14 * max(1.0, Bd[0, 0])


def test_csr_instances_threaded_conversion_matches_one_at_a_time(pkg):
    """The threaded conversion of a mixed list (dense complex and real
    arrays, a scipy matrix, an existing instance) gives the same CSR as
    converting each matrix alone, in the list's order."""
    import scipy.sparse as sp
    batch = pkg.batch
    rng = np.random.default_rng(5)
    mats = []
    for k in range(20):
        dim = 16 << (k % 3)
        A = np.zeros((dim, dim), complex if k % 2 else float)
        r, c = rng.integers(0, dim, 3 * dim), rng.integers(0, dim, 3 * dim)
        A[r, c] = rng.standard_normal(r.size) + (1j * rng.standard_normal(r.size) if k % 2 else 0)
        A = A + A.conj().T
        if k == 7:
            A = sp.csr_matrix(A)
        mats.append(A)
    mats.append(batch.CsrInstance(mats[0]))
    got = batch.csr_instances(mats)
    assert len(got) == len(mats) and got[-1] is mats[-1]
    for m, g in zip(mats, got):
        ref = m if isinstance(m, batch.CsrInstance) else batch.CsrInstance(m)
        assert (g.dim, g.n, g.complex) == (ref.dim, ref.n, ref.complex)
        assert np.array_equal(g.H.indptr, ref.H.indptr)
        assert np.array_equal(g.H.indices, ref.H.indices)
        assert np.array_equal(g.H.data, ref.H.data)
