"""Source-state spectra on the device: the reorthogonalised Lanczos
recurrence against the oracle's (two full Gram-Schmidt sweeps, as
spectrum.py:34-103), its quadrature against the exact spectral decomposition,
and the want='spectrum' dispatch against dense diagonalisation."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu


def _model(pkg, n, c6=10.0, De=1.2, phase=None):
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(c6), terms=('I1',))
    Om = np.ones(n) if phase is None else np.exp(1j * phase * np.arange(n))
    return calib, Om, -De * np.ones(n)


def _H(calib, Om, delta, codes=None):
    n = len(calib['atoms'])
    return orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes)


def _cheb_moments(H, phi, c, r, count):
    """<phi|T_p((H - c)/r)|phi>, p < count, by the Chebyshev recurrence on H."""
    t0 = np.asarray(phi, complex)
    t1 = (H @ t0 - c * t0) / r
    mu = [np.vdot(phi, t0).real, np.vdot(phi, t1).real]
    for _ in range(2, count):
        t0, t1 = t1, 2 * (H @ t1 - c * t1) / r - t0
        mu.append(np.vdot(phi, t1).real)
    return np.asarray(mu)


def _check_tridiagonal(pkg, out, H, phi, a, b, k):
    """The device tridiagonal against the oracle's and against H itself.

    Late Lanczos coefficients are ill-conditioned functions of the start
    vector once Ritz values converge: the oracle run from phi·(1 + 1e-16·g)
    differs from the unperturbed one by 4e-9 (relative) in alpha_28 and by
    4e-6 by step 200 for this Hamiltonian.  So the coefficients agree to
    1e-11 only over the first steps; what the recurrence fixes for every k is
    checked on well-conditioned quantities: the k-node Gauss quadrature
    integrates every polynomial of degree <= 2k-1 exactly (all Chebyshev
    moments of H on phi), the lowest Ritz values, and the broadened
    spectrum."""
    m = 20
    assert np.allclose(out['alphas'][:m], a[:m], rtol=1e-11, atol=1e-11)
    assert np.allclose(out['betas'][:m], b[:m], rtol=1e-11, atol=1e-11)
    nodes, w = pkg.quadrature(out['alphas'], out['betas'], out['norm2'])
    ev = np.linalg.eigvalsh(H.toarray())
    c, r = 0.5 * (ev[-1] + ev[0]), 0.51 * (ev[-1] - ev[0])
    mu = _cheb_moments(H, phi, c, r, 2 * k)
    quad = np.polynomial.chebyshev.chebvander((nodes - c) / r, 2 * k - 1).T @ w
    assert np.max(np.abs(quad - mu)) < 1e-11 * out['norm2']
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert np.allclose(nodes[:5], np.linalg.eigvalsh(T)[:5], rtol=1e-12, atol=1e-11)
    # a Lorentzian of width 0.5 MHz is resolved by the degree-(2k-1) exact
    # quadrature (narrower ones depend on unconverged nodes, which rounding moves)
    om = np.linspace(ev[0] - 1.0, ev[-1] + 1.0, 3000)
    S = pkg.lorentz_spectrum(out['alphas'], out['betas'], out['norm2'], om, 0.5)
    S_ref = pkg.lorentz_spectrum(a, b, out['norm2'], om, 0.5)
    assert np.max(np.abs(S - S_ref)) < 1e-10 * np.max(S_ref)


@pytest.mark.parametrize('k', [40, 200])
def test_reortho_tridiagonal_matches_oracle(pkg, dev, k):
    import torch
    n = 9
    calib, Om, delta = _model(pkg, n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    H = _H(calib, Om, delta)
    phi = np.random.default_rng(7).normal(size=1 << n)
    print('spectrum reortho n=%d k=%d' % (n, k), flush=True)
    out = pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), n_krylov=k)
    a, b, norm2 = orc.lanczos(lambda v: H @ v, phi, k)
    assert out['reortho'] and out['note'] == ''
    assert out['n'] == k == len(out['alphas']) and len(out['betas']) == k - 1
    assert not out['converged']
    assert np.isclose(out['norm2'], norm2, rtol=1e-14)
    _check_tridiagonal(pkg, out, H, phi, a, b, k)
    assert out['ortho_loss'] < 1e-12
    P.close()


def test_reortho_complex_source_and_phases(pkg, dev):
    import torch
    n = 8
    calib, Om, delta = _model(pkg, n, phase=0.7)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    assert not P.is_real
    H = _H(calib, Om, delta)
    rng = np.random.default_rng(3)
    phi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    out = pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), n_krylov=120)
    a, b, _ = orc.lanczos(lambda v: H @ v, phi, 120)
    _check_tridiagonal(pkg, out, H, phi, a, b, 120)
    P.close()


def test_full_krylov_quadrature_is_the_spectral_decomposition(pkg, dev):
    """k = dim: the nodes are the eigenvalues of H and the weights the
    squared overlaps |<e|phi>|^2 summed over each eigenvalue."""
    import torch
    n = 6
    calib, Om, delta = _model(pkg, n, c6=2.0, De=1.1)   # eigenvalue gaps >= 8e-4
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    Hd = _H(calib, Om, delta).toarray()
    w, V = np.linalg.eigh(Hd)
    phi = np.random.default_rng(11).normal(size=1 << n)
    out = pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), n_krylov=500)
    assert out['converged'] and out['n'] == 1 << n
    nodes, weights = pkg.quadrature(out['alphas'], out['betas'], out['norm2'])
    ov = np.abs(V.T @ phi) ** 2
    assert np.allclose(nodes, w, atol=1e-9)
    assert np.allclose(weights, ov, atol=1e-9)
    assert np.isclose(weights.sum(), phi @ phi, rtol=1e-12)
    P.close()


def test_breakdown_in_an_invariant_subspace(pkg, dev):
    """Om = 0 leaves H diagonal: phi on three basis states of distinct energy
    spans an invariant subspace, and the recurrence ends after three steps
    with converged = True (the reference's beta <= 1e-12)."""
    import torch
    n = 7
    calib, _, delta = _model(pkg, n)
    Om = np.zeros(n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    Hd = _H(calib, Om, delta).toarray()
    hd = np.diag(Hd)
    picks = [0, 1, 3]
    assert len(set(np.round(hd[picks], 9))) == 3
    phi = np.zeros(1 << n)
    phi[picks] = [1.0, 0.5, 0.25]
    out = pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), n_krylov=30)
    a, b, _ = orc.lanczos(lambda v: Hd @ v, phi, 30)
    assert out['converged'] and out['n'] == len(a) == 3
    assert len(out['betas']) == len(b) == 2
    assert np.allclose(out['alphas'], a, atol=1e-12) and np.allclose(out['betas'], b, atol=1e-12)
    nodes, weights = pkg.quadrature(out['alphas'], out['betas'], out['norm2'])
    order = np.argsort(hd[picks])
    assert np.allclose(nodes, hd[picks][order], atol=1e-12)
    assert np.allclose(weights, np.array([1.0, 0.25, 0.0625])[order], atol=1e-12)
    P.close()


def test_over_budget_downgrades_with_a_note(pkg, dev):
    import torch
    n = 10
    calib, Om, delta = _model(pkg, n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    H = _H(calib, Om, delta)
    phi = np.random.default_rng(5).normal(size=1 << n)
    # 1024 · 100 · 32 bytes = 3.1 MiB > 1 MiB
    out = pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), n_krylov=100,
                                   memory_budget_mb=1.0)
    assert not out['reortho'] and 'budget' in out['note'] and out['ortho_loss'] == 0.0
    a, b, _ = orc.lanczos(lambda v: H @ v, phi, 100)
    assert out['n'] == 100
    assert np.allclose(out['alphas'][:12], a[:12], rtol=1e-9, atol=1e-9)
    with pytest.raises(ValueError):
        pkg.spectrum_lanczos_hip(P, torch.as_tensor(phi, device=dev), 10, memory_budget_mb=0)
    with pytest.raises(ValueError):
        pkg.spectrum_lanczos_hip(P, torch.zeros(1 << n, device=dev, dtype=torch.float64), 10)
    P.close()


def test_source_spectrum_report(pkg, dev):
    import torch
    n = 8
    calib, Om, delta = _model(pkg, n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), device=dev)
    H = _H(calib, Om, delta)
    phi = np.random.default_rng(2).normal(size=1 << n)
    out = pkg.source_spectrum_hip(P, torch.as_tensor(phi, device=dev), shift=-1.0, eta=0.1,
                                  n_krylov=80, n_omega=4000)
    a, b, norm2 = orc.lanczos(lambda v: H @ v, phi, 80)
    for key in ('omega', 'spectral', 'evals', 'spectrum_weights', 'source_norm', 'eta', 'shift',
                'krylov', 'error_report'):
        assert key in out, key
    assert out['krylov']['n'] == 80 and out['krylov']['reortho']
    assert np.isclose(out['spectrum_weights'].sum(), norm2, rtol=1e-12)
    Tn = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert np.allclose(out['evals'][:5], np.linalg.eigvalsh(Tn)[:5] + 1.0, atol=1e-9)
    # the default window reaches 4 eta beyond the nodes: most of the weight
    ratio = out['error_report']['integral_ratio']
    assert 0.9 < ratio < 1.01
    # S(omega) at one point equals the pole sum of the quadrature
    om = out['omega'][1234]
    direct = np.sum(out['spectrum_weights'] * (0.1 / np.pi) /
                    ((om - out['evals']) ** 2 + 0.01))
    assert np.isclose(out['spectral'][1234], direct, rtol=1e-8)
    P.close()


def test_solve_equilibrium_spectrum_ground_source(pkg, dev):
    n = 8
    calib, Om, delta = _model(pkg, n)
    w_site = np.linspace(0.5, 1.5, n)
    sysd = dict(calib=calib, Om=Om, delta=delta, device=dev,
                source={'kind': 'site_weights', 'weights': w_site, 'subtract_mean': True},
                spectrum={'eta': 0.05, 'n_krylov': 60})
    print('spectrum dispatch ground source n=%d' % n, flush=True)
    out = pkg.solve_equilibrium_hip(sysd, 'spectrum', backend='hip-lanczos')
    Hd = _H(calib, Om, delta).toarray()
    e, V = np.linalg.eigh(Hd)
    g = V[:, 0]
    d = orc.bits_of(n, np.arange(1 << n)).astype(float).T @ w_site
    d = d - g @ (d * g)
    phi = d * g
    a, b, norm2 = orc.lanczos(lambda v: Hd @ v, phi, 60)
    assert out['want'] == 'spectrum' and out['gap'] is None
    assert abs(out['E0'] - e[0]) < 1e-9 * abs(e[0])
    assert np.isclose(out['source_norm'], norm2, rtol=1e-7)
    Tn = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert np.allclose(out['evals'][:3], np.linalg.eigvalsh(Tn)[:3] - e[0], atol=1e-7)
    # every polynomial moment of degree < 2k: set by phi, which the device
    # ground state fixes to its residual tolerance
    c, r = 0.5 * (e[-1] + e[0]), 0.51 * (e[-1] - e[0])
    mu = _cheb_moments(Hd, phi, c, r, 120)
    quad = np.polynomial.chebyshev.chebvander((out['evals'] + e[0] - c) / r, 119).T @ \
        out['spectrum_weights']
    assert np.max(np.abs(quad - mu)) < 1e-6 * norm2
    rep = out['error_report']
    assert rep['source_kind'] == 'site_weights' and rep['dim_basis'] == 1 << n
    assert rep['sector'] is None and 'subtracted_mean' in rep and 'sum_weights' in rep
    q = out['data_schema']['quantities']
    assert q['spectral']['quantity'] == 'spectral_function'
    assert q['spectral']['axes'] == ['energy'] and q['E0']['unit'] == 'MHz'
    assert out['data_schema']['axes']['energy']['field'] == 'energies'
    assert out['energies'].shape == out['spectral'].shape == (800,)


def test_solve_equilibrium_spectrum_sector_and_psi0(pkg, dev):
    """Om = 0 conserves the excitation number: the 3-excitation sector with
    an explicit initial state and a sparse source in that basis."""
    n = 8
    calib, _, delta = _model(pkg, n)
    Om = np.zeros(n)
    codes = orc.number_sector_basis(n, 3)
    dim = len(codes)
    rng = np.random.default_rng(4)
    psi = rng.normal(size=dim)
    psi /= np.linalg.norm(psi)
    rows = np.arange(dim)
    src = {'kind': 'sparse', 'rows': rows, 'cols': rows[::-1].copy(),
           'values': np.linspace(1, 2, dim), 'dim': dim}
    sysd = dict(calib=calib, Om=Om, delta=delta, device=dev, charge_sector=3, psi0=psi,
                source=src, spectrum={'n_krylov': dim})
    out = pkg.solve_equilibrium_hip(sysd, 'spectrum', backend='hip-lanczos')
    Hs = _H(calib, Om, delta, codes=codes).toarray()
    import scipy.sparse as sp
    O = sp.coo_matrix((src['values'], (rows, src['cols'])), shape=(dim, dim)).toarray()
    phi = O @ psi
    assert out['E0'] is None and out['dim_sub'] == dim
    assert np.array_equal(out['sector_idx'], codes)
    assert out['error_report']['sector'] == 3 and out['error_report']['dim_basis'] == dim
    assert out['krylov']['converged']
    # Om = 0 makes H diagonal with repeated values: the recurrence ends once
    # the distinct values phi touches are exhausted, and the quadrature then
    # reproduces every moment <phi|H^p|phi>
    nodes, wts = out['evals'], out['spectrum_weights']
    hd = np.diag(Hs)
    for p in range(6):
        assert np.isclose(np.sum(wts * nodes ** p), np.sum(phi ** 2 * hd ** p), rtol=1e-9), p
    with pytest.raises(ValueError):
        pkg.solve_equilibrium_hip(dict(sysd, Om=np.ones(n)), 'spectrum', backend='hip-lanczos')
    with pytest.raises(ValueError):
        pkg.solve_equilibrium_hip(dict(calib=calib, Om=Om, delta=delta, device=dev),
                                  'spectrum', backend='hip-lanczos')


def _flipflop_case():
    """Four sites with nearest-neighbour flip-flops J = 1 (I2), two
    excitations (dim 6), source D = n_0 + n_1 on the ground state: the exact
    spectrum is |<m|D|0>|^2 at E_m - E_0 (the case of
    checks/manybody_source_spectrum_selfcheck.py:31-52)."""
    n = 4
    calib = {'atoms': [{'id': i, 'pos': [float(i), 0.0, 0.0]} for i in range(n)],
             'pairs': [{'i': i, 'j': i + 1, 'J': 1.0} for i in range(n - 1)]}
    idx = orc.number_sector_basis(n, 2)
    bits = orc.bits_of(n, idx)
    diag = bits[0].astype(float) + bits[1].astype(float)
    Hd = orc.build_H_sparse(n, np.zeros(n), np.zeros(n), calib['pairs'], ['I2'],
                            codes=idx).toarray()
    w, V = np.linalg.eigh(Hd)
    phi = diag * V[:, 0]
    amp = np.abs(V.T @ phi) ** 2
    keep = amp > 1e-12
    nodes = w[keep] - w[0]
    order = np.argsort(nodes)
    return calib, idx, diag, phi, nodes[order], amp[keep][order]


def test_source_spectrum_flipflop_sector_exact(pkg, dev):
    calib, idx, diag, phi, nodes_x, weights_x = _flipflop_case()
    base = dict(calib=calib, Om=None, delta=None, terms=['I2'], device=dev, charge_sector=2,
                spectrum={'eta': 0.02, 'n_krylov': 6, 'n_omega': 1201,
                          'omega_min': -0.4, 'omega_max': 4.0})

    def run(**kw):
        return pkg.solve_equilibrium_hip(dict(base, **kw), 'spectrum', backend='hip-lanczos')

    res = run(source={'kind': 'diagonal', 'values': diag.tolist()})
    nodes = np.asarray(res['evals'])
    weights = np.asarray(res['spectrum_weights'])
    assert nodes.shape == nodes_x.shape and np.max(np.abs(nodes - nodes_x)) < 1e-9
    assert np.max(np.abs(weights - weights_x)) < 1e-9
    assert abs(weights.sum() - res['source_norm']) < 1e-10
    omega, s = np.asarray(res['energies']), np.asarray(res['spectral'])
    peaks = [omega[i] for i in range(1, len(s) - 1)
             if s[i] > s[i - 1] and s[i] >= s[i + 1] and s[i] > 0.05 * s.max()]
    assert peaks and all(min(abs(p - x) for x in nodes_x) < 2 * res['eta'] for p in peaks)
    q = res['data_schema']['quantities']
    assert {'spectral', 'source_norm'} <= set(q) and 'spectrum_weights' not in q
    assert res['krylov']['n'] >= 1 and res['error_report']['integral_ratio'] is not None
    assert res['dim_sub'] == 6
    # a vector source equal to D|0>
    res_v = run(source={'kind': 'vector', 'values': phi.tolist()})
    assert np.max(np.abs(np.asarray(res_v['spectrum_weights']) - weights_x)) < 1e-9
    # the initial state as the source (source omitted), given on all 2^4 codes
    full = np.zeros(16, complex)
    full[idx] = phi
    res_p = run(psi0=full)
    wv = np.asarray(res_v['spectrum_weights']) / res_v['source_norm']
    wp = np.asarray(res_p['spectrum_weights']) / res_p['source_norm']
    assert np.max(np.abs(wp - wv)) < 1e-9 and res_p['error_report']['source_kind'] == 'psi0'
    # a number-conserving operator on all 2^4 codes: no projected leak
    vals = [float(((s_ >> 3) & 1) + ((s_ >> 2) & 1)) for s_ in range(16)]
    res_f = run(source={'kind': 'full_sparse', 'rows': list(range(16)), 'cols': list(range(16)),
                        'values': vals, 'dim': 16})
    assert abs(res_f['error_report']['projected_leak']) < 1e-12
    assert np.max(np.abs(np.asarray(res_f['spectrum_weights']) - weights_x)) < 1e-9
    # truncation: two Krylov steps give two nodes
    res_k = run(source={'kind': 'diagonal', 'values': diag.tolist()},
                spectrum={'eta': 0.2, 'n_krylov': 2, 'n_omega': 201})
    assert len(res_k['evals']) == 2 and res_k['krylov']['n'] == 2
    for bad in ({'kind': 'nope'}, {'kind': 'diagonal', 'values': [1.0]},
                {'kind': 'sparse', 'rows': [0], 'cols': [1], 'values': [1.0]}):
        with pytest.raises(ValueError):
            run(source=bad)
