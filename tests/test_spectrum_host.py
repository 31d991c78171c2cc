"""Host logic of the source-spectrum path: the source forms of
``_resolve_source`` (eq_route.py:163-252), the Lorentzian/quadrature
post-processing against closed forms, and the data descriptors of
``describe_result`` (results.py:11-137).  No GPU."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc


def _psi(dim, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=dim) + 1j * rng.normal(size=dim)
    return v / np.linalg.norm(v)


def test_vector_psi0_and_diagonal_sources(pkg):
    n = 5
    psi = _psi(1 << n)
    v = np.arange(1 << n, dtype=float)
    phi, info = pkg.resolve_source({'kind': 'vector', 'values': v}, psi, n)
    assert np.array_equal(phi, v) and info == {'kind': 'vector', 'dim': 32}
    phi, info = pkg.resolve_source({'kind': 'state'}, psi, n)
    assert np.array_equal(phi, psi) and info['kind'] == 'psi0'
    phi, _ = pkg.resolve_source({'kind': 'diagonal', 'values': v}, psi, n)
    assert np.allclose(phi, v * psi)
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'diagonal', 'values': v[:-1]}, psi, n)
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'vector', 'values': v}, psi[:-1], n)
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'nope'}, psi, n)
    with pytest.raises(ValueError):
        pkg.resolve_source(['diagonal'], psi, n)


def test_site_weights_source_and_connected_part(pkg):
    n = 6
    codes = orc.number_sector_basis(n, 2)
    psi = _psi(len(codes), 1)
    w = np.array([0.3, -1.0, 2.0, 0.5, 0.0, 1.25])
    phi, info = pkg.resolve_source({'kind': 'site_weights', 'weights': w}, psi, n, codes)
    d = orc.bits_of(n, codes).astype(float).T @ w
    assert np.allclose(phi, d * psi, atol=1e-15)
    assert info['sum_weights'] == pytest.approx(w.sum())
    assert info['weight_sq'] == pytest.approx((w * w).sum())
    # (O - <O>)|0> is orthogonal to |0>: the elastic peak is removed
    phi_c, info_c = pkg.resolve_source({'kind': 'weights', 'weights': w, 'subtract_mean': True},
                                       psi, n, codes)
    assert abs(np.vdot(psi, phi_c)) < 1e-14
    assert info_c['subtracted_mean'] == pytest.approx(float(np.vdot(psi, d * psi).real))
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'site_weights', 'weights': w[:-1]}, psi, n, codes)


def test_sparse_sources_and_projected_leak(pkg):
    n = 4
    codes = orc.number_sector_basis(n, 2)
    dim = len(codes)
    psi = _psi(dim, 2)
    rows = np.arange(dim)
    phi, _ = pkg.resolve_source({'kind': 'sparse', 'rows': rows, 'cols': rows,
                                 'values': 2.0 * np.ones(dim), 'dim': dim}, psi, n, codes)
    assert np.allclose(phi, 2.0 * psi)
    with pytest.raises(ValueError):          # shape (2, 2) is not the basis
        pkg.resolve_source({'kind': 'sparse', 'rows': [0], 'cols': [1], 'values': [1.0]},
                           psi, n, codes)
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'sparse', 'rows': [], 'cols': [], 'values': []}, psi, n, codes)
    with pytest.raises(ValueError):
        pkg.resolve_source({'kind': 'sparse', 'rows': [0, 1], 'cols': [1], 'values': [1.0]},
                           psi, n, codes)
    # X on site 0 changes the excitation number: all of O|0> leaves the sector
    full = 1 << n
    xr = np.arange(full)
    xc = xr ^ (1 << (n - 1))
    phi, info = pkg.resolve_source({'kind': 'full_sparse', 'rows': xr, 'cols': xc,
                                    'values': np.ones(full), 'dim': full}, psi, n, codes)
    assert np.allclose(phi, 0.0) and info['projected_leak'] == pytest.approx(1.0)
    # n_0 conserves it: no leak, and O|0> equals the diagonal form
    n0 = ((xr >> (n - 1)) & 1).astype(float)
    phi, info = pkg.resolve_source({'kind': 'full_sparse', 'rows': xr, 'cols': xr,
                                    'values': n0, 'dim': full}, psi, n, codes)
    assert info['projected_leak'] == pytest.approx(0.0, abs=1e-15)
    assert np.allclose(phi, n0[codes] * psi)
    with pytest.raises(ValueError):
        big = 1 << 17
        pkg.resolve_source({'kind': 'full_sparse', 'rows': [0], 'cols': [0], 'values': [1.0],
                            'dim': big}, _psi(big), 17)


def test_quadrature_and_lorentzian_closed_forms(pkg):
    rng = np.random.default_rng(9)
    a = rng.normal(size=7)
    b = np.abs(rng.normal(size=6)) + 0.1
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    w, V = np.linalg.eigh(T)
    nodes, weights = pkg.quadrature(a, b, 2.5, shift=0.75)
    assert np.allclose(nodes, w - 0.75) and np.allclose(weights, 2.5 * V[0] ** 2)
    om = np.linspace(-4, 4, 9)
    S = pkg.lorentz_spectrum(a, b, 2.5, om, 0.3, shift=0.75)
    # -1/pi Im <e0|(z - T)^-1|e0> from the resolvent itself
    ref = [-2.5 / np.pi * np.linalg.inv((x + 0.75 + 0.3j) * np.eye(7) - T)[0, 0].imag for x in om]
    assert np.allclose(S, ref, rtol=1e-12)
    assert pkg.quadrature([], [], 1.0)[0].size == 0
    assert np.array_equal(pkg.lorentz_spectrum([], [], 1.0, om, 0.3), np.zeros_like(om))


def test_describe_result_descriptors(pkg):
    spec = dict(want='spectrum', n_sites=4, E0=-1.0, gap=None, evals=np.zeros(3),
                spectrum_weights=np.ones(3), energies=np.linspace(0, 1, 5),
                spectral=np.ones(5), source_norm=1.0)
    calib = {'atoms': [{'id': i, 'pos': [float(i), 0.0]} for i in range(4)]}
    out = pkg.describe_result(spec, calib)
    q = out['data_schema']['quantities']
    assert out['data_schema']['version'] == 1
    assert q['spectral'] == dict(quantity='spectral_function', axes=['energy'], unit='1/MHz')
    assert q['source_norm']['quantity'] == 'source_state_weight'
    assert q['evals'] == dict(quantity='eigenenergy', axes=['eigenstate'], unit='MHz')
    assert 'gap' not in q and 'spectrum_weights' not in q
    assert q['E0']['summary'] == {'mode': 'scalar'}
    assert out['sites'][2] == {'site_id': 2, 'pos': [2.0, 0.0]}
    assert out['data_schema']['axes']['site']['size'] == 4
    th = pkg.describe_result(dict(betas=np.ones(2), E=np.ones(2), C=np.ones(2), F=np.ones(2),
                                  n_atom=np.ones((2, 3)), gap=float('nan')))
    tq = th['data_schema']['quantities']
    assert tq['E']['axes'] == ['beta'] and tq['n_atom']['axes'] == ['beta', 'site']
    assert tq['gap']['missing_values'] == 'nonfinite'
    with pytest.raises(ValueError):
        pkg.describe_result(dict(pops=np.ones((3, 2)), states=[{'index': 0}]))
    ev = pkg.describe_result(dict(t=np.ones(3), pops=np.ones((3, 1)), states=[{'index': 0}],
                                  loss=np.zeros(3)))
    assert ev['pops_kind'] == 'basis_probability'
    assert ev['data_schema']['quantities']['loss']['alias_of'] == 'norm_drift'
    assert np.array_equal(ev['norm_drift'], np.zeros(3))
