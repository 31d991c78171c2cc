"""Host-side parts of the HIP dispatch: backend names and the spectrum
post-processing, pinned by the mathematics they restate (no GPU)."""
import numpy as np
import pytest


def test_backend_names_and_wants(pkg):
    assert pkg.normalize_backend('HIP-Lanczos', 'gs') == 'hip-lanczos'
    assert pkg.normalize_backend('kpm-hip', 'dos') == 'hip-kpm'
    with pytest.raises(ValueError):
        pkg.normalize_backend('hip-kpm', 'gs')
    with pytest.raises(ValueError):
        pkg.normalize_backend('gpu-lanczos', 'gs')
    assert set(pkg.EQ_BACKENDS_HIP) == {'hip-lanczos', 'hip-kpm', 'hip-batch'}


def _tri(seed, k):
    rng = np.random.default_rng(seed)
    return rng.normal(size=k), np.abs(rng.normal(size=k - 1)) + 0.1


def test_quadrature_is_the_spectral_measure_of_e0():
    a, b = _tri(0, 12)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    w, V = np.linalg.eigh(T)
    nodes, weights = pkg_quadrature()(a, b, 2.5, shift=0.3)
    assert np.allclose(nodes, w - 0.3)
    assert np.isclose(weights.sum(), 2.5)
    # moments of the measure are the moments of T seen from e0
    for p in range(6):
        assert np.isclose((weights * (nodes + 0.3) ** p).sum(),
                          2.5 * np.linalg.matrix_power(T, p)[0, 0])


def test_lorentz_continued_fraction_equals_resolvent():
    a, b = _tri(1, 9)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    om = np.linspace(-3, 3, 41)
    eta, shift, norm2 = 0.05, 0.2, 1.7
    S = pkg_lorentz()(a, b, norm2, om, eta, shift)
    ref = []
    for w in om:
        G = np.linalg.inv((w + shift + 1j * eta) * np.eye(len(a)) - T)
        ref.append(-norm2 * G[0, 0].imag / np.pi)
    assert np.allclose(S, ref, rtol=1e-10, atol=1e-12)


def pkg_quadrature():
    import __graft_entry__ as ge
    return ge.load_package().quadrature


def pkg_lorentz():
    import __graft_entry__ as ge
    return ge.load_package().lorentz_spectrum


def test_segment_rules_known_cases(pkg):
    segs = [{'t0': 0.0, 'Om': [1.0, 2.0], 'phase': [0.0, np.pi / 2], 'delta': [0.5, -0.5]},
            {'t0': 4.0, 'Om': [3.0, 0.0], 'delta': [1.5, 0.5]}]
    assert pkg.segment_intervals(1.0, 6.0, segs) == [(1.0, 4.0), (4.0, 6.0)]
    assert pkg.segment_intervals(4.0, 6.0, segs) == [(4.0, 6.0)]      # a start at a is no cut
    om, de = pkg.segment_drive(2.0, segs, 2)
    assert np.allclose(om, [1.0, 2.0j]) and np.allclose(de, [0.5, -0.5])
    om, de = pkg.segment_drive(4.0, segs, 2)                          # latest t0 <= t
    assert np.allclose(om, [3.0, 0.0]) and np.allclose(de, [1.5, 0.5])
    om, de = pkg.segment_drive(2.0, segs, 2, interp='midpoint')       # mean with the next
    assert np.allclose(om, [2.0, 1.0j]) and np.allclose(de, [1.0, 0.0])
    om, de = pkg.segment_drive(5.0, segs, 2, interp='midpoint')       # last one: its own
    assert np.allclose(om, [3.0, 0.0])
    with pytest.raises(ValueError):
        pkg.segment_drive(1.0, [{'t0': 0.0, 'Om': [1.0]}], 2)
    with pytest.raises(ValueError):
        pkg.segment_drive(1.0, segs, 2, interp='cubic')
