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
