"""hamiltonian_linear_operator (matrix_free.py:51-62) on the device against
the oracle's CSR of the same calib, including the loss diagonal (gamma != 0)
and a restricted basis given in an arbitrary order."""
import json
import os

import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _chain(n):
    return {'atoms': [{}] * n, 'terms': ['I1'],
            'pairs': [{'i': i, 'j': j, 'V': 12.0 / (j - i) ** 6} for i in range(n)
                      for j in range(i + 1, n)]}


@pytest.mark.parametrize('lossy', [False, True])
def test_linear_operator_full_and_restricted(pkg, dev, lossy):
    n = 9
    calib = _chain(n)
    rng = np.random.default_rng(9)
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    gamma = rng.uniform(0.5, 5.0, n) if lossy else None
    op = pkg.hamiltonian_linear_operator_hip(calib, Om, delta, gamma, device=dev)
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], gamma=gamma)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    ref = H @ v
    assert op.shape == H.shape and op.dtype == np.complex128
    assert np.max(np.abs(op.matvec(v) - ref)) <= 1e-12 * np.max(np.abs(ref))
    # a restricted basis in a shuffled order (vectors in that order)
    codes = rng.permutation(1 << n)[:200]
    opr = pkg.hamiltonian_linear_operator_hip(calib, Om, delta, gamma, basis_indices=codes,
                                              device=dev)
    Hr = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes, gamma=gamma)
    w = rng.normal(size=codes.size) + 1j * rng.normal(size=codes.size)
    refr = Hr @ w
    assert np.max(np.abs(opr.matvec(w) - refr)) <= 1e-12 * np.max(np.abs(refr))


def test_linear_operator_exchange_channels(pkg, dev):
    """The reference's kagome N=12 XY calib (J channels)."""
    calib = json.load(open(os.path.join(HERE, 'golden', 'kagome_n12.json')))
    n = len(calib['atoms'])
    delta = np.array([a.get('delta_mhz', 0.0) for a in calib['atoms']])
    op = pkg.hamiltonian_linear_operator_hip(calib, np.zeros(n), delta, device=dev)
    H = orc.build_H_sparse(n, np.zeros(n), delta, calib['pairs'], ['I2'])
    v = np.random.default_rng(12).normal(size=1 << n) + 0j
    ref = H @ v
    assert np.max(np.abs(op.matvec(v) - ref)) <= 1e-12 * np.max(np.abs(ref))
