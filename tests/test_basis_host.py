"""Number sectors, the number-conservation check and J/A model building on
the host (no GPU), pinned by brute force and the reference's refusals."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize('n,k', [(0, 0), (1, 0), (1, 1), (6, 3), (10, 4), (12, 12), (16, 5)])
def test_number_sector_basis_brute_force(pkg, n, k):
    ref = np.array([c for c in range(1 << n) if bin(c).count('1') == k], np.int64)
    assert np.array_equal(pkg.number_sector_basis(n, k), ref)


def test_number_sector_basis_refusals(pkg):
    for bad in ((4, 5), (4, -1), (64, 1), (4, 1.5), (4, True)):
        with pytest.raises(ValueError):
            pkg.number_sector_basis(*bad)


def test_check_number_conserving(pkg):
    calib = {'atoms': [{}] * 3, 'pairs': [{'i': 0, 'j': 1, 'J': 1.0, 'A': 0.5}],
             'terms': ['I2', 'I3']}
    with pytest.raises(ValueError):
        pkg.check_number_conserving(calib, Om=[0, 1, 0])
    with pytest.raises(ValueError):
        pkg.check_number_conserving(calib)                 # active pair creation
    pkg.check_number_conserving(calib, terms=['I2'])      # exchange only: fine


def test_from_calib_pair_channels(pkg):
    calib = json.load(open(os.path.join(HERE, 'golden', 'kagome_n12.json')))
    m = pkg.from_calib(calib, None, None)
    assert m.n == 12 and len(m.pair_i) == 66
    assert m.has_pair_flips and np.all(m.pair_v == 0) and m.is_real
    mc = pkg.from_calib({'atoms': [{}] * 2, 'pairs': [{'i': 0, 'j': 1, 'J': 1j, 'A': 0.3}],
                         'terms': ['I2', 'I3']})
    assert not mc.is_real and mc.pair_A[0] == 0.3
    m1 = pkg.from_calib({'atoms': [{}] * 2, 'pairs': [{'i': 0, 'j': 1, 'V': 2.0, 'J': 1.0}],
                         'terms': ['I1']})
    assert not m1.has_pair_flips and m1.pair_v[0] == 2.0
