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


# ---- blockade graphs -------------------------------------------------------
def _grid_edges(w, h, diag=False):
    """Row-major w x h lattice: nearest neighbours (and diagonals)."""
    idx = lambda r, c: r * w + c
    out = []
    for r in range(h):
        for c in range(w):
            if c + 1 < w:
                out.append((idx(r, c), idx(r, c + 1)))
            if r + 1 < h:
                out.append((idx(r, c), idx(r + 1, c)))
                if diag and c + 1 < w:
                    out.append((idx(r, c), idx(r + 1, c + 1)))
                if diag and c > 0:
                    out.append((idx(r, c), idx(r + 1, c - 1)))
    return out


def _transfer_count(w, h, diag=False):
    """Independent sets of a w x h lattice by the row transfer matrix (an
    independent method: rows are independent sets of a path, consecutive
    rows may not share a column, nor a diagonal when diag)."""
    rows = [r for r in range(1 << w) if not (r & (r >> 1))]
    cnt = {r: 1 for r in rows}
    for _ in range(h - 1):
        new = dict.fromkeys(rows, 0)
        for r in rows:
            for q, c in cnt.items():
                if r & q or (diag and (r & (q << 1) or r & (q >> 1))):
                    continue
                new[r] += c
        cnt = new
    return sum(cnt.values())


def test_c_oracle_independent_sets_vs_brute_force():
    from oracle import hv_ref, rydberg_oracle as orc
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 6, 9, 12):
        for _ in range(4):
            edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.3]
            a = hv_ref.indep_codes(n, edges)
            assert np.array_equal(a, np.asarray(orc.independent_states_bruteforce(n, edges)))
            assert np.array_equal(a, np.asarray(orc.independent_states(n, edges)))


@pytest.mark.parametrize('w,h,diag', [(4, 4, False), (5, 5, False), (5, 5, True), (6, 6, True)])
def test_c_oracle_lattice_counts_vs_transfer_matrix(w, h, diag):
    from oracle import hv_ref
    assert hv_ref.indep_count(w * h, _grid_edges(w, h, diag)) == _transfer_count(w, h, diag)


def test_c_oracle_chain_count_is_fibonacci(pkg):
    from oracle import hv_ref
    for n in (1, 2, 10, 24):
        assert hv_ref.indep_count(n, pkg.chain_edges(n)) == pkg.fibonacci(n + 2)


def test_c_oracle_projected_hv_vs_channel_oracle():
    from oracle import hv_ref, rydberg_oracle as orc
    rng = np.random.default_rng(3)
    n = 10
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.25]
    codes = hv_ref.indep_codes(n, edges)
    pairs = [{'i': i, 'j': j, 'V': float(rng.uniform(0, 3))} for i in range(n) for j in range(i + 1, n)]
    Om = rng.uniform(0.2, 1.5, n)
    delta = rng.uniform(-2, 2, n)
    x = rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, Om, delta, pairs, ['I1'], x, codes=codes)
    d = hv_ref.diag(n, delta, pairs, codes)
    y = hv_ref.hv_codes(n, Om, d, codes, x)
    assert np.max(np.abs(y - ref.real)) <= 1e-12 * np.max(np.abs(ref))


def test_validate_edges_and_plan_edges_mirror_reference(pkg):
    assert pkg.validate_edges(4, [(2, 1), (1, 2), [0, 3]]) == [(0, 3), (1, 2)]
    for bad in ([(0, 0)], [(0, 4)], [(0, 1.5)], [(True, 1)], [(0, 1, 2)]):
        with pytest.raises(ValueError):
            pkg.validate_edges(4, bad)
    calib = {'atoms': [{}] * 4, 'pairs': [{'i': 0, 'j': 1, 'V': 50.0}, {'i': 1, 'j': 3, 'V': -60.0},
                                          {'i': 0, 'j': 2, 'V': 5.0}]}
    assert pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10}) == [(0, 1), (1, 3)]
    assert pkg.blockade_edges(calib, {'kind': 'pxp', 'edges': [(2, 3)]}) == [(2, 3)]
    with pytest.raises(ValueError):
        pkg.blockade_edges(calib, {'kind': 'blockade'})             # no threshold, no edges
    with pytest.raises(ValueError):
        pkg.blockade_edges(calib, {'kind': 'blockade', 'edges': 'all'})
    with pytest.raises(ValueError):
        pkg.build_plan_hip(calib, constraint={'kind': 'ring'})


def test_normalize_pauli_strings_mirrors_reference(pkg):
    out = pkg.normalize_pauli_strings([[[0, 'x'], [2, 'σy']], {'label': 'zz', 'terms': [[1, 'Z'], [3, 'sz']]}], 4)
    assert out == [{'label': 'x@0 y@2', 'terms': [(0, 'x'), (2, 'y')]},
                   {'label': 'zz', 'terms': [(1, 'z'), (3, 'z')]}]
    assert pkg.normalize_pauli_strings([]) == []
    assert pkg.normalize_pauli_strings(None) is None
    for bad in ([[]], [[[0, 'x'], [0, 'z']]], [[[5, 'x']]], [[[0, 'w']]], [[[True, 'x']]],
                [{'label': 'a'}], 'x@0'):
        with pytest.raises(ValueError):
            pkg.normalize_pauli_strings(bad, 4)


def test_state_vector_mirrors_reference(pkg):
    """Initial states as state.py:112-210: patterns, index, superposition,
    product dict, list of local states, 2^N vector; projected weight outside
    the basis is an error."""
    n = 3
    v = pkg.state_vector_np('grg', n)
    assert v[0b010] == 1 and np.count_nonzero(v) == 1
    assert np.array_equal(pkg.state_vector_np(5, n), np.eye(8)[5])
    s = pkg.state_vector_np(('sup', [('ggg', 1.0), ('rrr', 1.0)]), n)
    assert np.allclose(s[[0, 7]], 1 / np.sqrt(2)) and np.count_nonzero(s) == 2
    p = pkg.state_vector_np({'kind': 'product', 'amplitudes': [[1, 1], 'g', [0, 1]]}, n)
    assert np.allclose(p[[0b001, 0b101]], 1 / np.sqrt(2))
    q = pkg.state_vector_np(['r', 'g', 'g'], n)
    assert q[0b100] == 1
    full = np.arange(8, dtype=complex)
    assert np.allclose(pkg.state_vector_np(full, n), full / np.linalg.norm(full))
    codes = [0b000, 0b001, 0b100, 0b101]
    assert np.allclose(pkg.state_vector_np('rgr', n, codes), [0, 0, 0, 1])
    with pytest.raises(ValueError):
        pkg.state_vector_np('rrg', n, codes)          # outside the basis
    with pytest.raises(ValueError):
        pkg.state_vector_np('gx', n)
    with pytest.raises(ValueError):
        pkg.state_vector_np('ground', n)
    assert pkg.product_label(3, 0b101) == '|rgr>'


def test_from_positions_mirrors_reference(pkg):
    """Pairs from a point cloud as calib_build.from_positions: scalar kernel is
    pure J, tuples are (V, J, A), periodic images summed per pair, images
    past the cutoff dropped and counted, default terms ('I2',)."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    c = pkg.from_positions(pos, lambda r: 1.0 / r)
    assert c['terms'] == ['I2'] and len(c['pairs']) == 3
    assert c['pairs'][0] == {'i': 0, 'j': 1, 'V': 0.0, 'J': 1.0, 'A': 0.0}
    c = pkg.from_positions(pos, lambda r: (2.0 / r, 0.0, 0.5), terms=('I1', 'I3'))
    assert c['pairs'][1]['V'] == 1.0 and c['pairs'][1]['A'] == 0.5
    # periodic images: Lx = 3, Ly = 4, one image ring; sum of 1/r^2 by hand
    L = (3.0, 4.0)
    c = pkg.from_positions(pos, lambda r: {'V': 1.0 / r ** 2}, periods_um=L, images=1)
    want = 0.0
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            d = pos[1] - pos[0] + np.array([ix * L[0], iy * L[1], 0.0])
            want += 1.0 / float(d @ d)
    assert abs(c['pairs'][0]['V'] - want) < 1e-14
    assert c['notes']['images'] == 1 and c['notes']['periods_um'] == [3.0, 4.0]
    c = pkg.from_positions(pos, lambda r: 1.0, cutoff_um=1.5)
    assert [(p['i'], p['j']) for p in c['pairs']] == [(0, 1)]
    assert c['notes']['n_dropped_by_cutoff'] == 2
    for bad in (dict(periods_um=(1.0,)), dict(periods_um=(1.0, -1.0)),
                dict(periods_um=(1.0, 1.0), images=-1)):
        with pytest.raises(ValueError):
            pkg.from_positions(pos, lambda r: 1.0, **bad)
    with pytest.raises(ValueError):
        pkg.from_positions(np.zeros((0, 3)), lambda r: 1.0)
    with pytest.raises(ValueError):
        pkg.from_positions(np.zeros((3, 2)), lambda r: 1.0)


def _vortex_bonds(n, winding):
    """<σ+_a σ−_b> of the product state ⊗_k (|g> + e^{iφ_k}|r>)/√2 with
    φ_k = 2π·winding·k/n: (1/4)·e^{i(φ_b − φ_a)} for a ≠ b."""
    phi = 2 * np.pi * winding * np.arange(n) / n
    return 0.25 * np.exp(1j * (phi[None, :] - phi[:, None]))


@pytest.mark.parametrize('w', [1, -1, 2])
def test_loop_observables_count_vortex_windings(pkg, w):
    n = 6
    xy = _vortex_bonds(n, w)
    ring = list(range(n))
    ph, wd = pkg.loop_observables(xy, [ring, ring[::-1], [0, 3]])
    # every bond's phase step 2πw/6 lies in (−π, π] for |w| ≤ 2, so the
    # wrapped sum around the ring is exactly 2πw, reversed −2πw; a
    # two-site loop goes there and back (zero)
    assert np.allclose(ph, [2 * np.pi * w, -2 * np.pi * w, 0.0], atol=1e-12)
    assert np.allclose(wd, [w, -w, 0.0], atol=1e-12)
    xy[2, 3] = 1e-15
    ph, wd = pkg.loop_observables(xy, [ring, [0, 1, 2]])
    assert np.isnan(ph[0]) and np.isnan(wd[0]) and np.isfinite(ph[1])
    assert pkg.loop_observables(xy, None) == ([], [])
    for bad in ([0], [0, n], [-1, 2]):
        with pytest.raises(ValueError):
            pkg.loop_observables(xy, [bad])


def test_link_loop_observables_orientation_and_gaps(pkg):
    n = 4
    xys = [_vortex_bonds(n, 1), _vortex_bonds(n, -1)]
    xys[1][3, 0] = 0.0                 # no phase on the closing bond at t1
    pairs = [(0, 1), (2, 1), (2, 3), (3, 0), (1, 0)]
    series = np.array([[xy[a, b] for a, b in pairs] for xy in xys])
    loops = [[0, 1, 2, 3], [3, 2, 1, 0], [0, 2, 3]]
    ph, wd = pkg.link_loop_observables(series, pairs, loops)
    assert ph.shape == wd.shape == (2, 3)
    # (2, 1) walked as 1 -> 2 takes the conjugate; (0, 1) is listed before
    # (1, 0), so the bond 1 -> 0 is conj(<σ+_0 σ−_1>)
    assert np.allclose(wd[0, :2], [1, -1], atol=1e-12)
    assert np.isnan(wd[1, :2]).all()            # a bond of zero modulus
    assert np.isnan(ph[:, 2]).all()             # bond 0 -> 2 not measured
    # the same loops from the full matrix agree where every bond is measured
    ph0, _ = pkg.loop_observables(xys[0], loops[:2])
    assert np.allclose(ph[0, :2], ph0, atol=1e-12)
    one, _ = pkg.link_loop_observables(series[0], pairs, loops[:1])
    assert one.shape == (1, 1) and np.isclose(one[0, 0], 2 * np.pi)
    e1, e2 = pkg.link_loop_observables(series, pairs, None)
    assert e1.shape == (0, 0) and e2.shape == (0, 0)
    with pytest.raises(ValueError):
        pkg.link_loop_observables(series, pairs, [[1]])
