"""Independent-set bases of blockade graphs enumerated and ranked on the
device (RYD_BASIS_GRAPH) against the oracle's depth-first enumeration, the
transfer-matrix counts of square lattices, the oracle's projected H·v and
dense diagonalisation."""
import numpy as np
import pytest

from oracle import hv_ref, rydberg_oracle as orc
from test_basis_host import _grid_edges, _transfer_count

pytestmark = pytest.mark.gpu


def _random_edges(n, p, seed):
    rng = np.random.default_rng(seed)
    return [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]


def _codes(pkg, dev, n, edges):
    return pkg.independent_states_hip(n, edges, device=dev).cpu().numpy()


@pytest.mark.parametrize('n,edges', [
    (1, []), (2, [(0, 1)]), (3, []), (7, [(0, 1), (1, 2), (5, 6)]),
    (16, _random_edges(16, 0.2, 1)), (17, _random_edges(17, 0.2, 2)),
    (20, [(i, i + 1) for i in range(19)]), (20, [(i, j) for i in range(20) for j in range(i + 1, 20)]),
    (18, []), (22, _random_edges(22, 0.15, 3)), (24, _random_edges(24, 0.3, 4)),
    (20, _grid_edges(5, 4)), (25, _grid_edges(5, 5, diag=True)),
    (33, [(i, i + 1) for i in range(32)] + [(i, i + 2) for i in range(31)]),
])
def test_enumeration_matches_oracle(pkg, dev, n, edges):
    got = _codes(pkg, dev, n, edges)
    ref = hv_ref.indep_codes(n, edges)
    assert got.dtype == np.int64 and np.array_equal(got, ref)


def test_counts_and_zero_sites(pkg, dev):
    assert pkg.independent_set_count_hip(0, [], device=dev) == 1
    assert pkg.independent_states_hip(0, [], device=dev).cpu().tolist() == [0]
    assert pkg.independent_set_count_hip(36, _grid_edges(6, 6), device=dev) == 5598861
    assert pkg.independent_set_count_hip(30, [(i, i + 1) for i in range(29)], device=dev) == \
        pkg.fibonacci(32)
    with pytest.raises(ValueError):
        pkg.independent_states_hip(10, [(0, 1)], max_states=10, device=dev)
    with pytest.raises(ValueError):
        pkg.independent_states_hip(61, [], device=dev)
    with pytest.raises(ValueError):
        pkg.independent_states_hip(4, [(0, 4)], device=dev)


@pytest.mark.parametrize('budget', [None, 0])
def test_rank_round_trip_and_membership(pkg, dev, budget, monkeypatch):
    """rank(code) over the basis is 0..D-1; codes outside it rank -1; both the
    table ranks and the binary-search fallback (prefix budget 0)."""
    import torch
    if budget is not None:
        monkeypatch.setenv('RYD_GRAPH_PREFIX_MAX_BYTES', str(budget))
    n = 26
    edges = _random_edges(n, 0.12, 9)
    P = pkg.graph_plan(n, edges, device=dev)
    assert P.rank_mode == ('search' if budget == 0 else 'graph')
    codes = pkg.sector_codes(P)
    r = pkg.sector_rank(P, codes)
    assert torch.equal(r, torch.arange(P.dim, device=dev))
    rng = np.random.default_rng(5)
    probe = rng.integers(0, 1 << n, 200000, dtype=np.int64)
    ref = hv_ref.indep_codes(n, edges)
    pos = np.searchsorted(ref, probe)
    inside = (pos < len(ref)) & (ref[np.minimum(pos, len(ref) - 1)] == probe)
    want = np.where(inside, pos, -1)
    got = pkg.sector_rank(P, torch.as_tensor(probe, device=dev)).cpu().numpy()
    assert np.array_equal(got, want)
    ranks = torch.as_tensor(rng.integers(-3, P.dim + 3, 1000), device=dev)
    un = pkg.sector_codes(P, ranks=ranks).cpu().numpy()
    rr = ranks.cpu().numpy()
    assert np.array_equal(un, np.where((rr >= 0) & (rr < P.dim), ref[np.clip(rr, 0, P.dim - 1)], -1))


def _lattice_calib(w, h, c6=20.0, cutoff=2.9):
    pos = [(c, r) for r in range(h) for c in range(w)]
    pairs = []
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = np.hypot(pos[i][0] - pos[j][0], pos[i][1] - pos[j][1])
            if d <= cutoff:
                pairs.append({'i': i, 'j': j, 'V': c6 / d ** 6})
    return {'atoms': [{}] * len(pos), 'pairs': pairs, 'terms': ['I1']}


@pytest.mark.parametrize('budget', [None, 0])
def test_hv_6x6_lattice_vs_c_oracle(pkg, dev, budget, monkeypatch):
    """FP64 projected H·v on the 6x6 nearest-neighbour blockade basis
    (D = 5,598,861) against the oracle's binary-search matvec."""
    import torch
    if budget is not None:
        monkeypatch.setenv('RYD_GRAPH_PREFIX_MAX_BYTES', str(budget))
    w = h = 6
    n = w * h
    calib = _lattice_calib(w, h)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    assert edges == sorted(_grid_edges(w, h))
    rng = np.random.default_rng(11)
    Om = rng.uniform(0.5, 1.5, n)
    delta = rng.uniform(-2, 2, n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), basis='graph', edges=edges, device=dev)
    assert P.dim == _transfer_count(w, h)
    codes = hv_ref.indep_codes(n, edges)
    x = rng.normal(size=P.dim)
    d = hv_ref.diag(n, delta, calib['pairs'], codes)
    ref = hv_ref.hv_codes(n, Om, d, codes, x)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    y32 = P.hv(torch.as_tensor(x.astype(np.float32), device=dev)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 2e-5 * np.max(np.abs(ref))


def test_complex_drive_and_pair_flips_on_graph_basis(pkg, dev):
    """Complex Ω plus J/A channels projected onto an independent-set basis,
    against the oracle's channel matvec on the same codes."""
    import torch
    n = 12
    rng = np.random.default_rng(4)
    edges = _random_edges(n, 0.2, 8)
    pairs = [{'i': i, 'j': j, 'V': float(rng.uniform(0, 2)), 'J': complex(rng.normal(), rng.normal()),
              'A': complex(rng.normal(), rng.normal()) / 2} for i in range(n) for j in range(i + 1, n)]
    terms = ['I1', 'I2', 'I3']
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': terms}
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 6.3, n))
    delta = rng.uniform(-1, 1, n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta, None, terms), basis='graph', edges=edges,
                 device=dev)
    codes = np.asarray(orc.independent_states(n, edges))
    x = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    ref = orc.channel_matvec(n, Om, delta, pairs, terms, x, codes=codes)
    y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_ground_state_graph_basis_vs_dense(pkg, dev):
    n = 14
    calib = _lattice_calib(7, 2, c6=30.0)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    Om = np.ones(n)
    delta = 1.3 * np.ones(n)
    codes = np.asarray(orc.independent_states(n, edges))
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes).toarray()
    ev, vec = np.linalg.eigh(H)
    out = pkg.hip_lanczos_ground_state(calib, Om, delta, blockade_edges=edges, device=dev,
                                       return_state=True)
    assert out['converged'] and out['dim'] == len(codes)
    assert abs(out['E0'] - ev[0]) <= 1e-9 * max(1.0, abs(ev[0]))
    assert abs(out['gap'] - (ev[1] - ev[0])) <= 1e-8
    assert abs(abs(np.vdot(out['vec'], vec[:, 0])) - 1.0) <= 1e-8
    occ = orc.site_occupations(n, vec[:, 0], codes=codes)
    assert np.max(np.abs(out['n_atom'] - occ)) <= 1e-8


def test_dispatch_projected_plans(pkg, dev):
    """A device-built plan (states on the device) and a reference-style plan
    (host state list + blockade_edges) give the dense projected E0; the
    returned basis indices are the plan's states."""
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    cons = {'kind': 'blockade', 'threshold_mhz': 10.0}
    Om, delta = np.ones(n), 1.1 * np.ones(n)
    plan = pkg.build_plan_hip(calib, Om, delta, constraint=cons, device=dev)
    edges = [tuple(e) for e in plan['blockade_edges']]
    codes = np.asarray(orc.independent_states(n, edges))
    assert plan['projected'] and plan['dim'] == len(codes)
    assert np.array_equal(plan['states'].cpu().numpy(), codes)
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes).toarray()
    e0 = np.linalg.eigvalsh(H)[0]
    ref_plan = dict(plan, states=[int(c) for c in codes])
    for p in (plan, ref_plan):
        sysd = {'calib': calib, 'Om': Om, 'delta': delta, 'plan': p, 'device': dev}
        r = pkg.solve_equilibrium_hip(sysd, 'gs', backend='hip-lanczos', return_state=True)
        assert abs(r['E0'] - e0) <= 1e-9 * max(1.0, abs(e0))
        assert np.array_equal(np.asarray(r['basis']['indices']), codes)
        assert r['basis']['dim'] == len(codes)


def test_lattice_6x7_properties(pkg, dev):
    """6x7 nearest-neighbour blockade lattice, D = 69,050,253 (transfer
    matrix): strictly ascending codes, every code independent, ranks
    0..D-1, and a Hermitian projected H (<y, Hx> = <Hy, x>)."""
    import torch
    w, h = 6, 7
    n = w * h
    calib = _lattice_calib(w, h)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    rng = np.random.default_rng(2)
    P = pkg.Plan(pkg.from_calib(calib, rng.uniform(0.5, 1.5, n), rng.uniform(-1, 1, n)),
                 basis='graph', edges=edges, device=dev)
    assert P.dim == _transfer_count(w, h) == 69050253
    codes = pkg.sector_codes(P)
    assert bool((codes[1:] > codes[:-1]).all())
    bad = torch.zeros_like(codes)
    for i, j in edges:
        bad |= (codes >> (n - 1 - i)) & (codes >> (n - 1 - j)) & 1
    assert int(bad.max()) == 0
    assert torch.equal(pkg.sector_rank(P, codes), torch.arange(P.dim, device=dev))
    del bad
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn(P.dim, dtype=torch.float64, device=dev, generator=g)
    yv = torch.randn(P.dim, dtype=torch.float64, device=dev, generator=g)
    a = torch.dot(yv, P.hv(x)).item()
    b = torch.dot(P.hv(yv), x).item()
    assert abs(a - b) <= 1e-10 * abs(a)


@pytest.mark.parametrize('doubling', [False, True])
def test_kpm_on_graph_basis_vs_reference_estimator(pkg, dev, doubling):
    """KPM on a projected plan (the reference compiles plan['states'] into the
    CSR, eq_route.py:73-74): same probes, same moments as the oracle's
    estimator on the projected H."""
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    Om, delta = np.ones(n), -0.8 * np.ones(n)
    P = pkg.Plan(pkg.from_calib(calib, Om, delta), basis='graph', edges=edges, device=dev)
    codes = np.asarray(orc.independent_states(n, edges))
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes)
    emin, emax, c, r = orc.kpm_bounds(H, margin=0.01)
    mref, _, _ = orc.kpm_moments(H, 40, 3, seed=5, c=c, r=r)
    mu, _, _ = pkg.kpm_moments_hip(plan=P, num=40, nvec=3, seed=5, c=c, r=r, doubling=doubling)
    assert np.max(np.abs(mu - mref)) < 1e-10
    sysd = {'calib': calib, 'Om': Om, 'delta': delta, 'device': dev,
            'plan': pkg.build_plan_hip(calib, Om, delta, constraint={'kind': 'blockade',
                                                                      'threshold_mhz': 10.0},
                                       device=dev)}
    th = pkg.solve_equilibrium_hip(sysd, 'thermal', backend='hip-kpm', betas=[0.5], num=200,
                                   nvec=40)
    ev = np.linalg.eigvalsh(H.toarray())
    w = np.exp(-0.5 * (ev - ev.min()))
    e_exact = float(np.sum(ev * w) / np.sum(w))
    assert abs(th['E'][0] - e_exact) < 0.05 * max(1.0, abs(e_exact))


@pytest.mark.parametrize('how', ['edges', 'states', 'plan'])
def test_segments_on_graph_basis_vs_exact(pkg, dev, how):
    """Piecewise-constant evolution restricted to the independent sets of a
    4x3 lattice (engine_ed.py:191-193, :288) against the oracle's exact
    exponentials on the same projected H; psi0 = |g...g>, then a full 2^n
    vector restricted to the basis."""
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    cons = {'kind': 'blockade', 'threshold_mhz': 10.0}
    edges = pkg.blockade_edges(calib, cons)
    codes = np.asarray(orc.independent_states(n, edges))
    rng = np.random.default_rng(2)
    segs = [dict(t0=t0, Om=rng.uniform(0.5, 3.0, n), phase=rng.uniform(0, 2 * np.pi, n),
                 delta=rng.uniform(-2, 2, n)) for t0 in (0.0, 60.0)]
    t = np.array([0.0, 45.0, 120.0])
    kw = {'edges': dict(blockade_edges=edges), 'states': dict(basis_indices=codes[::-1].copy()),
          'plan': dict(plan=pkg.build_plan_hip(calib, constraint=cons, device=dev))}[how]
    out = pkg.propagate_segments_hip(calib, segs, t, device=dev, return_states=True, **kw)
    psi0 = np.zeros(len(codes), complex)
    psi0[0] = 1.0
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0, t, codes=codes)
    got = out['states'][:, ::-1] if how == 'states' else out['states']
    assert np.max(np.abs(got - ref)) < 1e-10
    assert np.array_equal(np.sort(out['basis_states']), codes)
    pops = np.array([orc.site_occupations(n, v, codes=codes) for v in ref])
    assert np.allclose(out['pops_atom'], pops, atol=1e-10)
    full = np.zeros(1 << n, complex)
    full[codes[3]] = 2.0j
    o2 = pkg.propagate_segments_hip(calib, segs, t[:2], psi0=full, blockade_edges=edges,
                                    device=dev)
    p2 = np.zeros(len(codes), complex)
    p2[3] = 1.0j
    r2 = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, p2, t[:2], codes=codes)[-1]
    assert np.max(np.abs(o2['psi'] - r2)) < 1e-10
    bad = np.zeros(1 << n, complex)
    bad[(1 << (n - 1)) | (1 << (n - 2))] = 1.0       # sites 0 and 1 both excited: blockaded
    with pytest.raises(ValueError):
        pkg.propagate_segments_hip(calib, segs, t[:2], psi0=bad, blockade_edges=edges, device=dev)


def test_generic_evolution_matches_fused_full_basis(pkg, dev):
    """The generic slice path (here a code-list plan holding all 2^n codes)
    reproduces the fused full-basis slice kernels."""
    n = 10
    calib = _lattice_calib(5, 2, c6=15.0)
    rng = np.random.default_rng(8)
    segs = [dict(t0=t0, Om=rng.uniform(0.5, 2.0, n), phase=rng.uniform(0, 6.3, n),
                 delta=rng.uniform(-1, 1, n)) for t0 in (0.0, 30.0)]
    t = np.array([0.0, 50.0])
    full = pkg.propagate_segments_hip(calib, segs, t, device=dev)
    codes = np.arange(1 << n, dtype=np.int64)
    gen = pkg.propagate_segments_hip(calib, segs, t, device=dev, basis_indices=codes)
    assert np.max(np.abs(full['psi'] - gen['psi'])) < 1e-12


@pytest.mark.parametrize('w,h,diag', [(5, 4, False), (6, 4, True), (5, 5, False)])
def test_split_graph_hv_vs_c_oracle(pkg, dev, w, h, diag, monkeypatch):
    """The two-pass graph H·v (top high bits in a second pass), forced on at
    small sizes, against the oracle's projected H·v, and against the one-pass
    kernel on the same plan."""
    import torch
    n = w * h
    calib = _lattice_calib(w, h)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 2.0 if diag else 10.0})
    rng = np.random.default_rng(n)
    Om, delta = rng.uniform(0.5, 1.5, n), rng.uniform(-2, 2, n)
    codes = hv_ref.indep_codes(n, edges)
    x = rng.normal(size=len(codes))
    ref = hv_ref.hv_codes(n, Om, hv_ref.diag(n, delta, calib['pairs'], codes), codes, x)
    out = {}
    for key, mind in (('split', '0'), ('one', str(1 << 40))):
        monkeypatch.setenv('RYD_GRAPH_SPLIT_MIN_DIM', mind)
        P = pkg.Plan(pkg.from_calib(calib, Om, delta), basis='graph', edges=edges, device=dev)
        assert P.dim == len(codes)
        out[key] = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
        P.close()
    for key, y in out.items():
        assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref)), key
    assert np.max(np.abs(out['split'] - out['one'])) <= 1e-13 * np.max(np.abs(ref))


def test_split_graph_ground_state_vs_eigsh(pkg, dev, monkeypatch):
    """Lanczos through the split H·v (three-term steps with z and fused
    reductions in the second pass) on the 5x5 lattice basis (D = 55,447)."""
    monkeypatch.setenv('RYD_GRAPH_SPLIT_MIN_DIM', '0')
    n = 25
    calib = _lattice_calib(5, 5)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    Om, delta = np.ones(n), 1.2 * np.ones(n)
    codes = hv_ref.indep_codes(n, edges)
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'], codes=codes)
    ref = orc.ground_state_eigsh(H, k=2)
    out = pkg.hip_lanczos_ground_state(calib, Om, delta, blockade_edges=edges, device=dev)
    assert out['converged'] and out['dim'] == len(codes)
    assert abs(out['E0'] - ref['E0']) <= 1e-9 * max(1.0, abs(ref['E0']))
    assert abs(out['gap'] - ref['gap']) <= 1e-8


@pytest.mark.parametrize('wide', ['0', '1'])
def test_graph_hv_row_table_widths(pkg, dev, wide, monkeypatch):
    """16-bit (default below 65535 rows per segment) and 32-bit row tables of
    the graph kernel against the C oracle, one-pass and split."""
    import torch
    monkeypatch.setenv('RYD_GRAPH_WIDE_ROWS', wide)
    n = 30
    calib = _lattice_calib(6, 5)
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    rng = np.random.default_rng(7)
    Om, delta = rng.uniform(0.5, 1.5, n), rng.uniform(-2, 2, n)
    codes = hv_ref.indep_codes(n, edges)
    x = rng.normal(size=len(codes))
    ref = hv_ref.hv_codes(n, Om, hv_ref.diag(n, delta, calib['pairs'], codes), codes, x)
    for mind in ('0', str(1 << 40)):
        monkeypatch.setenv('RYD_GRAPH_SPLIT_MIN_DIM', mind)
        P = pkg.Plan(pkg.from_calib(calib, Om, delta), basis='graph', edges=edges, device=dev)
        assert P.rank_mode == 'graph'
        y = P.hv(torch.as_tensor(x, device=dev)).cpu().numpy()
        P.close()
        assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref)), mind
