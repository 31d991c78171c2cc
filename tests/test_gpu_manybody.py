"""evolve_manybody on the device (the result contract of engine_ed.py:153-372)
against the oracle's exact propagation and its restatements of the
observables (spin observables, entropies, C_ij, Pauli strings)."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc
from test_gpu_graph import _lattice_calib

pytestmark = pytest.mark.gpu


def _segments(n, rng):
    return [dict(t0=t0, Om=rng.uniform(0.5, 3.0, n), phase=rng.uniform(0, 2 * np.pi, n),
                 delta=rng.uniform(-2, 2, n)) for t0 in (0.0, 55.0)]


def _check(out, ref_states, n, codes, cut, pauli=None, link=None, tol=1e-9):
    t = out['t']
    for k in range(len(t)):
        v = ref_states[k]
        p = np.abs(v) ** 2
        assert abs(out['norms'][k] - p.sum()) < tol
        assert np.max(np.abs(out['pops_atom'][k] - orc.site_occupations(n, v, codes))) < tol
        assert abs(out['S'][k] - orc.entropy_at_cut(n, v, cut, codes)) < 1e-8
        ob = orc.spin_observables(n, v, codes)
        O = out['observables']
        for key in ('sx', 'sy', 'sz'):
            assert np.max(np.abs(O[key][k] - ob[key])) < tol, key
        assert np.max(np.abs(O['sm'][k] - ob['sm'])) < tol
        if pauli:
            for q, s in enumerate(pauli):
                assert abs(O['pauli_strings'][k, q] - orc.pauli_expect(n, v, s['terms'], codes)) < tol
        if link:
            for q, (a, b) in enumerate(link):
                assert abs(O['link_corr'][k, q] - ob['xy_corr'][a, b]) < tol
    last = spin = orc.spin_observables(n, ref_states[-1], codes)
    assert np.max(np.abs(out['observables']['xy_corr'] - last['xy_corr'])) < tol
    assert np.max(np.abs(out['corr'] - orc.corr_matrix(n, ref_states[-1], codes))) < tol
    assert np.allclose(out['loss'], 1 - out['norms'])


def test_evolve_manybody_full_basis(pkg, dev):
    n = 8
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    rng = np.random.default_rng(5)
    segs = _segments(n, rng)
    t = np.array([0.0, 30.0, 90.0])
    pauli = [[[0, 'x'], [1, 'y']], [[2, 'z']]]
    link = [(0, 1), (3, 5)]
    out = pkg.evolve_manybody_hip(calib, psi0='grgggggg', t_grid=t, segments=segs,
                                  pauli_strings=pauli, link_pairs=link, record_psi=True,
                                  diag=True, device=dev)
    psi0 = orc.propagate_segments(n, calib['pairs'], ['I1'], segs,
                                  pkg.state_vector_np('grgggggg', n), t)
    assert np.max(np.abs(out['psi_amp'] - psi0[-1])) < 1e-9
    norm = pkg.normalize_pauli_strings(pauli, n)
    _check(out, psi0, n, None, n // 2, pauli=norm, link=link)
    assert out['pops'].shape == (len(t), 1 << n) and out['states'][5]['label'] == '|gggggrgr>'
    assert out['backend'] == 'hip-ed' and out['method'] == 'manybody'
    assert out['basis'] == dict(kind='full', bit_order='site0-msb', n_atoms=n, dim=1 << n,
                                full_dim=1 << n)
    assert out['error_report']['max_norm_dev'] < 1e-9 and out['diag']['dim'] == 1 << n
    q = out['data_schema']['quantities']
    assert q['pops']['quantity'] == 'basis_probability' and q['pops']['axes'] == ['time', 'basis']
    assert q['loss']['alias_of'] == 'norm_drift' and np.array_equal(out['norm_drift'], out['loss'])
    assert q['observables.sx']['axes'] == ['time', 'site'] and len(out['sites']) == n


@pytest.mark.parametrize('via', ['edges', 'states'])
def test_evolve_manybody_projected_plan(pkg, dev, via):
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    cons = {'kind': 'blockade', 'threshold_mhz': 10.0}
    plan = pkg.build_plan_hip(calib, constraint=cons, device=dev)
    if via == 'states':
        plan = dict(plan, states=[int(c) for c in plan['states'].cpu().numpy()])
    codes = plan['states'].cpu().numpy() if via == 'edges' else np.asarray(plan['states'])
    rng = np.random.default_rng(6)
    segs = _segments(n, rng)
    t = np.array([0.0, 40.0, 100.0])
    out = pkg.evolve_manybody_hip(calib, t_grid=t, segments=segs, plan=plan, device=dev,
                                  entropy_cut=5)
    psi0 = pkg.state_vector_np(None, n, codes)
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0, t, codes=codes)
    _check(out, ref, n, codes, 5)
    assert out['backend'] == 'hip-ed-plan-blockade'
    assert np.array_equal(np.asarray(out['basis_states']), codes)
    assert out['pops'].shape == (len(t), n)                      # per-atom marginals


def test_evolve_manybody_ground_start_static_drive(pkg, dev):
    """psi0='ground' at a static drive stays stationary (up to phase)."""
    n = 8
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    Om, delta = np.ones(n), -1.0 * np.ones(n)
    out = pkg.evolve_manybody_hip(calib, psi0='ground', t_grid=[0.0, 20.0, 50.0], Om=Om,
                                  delta=delta, device=dev)
    assert np.max(np.abs(out['pops_atom'] - out['pops_atom'][0])) < 1e-8
    assert np.max(np.abs(out['S'] - out['S'][0])) < 1e-7
    with pytest.raises(ValueError):
        pkg.evolve_manybody_hip(calib, t_grid=[0.0, 1.0], entropy_cut=n, device=dev)


def test_evolve_manybody_loops_vortex(pkg, dev):
    """loops: a vortex product state winds once at t0 (closed form); later
    times follow the oracle's propagated state."""
    n = 4
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    phi = 2 * np.pi * np.arange(n) / n
    local = [np.array([1.0, np.exp(1j * p)]) / np.sqrt(2) for p in phi]
    psi0 = local[0]
    for v in local[1:]:
        psi0 = np.kron(psi0, v)
    rng = np.random.default_rng(9)
    segs = _segments(n, rng)
    t = np.array([0.0, 0.2, 0.5])
    loops = [[0, 1, 2, 3], [3, 2, 1, 0], [0, 2]]
    pairs = [(0, 1), (2, 1), (2, 3), (3, 0)]
    out = pkg.evolve_manybody_hip(calib, psi0=psi0, t_grid=t, segments=segs, loops=loops,
                                  link_pairs=pairs, device=dev)
    O = out['observables']
    assert np.allclose(O['loop_winding_series'][0, :2], [1.0, -1.0], atol=1e-9)
    assert np.isnan(O['loop_winding_series'][:, 2]).all()       # (0, 2) is not a listed link
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0, t)
    for k in range(len(t)):
        xy = orc.spin_observables(n, ref[k])['xy_corr']
        ph, _ = pkg.loop_observables(xy, loops[:2])
        assert np.allclose(O['loop_phase_series'][k, :2], ph, atol=1e-8)
    ph, wd = pkg.loop_observables(orc.spin_observables(n, ref[-1])['xy_corr'], loops)
    assert np.allclose(O['loop_phase'], ph, atol=1e-8) and np.allclose(O['loop_winding'], wd, atol=1e-8)


def test_evolve_manybody_lossy_full_basis(pkg, dev):
    """gamma != 0: the device Taylor propagator of H − (i/2)·γ·1e-3/CONV·n
    against the oracle's exact exponential of the same non-Hermitian H;
    every observable is taken on the decaying (unnormalised) state, as the
    reference does."""
    n = 8
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    rng = np.random.default_rng(21)
    segs = _segments(n, rng)
    gamma = rng.uniform(0.5, 6.0, n)                 # 1/µs
    t = np.array([0.0, 30.0, 90.0])
    link = [(0, 1), (2, 5)]
    out = pkg.evolve_manybody_hip(calib, psi0='grgrgggg', t_grid=t, segments=segs, gamma=gamma,
                                  link_pairs=link, record_psi=True, diag=True, device=dev)
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs,
                                 pkg.state_vector_np('grgrgggg', n), t, gamma=gamma)
    assert np.max(np.abs(out['psi_amp'] - ref[-1])) < 1e-10
    _check(out, ref, n, None, n // 2, link=link)
    assert np.all(np.diff(out['norms']) < 0) and out['norms'][-1] < 0.9
    assert np.allclose(out['loss'], 1 - out['norms'])
    assert out['backend'] == 'hip-ed-nonhermitian' and out['error_report']['hermitian'] is False
    assert out['error_report']['max_norm_dev'] == 0.0 and out['diag']['hermitian'] is False
    assert out['data_schema']['quantities']['loss']['alias_of'] == 'physical_loss'
    assert np.array_equal(out['physical_loss'], out['loss'])
    # a scalar rate applies to every atom
    o2 = pkg.evolve_manybody_hip(calib, psi0='grgrgggg', t_grid=t, segments=segs, gamma=2.0,
                                 device=dev, want_observables=False)
    r2 = orc.propagate_segments(n, calib['pairs'], ['I1'], segs,
                                pkg.state_vector_np('grgrgggg', n), t, gamma=np.full(n, 2.0))
    assert np.max(np.abs(o2['norms'] - np.sum(np.abs(r2) ** 2, axis=1))) < 1e-10


def test_evolve_manybody_lossy_projected_plan(pkg, dev):
    """gamma != 0 on a blockade-graph basis (the loss diagonal summed over
    each code's excited atoms on the device)."""
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    cons = {'kind': 'blockade', 'threshold_mhz': 10.0}
    plan = pkg.build_plan_hip(calib, constraint=cons, device=dev)
    codes = plan['states'].cpu().numpy()
    rng = np.random.default_rng(22)
    segs = _segments(n, rng)
    gamma = rng.uniform(0.5, 6.0, n)
    t = np.array([0.0, 40.0, 100.0])
    out = pkg.evolve_manybody_hip(calib, t_grid=t, segments=segs, plan=plan, gamma=gamma,
                                  device=dev, entropy_cut=5)
    psi0 = pkg.state_vector_np(None, n, codes)
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0, t, codes=codes,
                                 gamma=gamma)
    _check(out, ref, n, codes, 5)
    assert out['backend'] == 'hip-ed-plan-blockade' and out['error_report']['hermitian'] is False


@pytest.mark.parametrize('basis', ['full', 'chain_fib', 'codes'])
def test_propagate_lossy_every_basis_kind(pkg, dev, basis):
    """The loss diagonal of the Taylor propagator reads each rank's code
    through the plan's own basis (identity, Zeckendorf unranking, code
    list): one piece against the oracle's exponential of H − (i/2)γ·n."""
    import torch
    from oracle import hv_ref
    from energy_level_amd import evolve as ev
    n = 10
    rng = np.random.default_rng(31)
    Om = rng.uniform(0.5, 2.0, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.normal(size=n)
    pairs = [{'i': i, 'j': j, 'V': 8.0 / (j - i) ** 6} for i in range(n) for j in range(i + 1, n)]
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': ['I1']}
    gamma = rng.uniform(0.5, 8.0, n)
    model = pkg.from_calib(calib, Om, delta)
    if basis == 'full':
        codes = None
        P = pkg.Plan(model, device=dev)
    elif basis == 'chain_fib':
        codes = hv_ref.fib_codes(n)
        P = pkg.Plan(model, basis='chain_fib', device=dev)
    else:
        codes = hv_ref.fib_codes(n)[::2].copy()        # an arbitrary sorted code list
        P = pkg.Plan(model, basis='codes', codes=codes, device=dev)
    dim = P.dim
    x0 = rng.normal(size=dim) + 1j * rng.normal(size=dim)
    x0 /= np.linalg.norm(x0)
    psi = torch.as_tensor(x0, device=dev)
    dt = 120.0
    ev.propagate_lossy(P, model, psi, dt, gamma)
    seg = [dict(t0=0.0, Om=np.abs(Om), phase=np.angle(Om), delta=delta)]
    ref = orc.propagate_segments(n, pairs, ['I1'], seg, x0, np.array([0.0, dt]), codes=codes,
                                 gamma=gamma)[-1]
    assert np.max(np.abs(psi.cpu().numpy() - ref)) < 1e-11
    assert np.linalg.norm(ref) < 0.95                    # the loss acted
    P.close()


def test_progress_protocol_and_cancellation(pkg, dev):
    """The reference engines' three-slot progress array (coupling.py:24-36,
    engine_ed.py:283-307): arr[0] in-stage fraction, arr[1] cancel flag,
    arr[2] stage; a set cancel flag stops the run at the next output time
    (runtime.check_cancel raises RuntimeError)."""
    n = 6
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    segs = [dict(t0=0.0, Om=np.ones(n), phase=np.zeros(n), delta=np.zeros(n))]
    t = np.linspace(0.0, 40.0, 5)
    arr = np.zeros(3)
    pkg.evolve_manybody_hip(calib, t_grid=t, segments=segs, device=dev, progress_arr=arr)
    assert arr[0] == 1.0 and arr[1] == 0.0 and arr[2] == 2.0
    two = np.zeros(2)                    # a two-slot array gets the fraction only
    pkg.evolve_manybody_hip(calib, t_grid=t, segments=segs, device=dev, progress_arr=two)
    assert two[0] == 1.0 and two[1] == 0.0
    stop = np.array([0.0, 1.0, 0.0])
    with pytest.raises(RuntimeError):
        pkg.evolve_manybody_hip(calib, t_grid=t, segments=segs, device=dev, progress_arr=stop)
