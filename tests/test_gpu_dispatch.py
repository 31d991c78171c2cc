"""The HIP equilibrium dispatch against the oracle: result keys of the
reference's solve_equilibrium and values of its ED/KPM paths."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu


def _system(pkg, n, c6=10.0, De=1.2, dev='cuda:0'):
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(c6), terms=('I1',))
    return dict(calib=calib, Om=np.ones(n), delta=-De * np.ones(n), device=dev)


def test_gs_hip_lanczos_keys_and_values(pkg, dev):
    n = 10
    sysd = _system(pkg, n, dev=dev)
    out = pkg.solve_equilibrium_hip(sysd, 'gs', backend='hip-lanczos', return_state=True)
    H = orc.build_H_sparse(n, sysd['Om'], sysd['delta'], sysd['calib']['pairs'], ['I1'])
    ref = orc.ground_state_dense(H)
    for key in ('want', 'backend', 'n_sites', 'E0', 'gap', 'evals', 'residuals', 'converged',
                'dim', 'krylov_dim', 'restarts', 'tolerance', 'vec', 'n_atom',
                'requested_backend', 'resolved_backend', 'basis'):
        assert key in out, key
    # as dict(base, **out) in the reference: the solver's backend string wins
    assert out['want'] == 'gs' and out['n_sites'] == n
    assert out['backend'] == out['resolved_backend'] == 'hip-matrixfree-lanczos'
    assert out['requested_backend'] == 'hip-lanczos'
    assert out['basis']['bit_order'] == 'site0-msb' and out['basis']['dim'] == 1 << n
    assert abs(out['E0'] - ref['E0']) < 1e-10 * abs(ref['E0'])
    assert abs(out['gap'] - ref['gap']) < 1e-8
    assert np.allclose(out['n_atom'], orc.site_occupations(n, ref['vec']), atol=1e-9)


def test_gs_hip_lanczos_projected_chain_plan(pkg, dev):
    n = 12
    sysd = _system(pkg, n, c6=30.0, De=1.5, dev=dev)
    edges = [(i, i + 1) for i in range(n - 1)]
    states = np.asarray(orc.independent_states(n, edges))
    sysd['plan'] = dict(projected=True, dim=len(states), states=states)
    out = pkg.solve_equilibrium_hip(sysd, 'gs', backend='hip-lanczos', return_state=True)
    H = orc.build_H_sparse(n, sysd['Om'], sysd['delta'], sysd['calib']['pairs'], ['I1'], states)
    ref = orc.ground_state_dense(H)
    assert abs(out['E0'] - ref['E0']) < 1e-10 * abs(ref['E0'])
    assert np.array_equal(out['basis']['indices'], states)
    assert np.allclose(out['n_atom'], orc.site_occupations(n, ref['vec'], states), atol=1e-9)


def test_gs_hip_batch_single_system(pkg, dev):
    n = 9
    sysd = _system(pkg, n, dev=dev)
    out = pkg.solve_equilibrium_hip(sysd, 'gs', backend='hip-batch')
    H = orc.build_H_sparse(n, sysd['Om'], sysd['delta'], sysd['calib']['pairs'], ['I1'])
    ref = orc.ground_state_dense(H)
    assert abs(out['E0'] - ref['E0']) < 1e-9 * abs(ref['E0'])
    assert np.allclose(out['n_atom'], orc.site_occupations(n, ref['vec']), atol=1e-6)


def test_gs_hip_batch_exchange_and_complex_drive(pkg, dev):
    """A system the batch kernels do not take (J and A pairs, a complex
    drive) still answers hip-batch, as the reference's dense batch does."""
    n = 9
    rng = np.random.default_rng(9)
    terms = ['I1', 'I2', 'I3']
    pairs = [{'i': i, 'j': j, 'V': 2.0 / (j - i) ** 6, 'J': 0.7 / (j - i) ** 3,
              'A': 0.2 / (j - i) ** 3} for i in range(n) for j in range(i + 1, n)]
    calib = {'atoms': [{}] * n, 'pairs': pairs, 'terms': terms}
    Om = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.uniform(-1, 1, n)
    sysd = dict(calib=calib, Om=Om, delta=delta, terms=terms, device=dev)
    out = pkg.solve_equilibrium_hip(sysd, 'gs', backend='hip-batch')
    H = orc.build_H_sparse(n, Om, delta, pairs, terms)
    ref = orc.ground_state_dense(H)
    assert out['backend'] == 'hip-batch' and out['n_sites'] == n
    assert abs(out['E0'] - ref['E0']) < 1e-9 * abs(ref['E0'])
    assert abs(out['gap'] - ref['gap']) < 1e-6
    assert np.allclose(out['n_atom'], orc.site_occupations(n, ref['vec']), atol=1e-6)


@pytest.mark.parametrize('want', ['thermal', 'dos'])
def test_hip_kpm_thermal_and_dos_vs_oracle(pkg, dev, want):
    n = 8
    sysd = _system(pkg, n, dev=dev)
    num, nvec, seed = 60, 6, 3
    out = pkg.solve_equilibrium_hip(sysd, want, backend='hip-kpm', betas=[0.5, 2.0], num=num,
                                    nvec=nvec, seed=seed)
    H = orc.build_H_sparse(n, sysd['Om'], sysd['delta'], sysd['calib']['pairs'], ['I1'])
    # same bounds as the device run, then the reference recurrence and post-processing
    P = pkg.Plan(pkg.from_calib(sysd['calib'], sysd['Om'], sysd['delta']), device=dev)
    emin, emax, c, r = pkg.spectral_bounds(P, margin=1e-3)
    mu, _, _ = orc.kpm_moments(H, num, nvec, seed=seed, c=c, r=r)
    if want == 'thermal':
        th = orc.thermal_from_moments(H.shape[0], np.array([0.5, 2.0]), mu, c, r, num)
        for k in ('E', 'C', 'F'):
            assert np.allclose(out[k], th[k], rtol=1e-8, atol=1e-9), k
    else:
        assert out['energies'].shape == (400,)
        rho = orc.dos_from_moments(mu, c, r, out['energies'], num)
        assert np.allclose(out['rho'], rho, rtol=1e-8, atol=1e-10)
