"""Pauli-string expectations and site correlations on the device, on every
plan kind (full basis, chain sector, code lists, blockade graphs), against
the oracle's restatement of PauliPlan and corr_matrix."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc
from test_gpu_graph import _lattice_calib

pytestmark = pytest.mark.gpu


def _strings(n, rng, count=19):
    out = [[(0, 'z')], [(n - 1, 'x')], [(n // 2, 'y')], [(i, 'x') for i in range(n)]]
    while len(out) < count:
        k = int(rng.integers(1, min(n, 6) + 1))
        sites = rng.choice(n, k, replace=False)
        out.append([(int(s), str(rng.choice(['x', 'y', 'z']))) for s in sites])
    return out


def _plans(pkg, dev):
    """(name, plan, codes or None) for each plan kind at small sizes."""
    n = 12
    calib = _lattice_calib(4, 3, c6=30.0)
    m = pkg.from_calib(calib, np.ones(n), np.zeros(n))
    edges = pkg.blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 10.0})
    full = pkg.Plan(m, device=dev)
    chain = pkg.Plan(m, basis='chain_fib', device=dev)
    graph = pkg.Plan(m, basis='graph', edges=edges, device=dev)
    num = pkg.number_sector_basis(n, 4)
    nplan = pkg.Plan(m, basis='codes', codes=num, device=dev)
    return [('full', full, None),
            ('chain', chain, pkg.sector_codes(chain).cpu().numpy()),
            ('graph', graph, np.asarray(orc.independent_states(n, edges))),
            ('number', nplan, num)]


@pytest.mark.parametrize('dtype', ['complex128', 'complex64', 'float64'])
def test_pauli_strings_all_plan_kinds(pkg, dev, dtype):
    import torch
    rng = np.random.default_rng(1)
    n = 12
    specs = pkg.normalize_pauli_strings([[list(t) for t in s] for s in _strings(n, rng)], n)
    tol = 1e-12 if dtype != 'complex64' else 3e-6
    for name, P, codes in _plans(pkg, dev):
        D = P.dim
        v = rng.normal(size=D) + (1j * rng.normal(size=D) if dtype != 'float64' else 0)
        v /= np.linalg.norm(v)
        x = torch.as_tensor(v, device=dev).to(getattr(torch, dtype))
        got = pkg.PauliPlanHip(P, specs).expect(x)
        for s, g in zip(specs, got):
            ref = orc.pauli_expect(n, v, s['terms'], codes)
            assert abs(g - ref) <= tol, (name, s['label'], g, ref)


def test_pauli_strings_62_site_code_list(pkg, dev):
    """A code list on 62 sites (bits beyond 32 in the masks)."""
    import torch
    n = 62
    rng = np.random.default_rng(3)
    base = rng.integers(0, 1 << 62, 300, dtype=np.int64)
    codes = np.unique(np.concatenate([base, base ^ (1 << 61), base ^ (1 << 40) ^ (1 << 3)]))
    m = pkg.RydModel(n=n, Om=np.zeros(n, complex), delta=np.zeros(n))
    P = pkg.Plan(m, basis='codes', codes=codes, device=dev)
    v = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    specs = pkg.normalize_pauli_strings([[[0, 'x']], [[21, 'y'], [58, 'x']], [[0, 'z'], [61, 'z']]], n)
    got = pkg.measure_pauli_strings_hip(torch.as_tensor(v, device=dev), specs, P)
    for s, g in zip(specs, got):
        assert abs(g - orc.pauli_expect(n, v, s['terms'], codes)) < 1e-10


@pytest.mark.parametrize('dtype', ['complex128', 'float32'])
def test_correlations_all_plan_kinds(pkg, dev, dtype):
    import torch
    rng = np.random.default_rng(5)
    n = 12
    for name, P, codes in _plans(pkg, dev):
        v = rng.normal(size=P.dim) + (1j * rng.normal(size=P.dim) if dtype != 'float32' else 0)
        v /= np.linalg.norm(v)
        x = torch.as_tensor(v, device=dev).to(getattr(torch, dtype))
        C, ni = pkg.corr_matrix_hip(P, x)
        ref = orc.corr_matrix(n, v, codes)
        tol = 1e-12 if dtype == 'complex128' else 2e-6
        assert np.max(np.abs(C - ref)) < tol, name
        assert np.max(np.abs(ni - orc.site_occupations(n, v, codes))) < tol
        C2, _ = pkg.corr_matrix_hip(P, x)
        assert np.array_equal(C, C2)                        # deterministic sums
        # a state of norm 0.7 (as under loss) is not renormalised
        C3, ni3 = pkg.corr_matrix_hip(P, 0.7 * x)
        assert np.max(np.abs(C3 - orc.corr_matrix(n, 0.7 * v, codes))) < tol, name
        assert np.max(np.abs(ni3 - 0.49 * orc.site_occupations(n, v, codes))) < tol


def test_correlations_64_sites_and_large_sector(pkg, dev):
    """All 2080 pairs of 64 sites (33 pairs per lane), and the chain sector
    at N = 30 (D = 2,178,309) against the bit matrix of its codes."""
    import torch
    rng = np.random.default_rng(6)
    n = 62
    codes = np.unique(rng.integers(0, 1 << 62, 5000, dtype=np.int64))
    m = pkg.RydModel(n=n, Om=np.zeros(n, complex), delta=np.zeros(n))
    P = pkg.Plan(m, basis='codes', codes=codes, device=dev)
    v = rng.normal(size=len(codes))
    v /= np.linalg.norm(v)
    C, _ = pkg.corr_matrix_hip(P, torch.as_tensor(v, device=dev))
    assert np.max(np.abs(C - orc.corr_matrix(n, v, codes))) < 1e-12
    n = 30
    mc = pkg.RydModel(n=n, Om=np.ones(n, complex), delta=np.zeros(n))
    Pc = pkg.Plan(mc, basis='chain_fib', device=dev)
    cc = pkg.sector_codes(Pc).cpu().numpy()
    w = rng.normal(size=Pc.dim)
    w /= np.linalg.norm(w)
    S, norm = pkg.site_correlation_sums(Pc, torch.as_tensor(w, device=dev))
    p = w ** 2
    bits = ((cc[None, :] >> (n - 1 - np.arange(n))[:, None]) & 1).astype(np.float64)
    ref = (bits * p) @ bits.T
    assert abs(norm - 1.0) < 1e-12
    assert np.max(np.abs(S - ref)) < 1e-11


def test_evolution_records_pauli_and_corr(pkg, dev):
    n = 8
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    rng = np.random.default_rng(9)
    segs = [dict(t0=t0, Om=rng.uniform(0.5, 2.0, n), phase=rng.uniform(0, 6.3, n),
                 delta=rng.uniform(-1, 1, n)) for t0 in (0.0, 40.0)]
    t = np.array([0.0, 30.0, 90.0])
    specs = [[[0, 'x']], [[1, 'y'], [2, 'y']], [[3, 'z'], [4, 'x'], [5, 'y']]]
    out = pkg.propagate_segments_hip(calib, segs, t, device=dev, pauli_strings=specs, corr=True,
                                     return_states=True)
    norm = pkg.normalize_pauli_strings(specs, n)
    for k, v in enumerate(out['states']):
        for q, s in enumerate(norm):
            assert abs(out['pauli'][k, q] - orc.pauli_expect(n, v, s['terms'])) < 1e-10
        assert np.max(np.abs(out['corr'][k] - orc.corr_matrix(n, v))) < 1e-10
    assert out['pauli_labels'] == [s['label'] for s in norm]
