"""Batched Lanczos over uneven-N instances against per-instance runs and dense
ground states; hipGraph replays against the eager sequence."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu


def _models(pkg, ns, seed=0):
    rng = np.random.default_rng(seed)
    out, calibs = [], []
    for i, n in enumerate(ns):
        pos = pkg.chain_positions(n) + rng.normal(0, 0.1, size=(n, 3))
        calib = pkg.from_positions(pos, pkg.c6_coupling(10.0), terms=('I1',))
        De = np.linspace(-2, 4, 32)[i % 32]
        out.append(pkg.from_calib(calib, np.ones(n), -De * np.ones(n)))
        calibs.append((calib, De))
    return out, calibs


def test_batch_matches_single_instance_recurrence(pkg, dev):
    import torch
    ns = [4, 7, 10, 11, 12, 13, 15, 9]
    models, calibs = _models(pkg, ns)
    steps = 90
    B = pkg.BatchLanczos(models, steps=steps, dtype='float64', device=dev)
    v0 = B.start_vectors(3)
    B.run(v0, use_graph=False)
    alpha, nrm, stop = B.coefficients()
    for b, m in enumerate(models):
        P = pkg.Plan(m, device=dev)
        seg = v0[B.offsets[b]:B.offsets[b + 1]].clone()
        R = pkg.lanczos_run(P, seg, steps, 1, 0.0, want=1)
        # identical recurrences until loss of orthogonality amplifies rounding
        k = min(R.j, (1 << m.n) // 2, 20)
        assert np.allclose(alpha[b, :k], R.a[:k], rtol=1e-10, atol=1e-10)
        assert np.allclose(nrm[b, 1:k], R.b[:k - 1], rtol=1e-10, atol=1e-10)
    gs = B.ground_states()
    for b, (calib, De) in enumerate(calibs):
        n = ns[b]
        H = orc.build_H_sparse(n, np.ones(n), -De * np.ones(n), calib['pairs'], ['I1'])
        e0 = (orc.ground_state_dense(H) if n <= 10 else orc.ground_state_eigsh(H))['E0']
        assert abs(gs['E0'][b] - e0) < 1e-8 * max(1, abs(e0)), (n, gs['E0'][b], e0)


def test_batch_fp32_graph_replays_and_autotune(pkg, dev):
    import torch
    models = pkg.disorder_ensemble(count=24, nmin=10, nmax=16, seed=5)
    B = pkg.BatchLanczos(models, steps=150, dtype='float32', device=dev)
    v0 = B.start_vectors(1)
    B.run(v0, use_graph=False)
    ref_alpha, ref_nrm, _ = B.coefficients()
    for rep in range(3):
        B.run(v0, use_graph=True)
        a, nr, _ = B.coefficients()
        assert np.array_equal(a, ref_alpha) and np.array_equal(nr, ref_nrm), rep
        torch.randn(1000, device=dev).sum().item()      # eager work between replays
    tb = B.autotune(v0)
    assert tb in (10, 11, 12) and set(B.autotune_ms) == {10, 11, 12}
    B.run(v0, use_graph=True)
    gs = B.ground_states()
    B64 = pkg.BatchLanczos(models, steps=150, dtype='float64', device=dev)
    B64.run(B64.start_vectors(1), use_graph=True)
    g64 = B64.ground_states()
    conv = g64['ritz_bound'] < 1e-6 * np.abs(g64['E0'])
    assert conv.sum() >= len(models) // 2
    rel = np.abs(gs['E0'] - g64['E0']) / np.abs(g64['E0'])
    assert np.max(rel[conv]) < 2e-5


@pytest.mark.parametrize('dtype,tol', [('float64', 1e-7), ('float32', 2e-3)])
def test_batch_occupations_vs_dense(pkg, dev, dtype, tol):
    """n_atom of every instance (Ritz replay) against dense eigenvectors."""
    ns = [4, 7, 10, 11, 12, 9, 6, 8]
    models, calibs = _models(pkg, ns, seed=2)
    out = pkg.batched_ground_state_hip(models, steps=120, dtype=dtype, device=dev, seed=4)
    assert len(out['n_atom']) == len(ns)
    checked = 0
    for b, (calib, De) in enumerate(calibs):
        n = ns[b]
        H = orc.build_H_sparse(n, np.ones(n), -De * np.ones(n), calib['pairs'], ['I1'])
        gs = orc.ground_state_dense(H)
        if gs['gap'] < 1e-3:
            continue                      # near-degenerate: the Ritz vector is not unique
        ref = orc.site_occupations(n, gs['vec'])
        assert out['n_atom'][b].shape == (n,)
        assert np.max(np.abs(out['n_atom'][b] - ref)) < tol, (n, out['n_atom'][b], ref)
        checked += 1
    assert checked >= 6


@pytest.mark.parametrize('dtype,rtol', [('float64', 1e-12), ('float32', 1e-5)])
def test_batch_stored_basis_matches_replay(pkg, dev, dtype, rtol):
    """A lent basis (every Lanczos row kept) gives the replay's recurrence bit
    for bit and its Ritz vectors and occupations to rounding, including an
    instance with a zero start vector (stopped before its first step, as a
    finished instance of a restarted batch is): its rows are never written,
    so the lent memory is filled with NaN first, and its zero coefficients
    must skip them."""
    import ctypes
    import torch
    ns = [4, 7, 10, 11, 12, 9]
    models, _ = _models(pkg, ns, seed=6)
    Bs = pkg.BatchLanczos(models, steps=90, dtype=dtype, device=dev, store_basis=True)
    Br = pkg.BatchLanczos(models, steps=90, dtype=dtype, device=dev, store_basis=False)
    assert Bs.stores_basis and not Br.stores_basis
    Bs._basis.view(Bs.tdt).fill_(float('nan'))
    v0 = Bs.start_vectors(2)
    v0[int(Bs.offsets[1]):int(Bs.offsets[2])] = 0.0
    for graph in (False, True):
        Bs.run(v0, use_graph=graph)
        Br.run(v0, use_graph=graph)
        a_s, n_s, st_s = Bs.coefficients()
        a_r, n_r, st_r = Br.coefficients()
        live = np.arange(len(ns)) != 1      # instance 1's rows past its stop are never set
        assert np.array_equal(a_s[live], a_r[live]) and np.array_equal(n_s[live], n_r[live])
        assert np.array_equal(st_s, st_r) and n_s[1, 0] == 0.0
    assert st_s[1] == 1
    Bs.ground_states()
    Br.ground_states()
    which = np.zeros(len(ns), np.int64)
    which[1] = -1
    x_s, occ_s = Bs.ritz(v0, Bs.restart_coefficients(which))
    x_r, occ_r = Br.ritz(v0, Br.restart_coefficients(which))
    assert torch.isfinite(x_s).all()
    assert not x_s[int(Bs.offsets[1]):int(Bs.offsets[2])].any()
    scale = torch.linalg.vector_norm(x_r.double()).item()
    assert torch.max(torch.abs(x_s.double() - x_r.double())).item() <= rtol * scale
    assert np.allclose(occ_s.cpu().numpy(), occ_r.cpu().numpy(), rtol=10 * rtol, atol=0)
    # too small a buffer is refused; detaching returns to the rotating rows
    L = pkg._native.lib()
    need = L.ryd_batch_basis_bytes(Bs._h)
    assert need == (90 + 1) * Bs.total * Bs._basis.element_size() * (8 if dtype == 'float64' else 4)
    assert L.ryd_batch_attach_basis(Bs._h, ctypes.c_void_p(Bs._basis.data_ptr()),
                                    need - 1) == pkg._native.RYD_ERR_INVALID
    pkg._native.check(L.ryd_batch_attach_basis(Bs._h, None, 0))
    Bs.run(v0, use_graph=True)
    x_d, _ = Bs.ritz(v0, Bs.restart_coefficients(which))
    assert torch.max(torch.abs(x_d.double() - x_r.double())).item() <= rtol * scale
    Bs.close()
    Br.close()


def test_certified_fp64_cycles_in_parts_match_whole_cycles(pkg, dev):
    """FP64 refinement cycles run in parts, an instance stopping at the part
    where its pairs pass, give the whole-cycle results (both certified to
    1e-9) in no more steps; E0 and gap checked against the oracle's eigsh."""
    ns = [10, 12, 13, 14, 15, 16, 11, 16, 14, 13, 15, 12]
    models, calibs = _models(pkg, ns, seed=5)
    kw = dict(steps=150, dtype='float32', tol=1e-9, bulk_cycles=2,
              refine_schedule=((300, 1), (700, 2)), seed=2, device=dev, autotune=False,
              tile_bits=12, observables=False, store_basis=True)
    whole = pkg.batch.certified_ground_states(models, refine_parts=None, **kw)
    parts = pkg.batch.certified_ground_states(models, refine_parts=100, **kw)
    assert np.all(whole['converged']) and np.all(parts['converged'])
    assert int(parts['steps_float64'].sum()) <= int(whole['steps_float64'].sum())
    for b, (calib, De) in enumerate(calibs):
        n = ns[b]
        s = max(1.0, abs(whole['E0'][b]))
        assert abs(parts['E0'][b] - whole['E0'][b]) <= 1e-9 * s
        assert abs(parts['gap'][b] - whole['gap'][b]) <= 2e-9 * s
        H = orc.build_H_sparse(n, np.ones(n), -De * np.ones(n), calib['pairs'], ['I1'])
        e0 = orc.ground_state_eigsh(H)['E0']
        assert abs(parts['E0'][b] - e0) <= 1e-9 * s, (n, parts['E0'][b], e0)


def test_run_parts_stops_at_the_part_and_reads_its_steps(pkg, dev):
    """An instance stopped by run_parts keeps the step count where it passed;
    its Ritz values equal those of a whole run cut at that step."""
    import torch
    ns = [10, 12, 14]
    models, _ = _models(pkg, ns, seed=7)
    B = pkg.BatchLanczos(models, steps=400, dtype='float64', device=dev, store_basis=True)
    v0 = B.start_vectors(4)
    need = np.ones(3, bool)
    B.run_parts(v0, 1e-9, 1e-10, need, need, part=50)
    gs = B.ground_states(1e-9, None, 1e-10)
    k = gs['steps']
    assert np.all(k % 50 == 0) and np.all(k >= 50) and np.all(k <= 400)
    W = pkg.BatchLanczos(models, steps=400, dtype='float64', device=dev, store_basis=True)
    W.run(v0, use_graph=False)
    a1, n1, _ = B.coefficients()
    a2, n2, _ = W.coefficients()
    for b in range(3):
        kb = int(k[b])
        assert np.array_equal(a1[b, :kb], a2[b, :kb])
        assert np.array_equal(n1[b, :kb + 1], n2[b, :kb + 1])
        if kb < 400:
            assert gs['converged'][b]
    x, _ = B.ritz(v0, B.restart_coefficients(np.zeros(3, np.int64)))
    assert bool(torch.isfinite(x).all())
    B.close()
    W.close()
