"""Piecewise-constant evolution: forward vs exact exponentials, adjoint gradient
vs FP64 central differences and a dense-expm autograd reference."""
import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu


def _model(pkg, n, c6=10.0):
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(c6), terms=('I1',))
    return calib, pkg.from_calib(calib, np.ones(n), np.zeros(n))


def _drive(M, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5, 1.5, M), rng.uniform(-1.0, 2.0, M)


def test_rabi_closed_form(pkg, dev):
    import torch
    n = 1
    Om = 2 * np.pi * 5.0
    model = pkg.RydModel(1, np.ones(1, complex), np.zeros(1))
    prob = pkg.PulseProblem(model, dt=1.0, K=30, dtype='complex128', device=dev)
    psi0 = torch.zeros(2, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    for M in (1, 7, 50):
        psi = prob.propagate(psi0, np.full(M, Om), np.zeros(M)).cpu().numpy()
        pred = np.sin(pkg.CONV * Om * M / 2) ** 2
        assert abs(abs(psi[1]) ** 2 - pred) < 1e-12


@pytest.mark.parametrize('n,dtype,tol', [(6, 'complex128', 1e-12), (13, 'complex128', 1e-11),
                                         (11, 'complex64', 2e-5), (13, 'complex64', 2e-5),
                                         (18, 'complex64', 2e-5), (16, 'complex128', 1e-11)])
def test_forward_vs_exact_segments(pkg, dev, n, dtype, tol):
    import torch
    calib, model = _model(pkg, n)
    M, dt = 12, 0.5
    Om, De = _drive(M, 1)
    prob = pkg.PulseProblem(model, dt=dt, K=15, dtype=dtype, device=dev)
    psi0 = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    got = prob.propagate(psi0, Om, De).cpu().numpy()
    segs = [{'t0': k * dt, 'Om': [Om[k]] * n, 'delta': [-De[k]] * n} for k in range(M)]
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0.cpu().numpy(),
                                 np.array([0.0, M * dt]),
                                 method='expm' if n <= 8 else 'expm_multiply')[-1]
    assert np.max(np.abs(got - ref)) < tol


def test_adjoint_vs_central_differences(pkg, dev):
    import torch
    n, M, dt = 8, 30, 0.05
    _, model = _model(pkg, n)
    Om, De = _drive(M, 2)
    prob = pkg.PulseProblem(model, dt=dt, K=15, dtype='complex128', device=dev, checkpoint=7)
    psi0 = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    tgt = pkg.z2_index(n)
    L, gO, gD, _ = prob.loss_and_grad(psi0, Om, De, tgt)
    h = 1e-5
    for k in (0, 13, M - 1):
        e = np.zeros(M)
        e[k] = h
        fd_o = (prob.loss_and_grad(psi0, Om + e, De, tgt)[0]
                - prob.loss_and_grad(psi0, Om - e, De, tgt)[0]) / (2 * h)
        fd_d = (prob.loss_and_grad(psi0, Om, De + e, tgt)[0]
                - prob.loss_and_grad(psi0, Om, De - e, tgt)[0]) / (2 * h)
        assert abs(gO[k] - fd_o) < 1e-7 * max(1.0, abs(fd_o)), (k, gO[k], fd_o)
        assert abs(gD[k] - fd_d) < 1e-7 * max(1.0, abs(fd_d)), (k, gD[k], fd_d)


def _per_slice_adjoint(pkg, prob, psi0, Om, De, tgt):
    """Gradient by the one-slice entry point ryd_evolve_adjoint: every slice
    input from a fresh forward, the adjoint sums copied to the host per
    slice (the straightforward form the stored sweep must reproduce)."""
    import torch
    from energy_level_amd import _native as nat
    P = prob.plan
    params = prob._slice_params(Om, De)
    Omv, dadd, cs, rs, coef = params
    M, D, K = len(Omv), prob.dim, prob.K
    work = torch.empty(4 * D, dtype=prob.tdt, device=dev_of(P))
    states = []
    psi = psi0.to(prob.tdt).clone()
    for k in range(M):
        states.append(psi.clone())
        prob._forward_slices(psi, params, k, k + 1, work)
    o = psi[tgt].item()
    g = torch.zeros(D, dtype=prob.tdt, device=dev_of(P))
    g[tgt] = -o
    tstore = torch.empty(K * D, dtype=prob.tdt, device=dev_of(P))
    vwork = torch.empty(3 * D, dtype=prob.tdt, device=dev_of(P))
    sums = torch.zeros(4 * K, dtype=torch.float64, device=dev_of(P))
    all_sums = np.zeros((M, K, 4))
    code = nat.dtype_code(prob.tdt)
    for k in range(M - 1, -1, -1):
        cf = np.ascontiguousarray(coef[k])
        nat.check(nat.lib().ryd_evolve_adjoint(
            P.handle, code, nat.ptr(states[k]), nat.ptr(g), float(Omv[k]), float(dadd[k]),
            float(cs[k]), float(rs[k]), K, cf.ctypes.data, nat.ptr(tstore), nat.ptr(vwork),
            nat.ptr(sums), nat.ptr(P.workspace()), nat.stream_ptr()))
        all_sums[k] = sums.view(K, 4).cpu().numpy()
    gO, gD = prob.grad_from_sums(all_sums, rs)
    return 1.0 - abs(o) ** 2, gO, gD, g


def dev_of(P):
    return P.device


@pytest.mark.parametrize('n,dtype,tol', [(9, 'complex128', 1e-13), (12, 'complex64', 2e-6),
                                         (14, 'complex128', 1e-13)])
def test_stored_sweep_matches_per_slice_adjoint(pkg, dev, n, dtype, tol):
    """The checkpointed sweep (segments recomputed once with their Chebyshev
    vectors stored, sums kept on the device) against the one-slice adjoint
    entry point, with whole segments and with the store forced down to one
    sub-step (chunks recomputed from the checkpoint); natural units so that
    slices take several sub-steps and checkpoints fall inside slices' runs."""
    import torch
    _, model = _model(pkg, n)
    Om, De = _drive(17, 5)
    psi0 = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    psi0[3] = 0.5
    psi0 /= torch.linalg.norm(psi0)
    tgt = pkg.z2_index(n)
    plan = pkg.Plan(model, device=dev)
    full = pkg.PulseProblem(model, dt=0.05, K=15, dtype=dtype, plan=plan, checkpoint=5,
                            units='natural')
    tiny = pkg.PulseProblem(model, dt=0.05, K=15, dtype=dtype, plan=plan, checkpoint=5,
                            units='natural', store_budget_bytes=1)
    L, gO, gD, _ = full.loss_and_grad(psi0, Om, De, tgt)
    assert len(full._owner) > 17          # sub-stepped slices
    L1, gO1, gD1, _ = tiny.loss_and_grad(psi0, Om, De, tgt)
    assert L1 == L and np.array_equal(gO1, gO) and np.array_equal(gD1, gD)
    Lr, gOr, gDr, _ = _per_slice_adjoint(pkg, full, psi0, Om, De, tgt)
    scale = max(np.max(np.abs(gOr)), np.max(np.abs(gDr)))
    assert abs(Lr - L) <= tol
    assert np.max(np.abs(gO - gOr)) <= tol * scale
    assert np.max(np.abs(gD - gDr)) <= tol * scale
    # torch autograd through the lazy backward equals the eager gradient
    om_t = torch.tensor(Om, requires_grad=True)
    de_t = torch.tensor(De, requires_grad=True)
    loss = pkg.PulseEvolution.apply(om_t, de_t, full, psi0, tgt)
    (2.0 * loss).backward()
    assert float(loss.detach()) == L
    assert np.array_equal(om_t.grad.numpy(), 2.0 * gO)
    assert np.array_equal(de_t.grad.numpy(), 2.0 * gD)
    plan.close()


def test_adjoint_vs_dense_expm_autograd(pkg, dev):
    """Reference gradient of the exact propagator by torch autograd through
    matrix_exp on the CPU (complex128)."""
    import torch
    n, M, dt = 6, 20, 0.05
    calib, model = _model(pkg, n)
    Om, De = _drive(M, 3)
    prob = pkg.PulseProblem(model, dt=dt, K=15, dtype='complex128', device=dev)
    psi0 = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    tgt = pkg.z2_index(n)
    L, gO, gD, _ = prob.loss_and_grad(psi0, Om, De, tgt)
    X = torch.as_tensor(orc.kron_hamiltonian(n, np.ones(n), np.zeros(n), [], []))
    Nop = torch.as_tensor(orc.kron_hamiltonian(n, np.zeros(n), np.ones(n), [], []))
    V = torch.as_tensor(orc.kron_hamiltonian(n, np.zeros(n), np.zeros(n), calib['pairs'], ['I1']))
    om_t = torch.tensor(Om, requires_grad=True)
    de_t = torch.tensor(De, requires_grad=True)
    psi = torch.zeros(1 << n, dtype=torch.complex128)
    psi[0] = 1
    for k in range(M):
        H = om_t[k] * X - de_t[k] * Nop + V
        psi = torch.linalg.matrix_exp(-1j * pkg.CONV * dt * H) @ psi
    loss = 1 - torch.abs(psi[tgt]) ** 2
    loss.backward()
    assert abs(float(loss.detach()) - L) < 1e-12
    assert np.max(np.abs(om_t.grad.numpy() - gO)) < 1e-10 * max(1, np.max(np.abs(gO)))
    assert np.max(np.abs(de_t.grad.numpy() - gD)) < 1e-10 * max(1, np.max(np.abs(gD)))


def test_autograd_function_and_complex64(pkg, dev):
    import torch
    n, M, dt = 10, 40, 0.05
    _, model = _model(pkg, n)
    Om, De = _drive(M, 4)
    psi0 = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    psi0[0] = 1
    tgt = pkg.z2_index(n)
    p64 = pkg.PulseProblem(model, dt=dt, dtype='complex128', device=dev)
    p32 = pkg.PulseProblem(model, dt=dt, dtype='complex64', device=dev)
    om_t = torch.tensor(Om, requires_grad=True)
    de_t = torch.tensor(De, requires_grad=True)
    loss = pkg.PulseEvolution.apply(om_t, de_t, p64, psi0, tgt)
    loss.backward()
    L32, g32, d32, _ = p32.loss_and_grad(psi0, Om, De, tgt)
    assert abs(L32 - float(loss.detach())) < 1e-5
    assert np.max(np.abs(g32 - om_t.grad.numpy())) < 1e-4 * max(1, np.max(np.abs(g32)))
    assert np.max(np.abs(d32 - de_t.grad.numpy())) < 1e-4 * max(1, np.max(np.abs(d32)))


@pytest.mark.parametrize('interp,dtype,tol', [('left', 'complex128', 1e-10),
                                              ('midpoint', 'complex128', 1e-10),
                                              ('left', 'complex64', 2e-5)])
def test_segments_match_exact_exponentials(pkg, dev, interp, dtype, tol):
    """Per-site complex drives in three segments, output grid off the jumps."""
    n = 6
    rng = np.random.default_rng(11)
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    segs = [dict(t0=t0, Om=rng.uniform(0.5, 3.0, n), phase=rng.uniform(0, 2 * np.pi, n),
                 delta=rng.uniform(-2, 2, n)) for t0 in (0.0, 70.0, 150.0)]
    t = np.array([0.0, 40.0, 100.0, 160.0, 230.0])
    out = pkg.propagate_segments_hip(calib, segs, t, interp=interp, dtype=dtype, device=dev,
                                     return_states=True)
    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1.0
    ref = orc.propagate_segments(n, calib['pairs'], ['I1'], segs, psi0, t, interp=interp)
    assert np.max(np.abs(out['states'] - ref)) < tol
    pops = np.array([orc.site_occupations(n, v) for v in ref])
    assert np.allclose(out['pops_atom'], pops, atol=tol)
    assert np.allclose(out['norms'], 1.0, atol=tol)
    assert out['basis_probabilities'].shape == (len(t), 1 << n)


def test_segments_single_atom_rabi(pkg, dev):
    """One atom, Ω = 2 MHz: P_r(t) = sin²(CONV·Ω·t/2) (reference m1 known answer)."""
    calib = {'atoms': [{}], 'pairs': [], 'terms': ['I1']}
    t = np.linspace(0.0, 400.0, 9)
    out = pkg.propagate_segments_hip(calib, [dict(t0=0.0, Om=[2.0], delta=[0.0])], t,
                                     device=dev)
    CONV = 2 * np.pi * 1e-3
    assert np.allclose(out['pops_atom'][:, 0], np.sin(CONV * 2.0 * t / 2) ** 2, atol=1e-12)
