"""ryd_lanczos on small full-basis plans (one launch per call for n <= 13,
complex128 n <= 12; complex128 at n = 13 takes the multi-launch form):
alpha/beta against the oracle's recurrence, a run split over several calls
equal to one call, the replay's Ritz vector against the Lanczos residual
identity, and breakdowns."""
import ctypes

import numpy as np
import pytest

from oracle import rydberg_oracle as orc

pytestmark = pytest.mark.gpu

DTYPES = ('float32', 'float64', 'complex64', 'complex128')


def _case(pkg, n, dtype, omega_scale=1.0):
    import torch
    cplx = dtype.startswith('complex')
    rng = np.random.default_rng(100 + n)
    calib = (pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
             if n > 1 else {'atoms': [{}], 'pairs': []})
    Om = rng.uniform(0.5, 1.5, n) * omega_scale
    if cplx:
        Om = Om * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    delta = rng.uniform(-1.0, 1.0, n)
    plan = pkg.Plan(pkg.from_calib(calib, Om, delta), device='cuda')
    H = orc.build_H_sparse(n, Om, delta, calib['pairs'], ['I1'])
    tdt = getattr(torch, dtype)
    return plan, H, tdt


def _lanczos(pkg, plan, tdt, v0, chunks, tiny=1e-10, replay_y=None, alpha=None, nrm=None):
    """ryd_lanczos over `chunks` (steps per call); returns alpha, nrm, stop, steps
    (host) and, with replay_y, the replayed Ritz vector."""
    import torch
    nat = pkg._native
    L = nat.lib()
    dev = plan.device
    K = int(sum(chunks))
    bufs = [torch.zeros(plan.dim, dtype=tdt, device=dev) for _ in range(3)]
    bufs[0].copy_(v0)
    if alpha is None:
        alpha = torch.full((K + 1,), float('nan'), dtype=torch.float64, device=dev)
        nrm = torch.full((K + 2,), float('nan'), dtype=torch.float64, device=dev)
    state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=dev)
    code = nat.dtype_code(tdt)
    ws, st = plan.workspace(), nat.stream_ptr()
    out = None
    if replay_y is None:
        j = 0
        for c in chunks:
            nat.check(L.ryd_lanczos(plan.handle, code, nat.ptr(bufs[0]), nat.ptr(bufs[1]),
                                    nat.ptr(bufs[2]), j, int(c), nat.ptr(alpha), nat.ptr(nrm),
                                    nat.ptr(state), float(tiny), None, None, None, None, 0,
                                    nat.ptr(ws), st))
            j += int(c)
    else:
        out = torch.zeros(plan.dim, dtype=tdt, device=dev)
        y = torch.as_tensor(np.ascontiguousarray(replay_y), dtype=torch.float64, device=dev)
        nat.check(L.ryd_lanczos(plan.handle, code, nat.ptr(bufs[0]), nat.ptr(bufs[1]),
                                nat.ptr(bufs[2]), 0, len(replay_y), nat.ptr(alpha), nat.ptr(nrm),
                                nat.ptr(state), 0.0, nat.ptr(out), None, nat.ptr(y), None, 1,
                                nat.ptr(ws), st))
    flag, nvalid = ctypes.c_int(0), ctypes.c_int64(0)
    nat.check(L.ryd_lanczos_status(nat.ptr(state), ctypes.byref(flag), ctypes.byref(nvalid), st))
    torch.cuda.synchronize()
    return (alpha.cpu().numpy(), nrm.cpu().numpy(), int(flag.value), int(nvalid.value),
            out, bufs, alpha, nrm)


def _start(plan, tdt, seed=0):
    import torch
    g = torch.Generator(device=plan.device)
    g.manual_seed(seed)
    return torch.randn(plan.dim, dtype=torch.float64, device=plan.device, generator=g).to(tdt)


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n', [1, 3, 9, 10, 12, 13])
def test_alpha_beta_match_the_oracle_recurrence(pkg, dev, n, dtype):
    plan, H, tdt = _case(pkg, n, dtype)
    v0 = _start(plan, tdt)
    D = plan.dim
    # below D the plain recurrence (no reorthogonalisation) stays within
    # rounding of the oracle's reorthogonalised one; at D = 2 both steps
    K = 2 if D == 2 else min(20, D - 1)
    a, nr, flag, nvalid, *_ = _lanczos(pkg, plan, tdt, v0, [min(7, K), K - min(7, K)])
    ra, rb, nrm0sq = orc.lanczos(lambda q: H @ q, v0.cpu().numpy().astype(complex), K)
    steps = nvalid if flag else K
    assert steps == len(ra), (steps, len(ra), flag)
    hn = abs(H).sum(axis=1).max()
    tol = 1e-8 if dtype in ('float64', 'complex128') else 2e-4
    assert abs(nr[0] ** 2 - nrm0sq) <= 1e-12 * nrm0sq
    assert np.max(np.abs(a[:steps] - ra)) <= tol * hn
    nb = len(rb)
    assert np.max(np.abs(nr[1:1 + nb] - rb), initial=0.0) <= tol * hn


@pytest.mark.parametrize('dtype', DTYPES)
def test_run_split_over_calls_equals_one_call(pkg, dev, dtype):
    plan, H, tdt = _case(pkg, 12, dtype)
    v0 = _start(plan, tdt, seed=3)
    a1, n1, f1, s1, _, b1, *_ = _lanczos(pkg, plan, tdt, v0, [60])
    a2, n2, f2, s2, _, b2, *_ = _lanczos(pkg, plan, tdt, v0, [1, 13, 20, 26])
    assert (f1, s1) == (f2, s2) == (0, 0)
    assert np.array_equal(a1[:60], a2[:60]) and np.array_equal(n1[:61], n2[:61])
    # the live vectors u_60 (b[0]) and u_59 (b[2]); b[1] is scratch
    for i in (60 % 3, 59 % 3):
        assert bool((b1[i] == b2[i]).all())


@pytest.mark.parametrize('dtype', ['float64', 'complex128', 'float32'])
def test_replayed_ritz_vector_satisfies_the_lanczos_residual_identity(pkg, dev, dtype):
    """x = Q y for the lowest Ritz pair (theta, y) of the steps' tridiagonal:
    |H x - theta x| = |beta_K y_K| and x^H H x = theta (Q orthonormal to
    rounding over 40 steps from a random start)."""
    import torch
    plan, H, tdt = _case(pkg, 11, dtype)
    v0 = _start(plan, tdt, seed=5)
    K = 40
    a, nr, flag, _, _, _, ta, tn = _lanczos(pkg, plan, tdt, v0, [K])
    assert flag == 0
    T = np.diag(a[:K]) + np.diag(nr[1:K], 1) + np.diag(nr[1:K], -1)
    w, Y = np.linalg.eigh(T)
    theta, y = w[0], Y[:, 0]
    *_, x, _, _, _ = _lanczos(pkg, plan, tdt, v0, [K], replay_y=y, alpha=ta, nrm=tn)
    xh = x.to(torch.complex128).cpu().numpy()
    hn = abs(H).sum(axis=1).max()
    tol = 1e-9 if dtype != 'float32' else 2e-4
    assert abs(np.linalg.norm(xh) - 1.0) <= tol * 10
    r = H @ xh - theta * xh
    assert abs(np.linalg.norm(r) - abs(nr[K] * y[-1])) <= tol * hn
    assert abs(np.vdot(xh, H @ xh).real - theta) <= tol * hn


@pytest.mark.parametrize('dtype', ['float64', 'complex64'])
def test_breakdown_stops_the_run_and_later_calls(pkg, dev, dtype):
    """Without drive a basis state is an eigenvector: the first step's
    residual vanishes, the run stops after one step, and a continuing call
    leaves alpha untouched.  A zero start vector stops before any step."""
    import torch
    plan, H, tdt = _case(pkg, 10, dtype, omega_scale=0.0)
    v0 = torch.zeros(plan.dim, dtype=tdt, device=plan.device)
    v0[37] = 2.0
    a, nr, flag, steps, *_ = _lanczos(pkg, plan, tdt, v0, [5, 5])
    assert (flag, steps) == (1, 1)
    assert abs(nr[0] - 2.0) < 1e-12 and nr[1] <= 1e-10
    assert abs(a[0] - H[37, 37].real) <= 1e-6 * max(1.0, abs(H[37, 37]))
    assert np.all(np.isnan(a[1:10]))
    z = torch.zeros(plan.dim, dtype=tdt, device=plan.device)
    a, nr, flag, steps, *_ = _lanczos(pkg, plan, tdt, z, [5])
    assert (flag, steps) == (1, 0) and nr[0] == 0.0 and np.all(np.isnan(a[:5]))


@pytest.mark.parametrize('dtype', DTYPES)
def test_stored_rows_equal_the_rotating_form(pkg, dev, dtype):
    """ryd_lanczos_stored over two chunks laid out as lanczos_run makes them
    (the second starts at row j-1 with u_{j-1}, u_j copied in) gives the
    alpha/beta of the rotating form bit for bit, and its rows are the
    unnormalised u_j (|row j| = nrm[j])."""
    import torch
    n = 11 if dtype == 'complex128' else 12
    plan, H, tdt = _case(pkg, n, dtype)
    nat = pkg._native
    L = nat.lib()
    D = plan.dim
    v0 = _start(plan, tdt, seed=7)
    K = 50
    a1, n1, f1, *_ = _lanczos(pkg, plan, tdt, v0, [K])
    assert f1 == 0
    alpha = torch.full((K + 1,), float('nan'), dtype=torch.float64, device=plan.device)
    nrm = torch.full((K + 2,), float('nan'), dtype=torch.float64, device=plan.device)
    state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=plan.device)
    code = nat.dtype_code(tdt)
    ws, st = plan.workspace(), nat.stream_ptr()
    A = torch.zeros((31, D), dtype=tdt, device=plan.device)
    A[0].copy_(v0)
    nat.check(L.ryd_lanczos_stored(plan.handle, code, nat.ptr(A), D, 0, 0, 30, nat.ptr(alpha),
                                   nat.ptr(nrm), nat.ptr(state), 1e-10, nat.ptr(ws), st))
    B = torch.zeros((22, D), dtype=tdt, device=plan.device)
    B[0].copy_(A[29])
    B[1].copy_(A[30])
    nat.check(L.ryd_lanczos_stored(plan.handle, code, nat.ptr(B), D, 29, 30, 20, nat.ptr(alpha),
                                   nat.ptr(nrm), nat.ptr(state), 1e-10, nat.ptr(ws), st))
    torch.cuda.synchronize()
    a2, n2 = alpha.cpu().numpy(), nrm.cpu().numpy()
    assert np.array_equal(a1[:K], a2[:K]) and np.array_equal(n1[:K + 1], n2[:K + 1])
    rows = [A[j] for j in range(31)] + [B[j - 29] for j in range(31, 51)]
    norms = np.array([float(torch.linalg.vector_norm(r.to(torch.complex128))) for r in rows])
    tol = 1e-12 if dtype in ('float64', 'complex128') else 1e-5
    assert np.max(np.abs(norms - n2[:51]) / n2[:51]) <= tol
