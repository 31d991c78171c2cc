"""The five benchmark configurations, built from seeded synthetic inputs.

Each ``run_cN`` builds its model (geometry and couplings as the reference's
model builders: chain ``(i·a, 0, 0)``, row-major square, Gaussian jitter on all
three coordinates, V = C6/r^6 over all pairs) and returns a dict of measured
quantities: timings use ``torch.cuda.synchronize`` around the measured work,
inputs already on the device.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as nat
from .model import from_positions, chain_positions, square_positions, c6_coupling, from_calib

__all__ = ['chain_calib', 'lattice_calib', 'warm_device_memory', 'run_d1', 'run_c1', 'run_c2', 'run_c3', 'run_c4', 'run_c5',
           'run_g1', 'run_k1', 'time_hv']


def chain_calib(n, c6):
    return from_positions(chain_positions(n), c6_coupling(c6), terms=('I1',))


def _sync():
    import torch
    torch.cuda.synchronize()


def lattice_calib(w, h, c6=20.0, cutoff=2.9):
    """Row-major w x h square lattice at unit spacing (site r*w + c at (c, r)),
    V = c6/r^6 for pairs within ``cutoff``."""
    pos = [(float(c), float(r), 0.0) for r in range(h) for c in range(w)]
    pairs = []
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = float(np.hypot(pos[i][0] - pos[j][0], pos[i][1] - pos[j][1]))
            if d <= cutoff:
                pairs.append({'i': i, 'j': j, 'V': c6 / d ** 6})
    return {'atoms': [{'id': i, 'pos': list(p)} for i, p in enumerate(pos)], 'pairs': pairs,
            'terms': ['I1']}


def time_hv(plan, dtype, steps=10, warmup=2, seed=0):
    """Mean kernel time (ms) of ``plan.hv`` over ``steps`` launches, HIP events
    on the current stream, x ~ N(0,1) resident on the device."""
    import torch
    g = torch.Generator(device=plan.device)
    g.manual_seed(seed)
    x = torch.randn(plan.dim, dtype=torch.float64, device=plan.device, generator=g).to(dtype)
    y = torch.empty_like(x)
    for _ in range(warmup):
        plan.hv(x, y)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        plan.hv(x, y)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def time_steps(plan, dtype, steps=10, warmup=2, seed=0):
    """Mean time (ms) of one Chebyshev step and of one Lanczos step on
    ``plan``, HIP events on the current stream, vectors resident.

    Chebyshev: y = 2·H̃x − z as one fused launch (``ryd_apply`` with z, the
    step of kpm.py:146-148).  Lanczos: ``steps`` device steps of
    ``ryd_lanczos`` (the apply with −β·u_{j−1} and the α/β sums, the
    orthogonalisation u_{j+1} = (w − α u_j)/β, the scalar finishes), as
    torch_lanczos_ground_state's Krylov loop (matrix_free.py:231-251) runs
    them.  Both compare with 3·D·b, the bytes of a step that reads two
    vectors and writes one."""
    import ctypes
    import torch
    g = torch.Generator(device=plan.device)
    g.manual_seed(seed)
    x = torch.randn(plan.dim, dtype=torch.float64, device=plan.device, generator=g).to(dtype)
    z = torch.randn(plan.dim, dtype=torch.float64, device=plan.device, generator=g).to(dtype)
    y = torch.empty_like(x)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    for _ in range(warmup):
        plan.apply(x, y, s=0.01, ch=2.0, cz=-1.0, z=z)
    a, b = ev(), ev()
    a.record()
    for _ in range(steps):
        plan.apply(x, y, s=0.01, ch=2.0, cz=-1.0, z=z)
    b.record()
    torch.cuda.synchronize()
    cheb = a.elapsed_time(b) / steps
    L = nat.lib()
    code = nat.dtype_code(dtype)
    alpha = torch.zeros(warmup + steps + 1, dtype=torch.float64, device=plan.device)
    nrm = torch.zeros(warmup + steps + 2, dtype=torch.float64, device=plan.device)
    state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=plan.device)
    tiny = 1e-6 if dtype == torch.float32 else 1e-13
    st = nat.stream_ptr()

    def run(j0, k):
        nat.check(L.ryd_lanczos(plan.handle, code, nat.ptr(x), nat.ptr(y), nat.ptr(z), j0, k,
                                nat.ptr(alpha), nat.ptr(nrm), nat.ptr(state), float(tiny), None,
                                None, None, None, 0, nat.ptr(plan.workspace()), st))
    run(0, warmup)
    a, b = ev(), ev()
    a.record()
    run(warmup, steps)
    b.record()
    torch.cuda.synchronize()
    flag, nvalid = ctypes.c_int(0), ctypes.c_int64(0)
    nat.check(L.ryd_lanczos_status(nat.ptr(state), ctypes.byref(flag), ctypes.byref(nvalid), st))
    if flag.value:
        raise RuntimeError('Lanczos timing run broke down at step %d' % nvalid.value)
    lz = a.elapsed_time(b) / steps
    bts = 3.0 * plan.dim * x.element_size()
    return dict(cheb_step_ms=cheb, cheb_step_gbps_3Db=bts / (cheb * 1e-3) / 1e9,
                lanczos_step_ms=lz, lanczos_step_gbps_3Db=bts / (lz * 1e-3) / 1e9,
                bytes_3Db=bts)


def _dense_h(m):
    """Dense complex128 H of a full-basis model with real drives and no J/A
    pairs (the input of the dense-stack workload): site i is bit n-1-i of the
    row index, diagonal Σ delta_i n_i + Σ V_ij n_i n_j, flips Ω_i/2."""
    n = int(m.n)
    om = np.asarray(m.Om, complex)
    if np.any(om.imag != 0) or m.has_pair_flips:
        raise ValueError('the dense-stack builder takes real drives without J/A pairs')
    D = 1 << n
    codes = np.arange(D)
    occ = ((codes[:, None] >> (n - 1 - np.arange(n))[None, :]) & 1).astype(float)
    diag = occ @ np.asarray(m.delta, float)
    for i, j, v in zip(m.pair_i, m.pair_j, m.pair_v):
        diag += v * occ[:, i] * occ[:, j]
    H = np.diag(diag).astype(np.complex128)
    for i in range(n):
        H[codes ^ (1 << (n - 1 - i)), codes] += 0.5 * om.real[i]
    return H


def run_d1(device='cuda', count=64, n=10, seed=0, sequential=16):
    """Dense-stack batched_ground_state (equilibrium.py:251-311 takes a
    (B, dim, dim) stack): ``count`` disordered chains at N=``n`` (jitter
    0.1a, Δ over 32 values in [−2, 4], Ω=1, C6=10) as dense complex128
    matrices of dim 2^n, solved by ``batched_ground_state_dense_hip`` (one
    CSR batch: E0 certified to 1e-9·max(1,|E0|), E1 counted with
    multiplicity by a locked batched run), timed after one warm-up call.
    Beside it the device Lanczos one instance at a time (the mirror's
    previous form) over the first ``sequential`` instances.  The dense
    builder is checked against the device plan's H·v, and E0 and the gap of
    four instances against numpy's eigvalsh."""
    import torch
    from .batch import batched_ground_state_dense_hip, disorder_ensemble
    from .lanczos import hip_lanczos_ground_state
    from .plan import CsrPlan, Plan
    models = disorder_ensemble(count=count, nmin=n, nmax=n, seed=seed)
    Hs = np.stack([_dense_h(m) for m in models])
    P = Plan(models[0], device=device)
    g = torch.Generator(device=P.device)
    g.manual_seed(seed)
    x = torch.randn(P.dim, dtype=torch.complex128, device=P.device, generator=g)
    y = P.hv(x).cpu().numpy()
    P.close()
    xh = x.cpu().numpy()
    builder_err = float(np.linalg.norm(Hs[0] @ xh - y) / np.linalg.norm(y))
    if builder_err > 1e-12:
        raise RuntimeError('dense builder disagrees with the device H·v (%.2e)' % builder_err)
    batched_ground_state_dense_hip(Hs, device=device)
    _sync()
    t0 = time.perf_counter()
    r = batched_ground_state_dense_hip(Hs, device=device)
    _sync()
    t_batch = time.perf_counter() - t0
    _sync()
    t0 = time.perf_counter()
    for b in range(min(sequential, count)):
        Pc = CsrPlan(Hs[b], device=device)
        hip_lanczos_ground_state(None, plan=Pc, device=device)
        Pc.close()
    _sync()
    t_seq = (time.perf_counter() - t0) / min(sequential, count)
    w = np.linalg.eigvalsh(Hs[:4])
    return dict(config='D1 dense-stack batched_ground_state, %d x N=%d (dim %d, complex128)'
                       % (count, n, 1 << n),
                count=count, dim=1 << n, seconds_batched=t_batch,
                solves_per_s_batched=count / t_batch,
                seconds_per_solve_sequential=t_seq, solves_per_s_sequential=1.0 / t_seq,
                speedup_vs_sequential=t_seq * count / t_batch,
                E0_err_vs_eigvalsh=float(np.max(np.abs(r['E0'][:4] - w[:, 0]))),
                gap_err_vs_eigvalsh=float(np.max(np.abs(r['gap'][:4] - (w[:, 1] - w[:, 0])))),
                builder_rel_err=builder_err)


def run_c1(device='cuda', steps=200, seed=0):
    """Plain FP64 Lanczos ground state, chain N=12, Ω=1, Δ=1.2, C6=10:
    ``steps`` device Lanczos steps from a seeded normalised N(0,1) start, E0
    from the FP64 tridiagonal, <n_i> from the replayed Ritz vector."""
    import torch
    from .plan import Plan
    from .lanczos import lanczos_run, ritz_vectors, rayleigh_residual, site_occupations, \
        _start_vector
    n = 12
    calib = chain_calib(n, 10.0)
    model = from_calib(calib, np.ones(n), -1.2 * np.ones(n))
    P = Plan(model, device=device)
    v0 = _start_vector(P.dim, seed, False, P.device)

    def solve():
        R = lanczos_run(P, v0, steps, 1, 0.0, want=2)
        x = ritz_vectors(P, R)
        E0, res = rayleigh_residual(P, x[0], R.theta[R.picks[0]])
        return R, E0, res, site_occupations(P, x[0])

    # the first call in a process also loads these kernels' code
    _sync()
    t0 = time.perf_counter()
    solve()
    _sync()
    t_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    R, E0, res, n_atom = solve()
    _sync()
    t = time.perf_counter() - t0
    return dict(config='C1 chain N=12 FP64 Lanczos', steps=int(R.j), E0=E0, residual=res,
                gap=float(R.theta[R.picks[1]] - R.theta[R.picks[0]]), n_atom=n_atom.tolist(),
                seconds=t, seconds_first_call=t_cold, calib=calib)


def run_c2(device='cuda', num=1024, nvec=16, seed=0, margin=0.01, dtypes=('float32', 'float64')):
    """KPM DOS, 5x5 square N=25, Ω=1, Δ=0.8, C6=20; bounds from FP64 Lanczos
    with a ``margin`` (1 %); ``num`` moments, ``nvec`` Rademacher probes
    (the reference stream), Jackson kernel; FP32 bulk against FP64."""
    import torch
    from .plan import Plan
    from .kpm import kpm_moments_hip, spectral_bounds, device_probes
    n = 25
    calib = from_positions(square_positions(n), c6_coupling(20.0), terms=('I1',))
    model = from_calib(calib, np.ones(n), -0.8 * np.ones(n))
    P = Plan(model, device=device)
    _sync()
    tb = time.perf_counter()
    emin, emax, c, r = spectral_bounds(P, margin=margin)
    _sync()
    t_bounds = time.perf_counter() - tb
    # probes generated once on the host (inputs), uploaded before timing
    out = dict(config='C2 square 5x5 N=25 KPM', num=num, nvec=nvec, emin=emin, emax=emax, c=c,
               r=r, bounds_seconds=t_bounds)
    mus = {}
    tp = time.perf_counter()
    probes = device_probes(P, nvec, seed)
    _sync()
    out['probe_generation_seconds_host'] = time.perf_counter() - tp
    for dt in dtypes:
        _sync()
        t0 = time.perf_counter()
        mu, _, _ = kpm_moments_hip(plan=P, num=num, nvec=nvec, seed=seed, dtype=dt, c=c, r=r,
                                   probes=probes)
        _sync()
        mus[dt] = mu
        out['seconds_' + dt] = time.perf_counter() - t0
    if 'float32' in mus and 'float64' in mus:
        err = np.abs(mus['float32'] - mus['float64'])
        out['fp32_max_abs_moment_error'] = float(err.max())
        out['fp32_mean_abs_moment_error'] = float(err.mean())
    out['mu'] = mus
    out['steps_per_probe'] = (num + 1) // 2
    out['cheb_steps_per_s_float32'] = nvec * ((num + 1) // 2) / out['seconds_float32']
    return out


def warm_device_memory(device='cuda', keep_free=8 << 30):
    """Take the device memory torch can get once (32 GiB blocks, filled and
    returned to torch's cache), leaving ``keep_free`` bytes; returns the
    seconds spent.  A solver that stores its Krylov basis in cached blocks
    then runs as a second solve in the same process does."""
    import torch
    from .lanczos import CHUNK_BYTES
    _sync()
    t0 = time.perf_counter()
    blocks = []
    while True:
        free, _ = torch.cuda.mem_get_info(device)
        if free < CHUNK_BYTES + keep_free:
            break
        try:
            b = torch.empty(CHUNK_BYTES, dtype=torch.uint8, device=device)
        except torch.OutOfMemoryError:
            break
        b.fill_(0)
        blocks.append(b)
    _sync()
    del blocks
    return time.perf_counter() - t0


def run_c3(device='cuda', seed=0, m=60, plain=True, warm=True):
    """Blockade-sector Lanczos, chain N=40 (D = F(42) = 267,914,296), Ω=1,
    Δ=1.5, V = 30/r^6 beyond nearest neighbours: FP32 bulk + FP64 refinement,
    against the plain FP64 path.  ``warm``: take the device memory into
    torch's cache first (``memory_warmup_seconds``) so that both solves store
    their Krylov basis in cached blocks, as a second solve in a process does."""
    from .lanczos import hip_lanczos_ground_state
    from .plan import Plan
    n = 40
    calib = chain_calib(n, 30.0)
    Om, delta = np.ones(n), -1.5 * np.ones(n)
    P = Plan(from_calib(calib, Om, delta), basis='chain_fib', device=device)
    out = dict(config='C3 chain N=40 blockade sector', dim=P.dim)
    if warm:
        out['memory_warmup_seconds'] = warm_device_memory(device)
    _sync()
    t0 = time.perf_counter()
    r32 = hip_lanczos_ground_state(calib, Om, delta, basis='chain_blockade', dtype='float32',
                                   m=m, seed=seed, plan=P)
    _sync()
    out['seconds_fp32_refined'] = time.perf_counter() - t0
    out['fp32_refined'] = {k: r32[k] for k in ('E0', 'gap', 'residuals', 'converged', 'krylov_dim',
                                              'precision')}
    if plain:
        _sync()
        t0 = time.perf_counter()
        r64 = hip_lanczos_ground_state(calib, Om, delta, basis='chain_blockade', dtype='float64',
                                       m=m, seed=seed, plan=P)
        _sync()
        out['seconds_fp64'] = time.perf_counter() - t0
        out['fp64'] = {k: r64[k] for k in ('E0', 'gap', 'residuals', 'converged', 'krylov_dim')}
        out['E0_rel_diff'] = abs(r32['E0'] - r64['E0']) / abs(r64['E0'])
        out['speedup_vs_fp64'] = out['seconds_fp64'] / out['seconds_fp32_refined']
    return out


def run_c4(device='cuda', count=512, steps=150, seed=0, plain=True, plain_count=128, tol=1e-9):
    """512 chain instances, N ~ U{10..20}, jitter σ = 0.1a, Δ over 32 values in
    [−2, 4], Ω=1, C6=10.  Batched: FP32 bulk cycles of ``steps`` Lanczos
    steps (one graph, autotuned tiles) and batched FP64 refinement cycles
    until each instance's two lowest Ritz pairs are within ``tol``·max(1,|E|)
    (``certified_ground_states``).  Plain: the FP64 device Lanczos per
    instance with the same certification (explicit residuals <= tol), and
    the batch machinery on one instance at a time (its cycles of 2·steps
    steps replayed as one hipGraph, 13 cycles at most), over the first
    ``plain_count`` instances (N uniform like the whole ensemble); the batch
    is compared per solve with the faster plain path among those that
    certify every instance.  All timed whole, start to certified E0 and gap; the
    autotuner runs once before the timed region, and the batched run keeps
    its Krylov basis in memory that the warm-up run left in torch's cache."""
    import torch
    from .batch import certified_ground_states, disorder_ensemble
    from .lanczos import hip_lanczos_ground_state
    from .plan import Plan
    models = disorder_ensemble(count=count, seed=seed)
    # the warm-up run also takes the stored Krylov basis into torch's cache
    # (store_basis=True); the timed run reuses it ('auto'), as a repeated
    # solve over an ensemble does
    warm = certified_ground_states(models, steps=steps, tol=tol, seed=seed, device=device,
                                   observables=False, store_basis=True)
    tb = warm['tile_bits']
    _sync()
    t0 = time.perf_counter()
    gs = certified_ground_states(models, steps=steps, tol=tol, seed=seed, device=device,
                                 observables=False, tile_bits=tb)
    _sync()
    t_batch = time.perf_counter() - t0
    out = dict(config='C4 512-instance disorder ensemble', count=count, steps_per_cycle=steps,
               tolerance=tol, ns=[int(m.n) for m in models], tile_bits=tb,
               autotune_ms=warm['autotune_ms'], seconds_batched=t_batch,
               solves_per_s_batched=count / t_batch, converged_batched=int(gs['converged'].sum()),
               bulk_converged=int(gs['bulk_converged'].sum()),
               steps_fp32_total=int(gs['steps_float32'].sum()),
               steps_fp64_total=int(gs['steps_float64'].sum()),
               cycles_fp32=int(gs.get('cycles_float32', 0)),
               cycles_fp64=int(gs.get('cycles_float64', 0)),
               converged_E0_batched=int(gs['converged_E0'].sum()),
               unconverged_batched=np.nonzero(~gs['converged'])[0].tolist(),
               E0_batched=gs['E0'].tolist())
    if plain:
        # two plain FP64 paths, one instance at a time, each certified to the
        # same tolerance: the device Lanczos (continued recurrence, host
        # convergence checks every CHECK_STEPS steps) and the same batch
        # machinery on a single instance (its whole cycle one replayed
        # hipGraph: no per-launch overhead).  The batch is compared with the
        # faster of the two.
        idx = list(range(count)) if plain_count is None else list(range(min(count, plain_count)))
        E64, conv = [], []
        _sync()
        t0 = time.perf_counter()
        for b in idx:
            P = Plan(models[b], device=device)
            r = hip_lanczos_ground_state(None, plan=P, dtype='float64', tol=tol, m=150,
                                         max_restarts=8, seed=seed)
            E64.append(float(r['E0']))
            conv.append(bool(r['converged']))
            P.close()
        _sync()
        t_plain = time.perf_counter() - t0
        Eg, convg = [], []
        _sync()
        t0 = time.perf_counter()
        for b in idx:
            g1 = certified_ground_states([models[b]], steps=2 * steps, dtype='float64', tol=tol,
                                         seed=seed, device=device, observables=False,
                                         tile_bits=tb, bulk_cycles=9)
            Eg.append(float(g1['E0'][0]))
            convg.append(bool(g1['converged'][0]))
        _sync()
        t_graph = time.perf_counter() - t0
        E64 = np.asarray(E64)
        rel = np.abs(gs['E0'][idx] - E64) / np.abs(E64)
        relg = np.abs(np.asarray(Eg) - E64) / np.abs(E64)
        # only a plain path that certifies every instance is a comparison at
        # the same accuracy
        cands = [(t, name) for t, name, ok in
                 ((t_plain, 'plain device Lanczos', all(conv)),
                  (t_graph, 'graph-captured single-instance batch', all(convg))) if ok]
        best, against = min(cands) if cands else (t_plain, 'plain device Lanczos (unconverged)')
        out.update(seconds_plain_fp64=t_plain, plain_count=len(idx), converged_plain=int(sum(conv)),
                   solves_per_s_plain_fp64=len(idx) / t_plain,
                   seconds_plain_graph_fp64=t_graph, converged_plain_graph=int(sum(convg)),
                   solves_per_s_plain_graph_fp64=len(idx) / t_graph,
                   plain_graph_vs_plain_E0_rel_max=float(relg.max()),
                   speedup_vs_plain_fp64=(count / t_batch) / (len(idx) / best),
                   speedup_against=against,
                   E0_rel_diff_max=float(rel.max()), E0_rel_diff_median=float(np.median(rel)))
    return out


def run_c5(device='cuda', n=20, slices=200, dt=0.05, K=15, seed=0, fd_components=(0, 99, 199),
           units='natural'):
    """Differentiable pulse: chain N=20, 200 slices of dt=0.05 ns, per-slice
    Ω_k ~ U(0.5,1.5), Δ_k ~ U(−1,2), C6=10, 16-term Chebyshev slices, loss
    1 − |<Z2|ψ(T)>|², adjoint gradient of all 400 parameters (complex64,
    checkpoints every 20 slices), against a plain FP64 forward and FP64
    central differences.  ``units='natural'`` takes dt as the phase
    exp(−iH·dt) (Ω·T = 10, a non-trivial Z2 overlap; about four 16-term
    sub-steps per slice).  ``units='reference'`` is the reference propagator
    exp(−i·CONV·H·dt): 10 ns in all, one 16-term series per slice, but each
    atom's excitation amplitude stays near CONV·Ω·T/2 ≈ 0.03, so the Z2
    overlap is ~1e-15 and the loss 1 − O(1e-30): the gradient is below FP64
    resolution and the finite-difference checks carry no information there."""
    import torch
    from .evolve import PulseProblem, z2_index
    from .plan import Plan
    calib = chain_calib(n, 10.0)
    model = from_calib(calib, np.ones(n), np.zeros(n))
    rng = np.random.default_rng(seed)
    Om = rng.uniform(0.5, 1.5, slices)
    De = rng.uniform(-1.0, 2.0, slices)
    P = Plan(model, device=device)
    p32 = PulseProblem(model, dt=dt, K=K, dtype='complex64', plan=P, units=units)
    p64 = PulseProblem(model, dt=dt, K=K, dtype='complex128', plan=P, units=units)
    psi0 = torch.zeros(P.dim, dtype=torch.complex128, device=P.device)
    psi0[0] = 1.0                         # |g...g>
    tgt = z2_index(n)
    p32.loss_and_grad(psi0, Om, De, tgt)  # warm-up
    _sync()
    t0 = time.perf_counter()
    L32, gO32, gD32, _ = p32.loss_and_grad(psi0, Om, De, tgt)
    _sync()
    t_grad = time.perf_counter() - t0
    p64.propagate(psi0, Om, De)
    _sync()
    t0 = time.perf_counter()
    psi = p64.propagate(psi0, Om, De)
    _sync()
    t_fwd64 = time.perf_counter() - t0
    L64, gO64, gD64, _ = p64.loss_and_grad(psi0, Om, De, tgt)
    fd = []
    h = 1e-4
    for k in fd_components:
        e = np.zeros(slices)
        e[k] = h

        def loss(o, d):
            ps = p64.propagate(psi0, o, d)
            return 1.0 - abs(ps[tgt].item()) ** 2
        fo = (loss(Om + e, De) - loss(Om - e, De)) / (2 * h)
        fdd = (loss(Om, De + e) - loss(Om, De - e)) / (2 * h)
        fd.append(dict(k=int(k), fd_dOm=fo, adj64_dOm=float(gO64[k]), adj32_dOm=float(gO32[k]),
                       fd_dDelta=fdd, adj64_dDelta=float(gD64[k]), adj32_dDelta=float(gD32[k])))
    g64 = np.concatenate([gO64, gD64])
    g32 = np.concatenate([gO32, gD32])
    p64._slice_params(Om, De)
    return dict(config='C5 N=20 pulse adjoint', slices=slices, K=K, units=units,
                substeps_total=int(len(p64._owner)), loss_fp64=L64, loss_c64=L32,
                seconds_grad_c64=t_grad, grad_per_s=1.0 / t_grad, seconds_forward_fp64=t_fwd64,
                grad_cost_in_fp64_forwards=t_grad / t_fwd64,
                grad_c64_vs_fp64_max_rel=float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))),
                fd_checks=fd)


def run_c5r(device='cuda', **kw):
    """The C5 pulse in the reference's units (exp(−i·CONV·H·dt)), timed
    only: see run_c5 for why its gradient carries no information."""
    r = run_c5(device, units='reference', fd_components=(), **kw)
    r['config'] = 'C5 N=20 pulse adjoint, reference units (timing)'
    return r


def run_g1(device='cuda', w=6, h=7, diag=False, steps=10, seed=0, ground_state=True):
    """Blockade-graph basis: w x h square lattice (6 x 7: N = 42), V = 20/r^6
    within 2.9a, blockade edges |V| >= 10 MHz (nearest neighbours; with
    ``diag`` also the diagonals, threshold 2 MHz), Ω = 1, Δ = 1.5.  The
    basis (D = 69,050,253 for 6 x 7) is enumerated on the device; reports the
    plan build time, the H·v time in FP64 and FP32 (2·D·b algorithmic bytes)
    and the FP64 ground state."""
    from .basis import blockade_edges
    from .lanczos import hip_lanczos_ground_state
    from .plan import Plan
    n = w * h
    calib = lattice_calib(w, h)
    edges = blockade_edges(calib, {'kind': 'blockade', 'threshold_mhz': 2.0 if diag else 10.0})
    Om, delta = np.ones(n), -1.5 * np.ones(n)
    _sync()
    t0 = time.perf_counter()
    P = Plan(from_calib(calib, Om, delta), basis='graph', edges=edges, device=device)
    _sync()
    out = dict(config='G1 %dx%d lattice blockade graph%s' % (w, h, ' (diagonals)' if diag else ''),
               n_sites=n, edges=len(edges), dim=P.dim, rank_mode=P.rank_mode,
               build_seconds=time.perf_counter() - t0)
    import torch
    for name, dt, b in (('f64', torch.float64, 8), ('f32', torch.float32, 4)):
        ms = time_hv(P, dt, steps=steps, seed=seed)
        out['hv_ms_' + name] = ms
        out['hv_gbps_' + name] = 2.0 * P.dim * b / (ms * 1e-3) / 1e9
    if ground_state:
        _sync()
        t0 = time.perf_counter()
        r = hip_lanczos_ground_state(calib, Om, delta, blockade_edges=edges, dtype='float64',
                                     seed=seed, plan=P, basis='graph')
        _sync()
        out['seconds_fp64_gs'] = time.perf_counter() - t0
        out['fp64'] = {k: r[k] for k in ('E0', 'gap', 'residuals', 'converged', 'krylov_dim')}
    P.close()
    return out


def run_k1(device='cuda', calib=None, steps=10, seed=0, ground_state=True):
    """Exchange (J) channels: the dipolar XY kagome calib with 24 atoms held
    by the reference's checks (checks/fixtures/kagome_n24.json; 276 pairs
    with J, per-atom delta_mhz), Ω = 0.  Reports the H·v time of the
    full-basis plan (D = 2^24) in FP64 and FP32, and the FP64 ground state of
    the half-filled number sector (12 excitations, D = 2,704,156) on a code
    list (the sector scan's basis)."""
    import json
    import os
    import torch
    from .basis import number_sector_basis
    from .lanczos import hip_lanczos_ground_state
    from .plan import Plan
    if calib is None:
        here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        with open(os.path.join(here, 'tests', 'golden', 'kagome_n24.json')) as f:
            calib = json.load(f)
    n = len(calib['atoms'])
    delta = np.array([float(a.get('delta_mhz', 0.0)) for a in calib['atoms']])
    Om = np.zeros(n)
    model = from_calib(calib, Om, delta)
    P = Plan(model, device=device)
    # off-diagonal entries per H·v: an exchange pair acts on the half of the
    # codes where its two bits differ, a pair-creation pair on the other half
    nj = int(np.count_nonzero(np.abs(np.asarray(model.pair_J)) > 0))
    na = int(np.count_nonzero(np.abs(np.asarray(model.pair_A)) > 0))
    offdiag = (nj + na) * (P.dim // 2)
    out = dict(config='K1 kagome N=%d XY (J channels)' % n, n_sites=n,
               pairs=len(calib['pairs']), dim=P.dim, offdiag_per_hv=offdiag)
    for name, dt, b in (('f64', torch.float64, 8), ('f32', torch.float32, 4)):
        ms = time_hv(P, dt, steps=steps, seed=seed)
        out['hv_ms_' + name] = ms
        out['hv_gbps_' + name] = 2.0 * P.dim * b / (ms * 1e-3) / 1e9
        out['offdiag_per_s_' + name] = offdiag / (ms * 1e-3)
    P.close()
    if ground_state:
        codes = number_sector_basis(n, n // 2)
        _sync()
        t0 = time.perf_counter()
        r = hip_lanczos_ground_state(calib, Om, delta, basis_indices=codes, dtype='float64',
                                     seed=seed, device=device)
        _sync()
        out['sector'] = n // 2
        out['sector_dim'] = int(len(codes))
        out['seconds_fp64_gs'] = time.perf_counter() - t0
        out['fp64'] = {k: r[k] for k in ('E0', 'gap', 'residuals', 'converged', 'krylov_dim')}
    return out
