"""Matrix-free Lanczos ground state on the HIP engine.

Drop-in for ``torch_lanczos_ground_state``
(tomisnot/energy-level src/core/manybody/matrix_free.py:118-286): same
arguments, same result keys, backend ``'hip-matrixfree-lanczos'``.

Algorithm (not the reference's block Krylov; only converged outputs are
compared):

* the three-term recurrence runs on the device with FP64 alpha/beta and no
  host synchronisation between convergence checks (every 10 steps,
  ``ryd_lanczos``); the recurrence is continued, not restarted, up to
  ``m·(max_restarts+1)`` steps;
* after each cycle the FP64 tridiagonal matrix is diagonalised on the host;
  duplicate and Cullum–Willoughby spurious Ritz values are dropped; the two
  lowest genuine ones are tested with the bound |beta_k · y_k[-1]|;
* the Krylov basis can stay in HBM (blocks of rows; automatically for
  dim >= 2^22 when torch's allocator holds unused blocks, e.g. from a previous
  solve) and a Ritz vector is then one streaming combination of it;
  otherwise the recurrence keeps three vectors and Ritz vectors come from a
  replay with the stored coefficients (only the unstored tail when storage
  stopped partway);
* FP32 bulk (``dtype='float32'``/``'complex64'``, ``refine=True``): the FP32
  Ritz pair seeds a short FP64 recurrence, so E0 and the residuals are
  certified in FP64 (``precision`` in the result says which parts ran in
  which precision);
* each returned energy is the FP64 Rayleigh quotient of its Ritz vector and
  ``residuals`` are the explicit |H x − E x| of the normalised vectors
  (certification rule of matrix_free.py:258-266: residual <= tol·max(1,|E|)).
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .model import from_calib
from .plan import Plan, cast

__all__ = ['hip_lanczos_ground_state', 'lanczos_run', 'ritz_vectors', 'rayleigh_residual',
           'site_occupations', 'restricted_plan', 'lanczos_vector',
           'hamiltonian_linear_operator_hip', 'ground_is_unique', 'locked_lowest']


def _tridiag(a, b):
    """Eigenpairs of the FP64 tridiagonal (alpha a, beta b); LAPACK's MRRR
    driver can fail on off-diagonals near underflow after a breakdown, then
    the dense symmetric solver takes over."""
    from scipy.linalg import eigh_tridiagonal
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if len(a) == 1:
        return np.array([a[0]]), np.ones((1, 1))
    try:
        return eigh_tridiagonal(a, b)
    except np.linalg.LinAlgError:
        return np.linalg.eigh(np.diag(a) + np.diag(b, 1) + np.diag(b, -1))


def _genuine_lowest(a, b, theta, want=2, bounds=None, conv_tol=None, that=None, merge_rel=1e-10):
    """Indices (into theta) of the ``want`` lowest distinct, non-spurious Ritz values.

    Duplicates (ghost copies of converged values, equal to ``merge_rel``·|T|:
    the rounding of the recurrence, 1e-10 for FP64 vectors and 1e-6 for FP32
    ones) merge; a simple value that
    is also an eigenvalue of T with its first row/column removed is spurious
    (Cullum–Willoughby), with two exceptions that test cannot see:

    * the lowest Ritz value is always kept.  A spurious value is a copy in
      the making of an eigenvalue that has already converged, so it never
      lies below the lowest genuine one; the test itself misfires on a
      genuine value whose eigenvector has a tiny component in the start
      vector, and dropping the lowest value then reports the first excited
      level as E0;
    * a value whose residual bound ``bounds[i]`` (|beta_k · y_k[-1]|) is at
      most ``conv_tol``·max(1, |theta_i|) has converged, so it lies within
      that bound of an eigenvalue of H (Paige) and is kept.

    ``theta`` may hold only the lowest eigenvalues of T, with ``that`` the
    lowest eigenvalues of T-hat that interlace them."""
    from scipy.linalg import eigvalsh_tridiagonal
    k = len(a)
    scale = max(1.0, float(np.max(np.abs(theta))))
    if that is not None and k > 1:
        # only the lowest values are at hand: bound |T| by Gershgorin
        scale = max(scale, float(np.max(np.abs(a)) + 2.0 * np.max(np.abs(b))))
    tol = max(merge_rel * scale, 100.0 * k * np.finfo(float).eps * scale)
    if that is not None:
        pass
    elif k > 2:
        that = eigvalsh_tridiagonal(a[1:], b[1:])
    elif k == 2:
        that = a[1:]
    else:
        that = np.zeros(0)
    order = np.argsort(theta)
    kept = []
    for i in order:
        th = theta[i]
        if kept and abs(th - theta[kept[-1]]) <= tol:
            continue
        mult = int(np.sum(np.abs(theta - th) <= tol))
        if kept and mult == 1 and that.size and np.min(np.abs(that - th)) <= tol:
            converged = (bounds is not None and conv_tol is not None and
                         bounds[i] <= conv_tol * max(1.0, abs(th)))
            if not converged:
                continue
        kept.append(int(i))
        if len(kept) == want:
            break
    return kept


def _ritz_bounds(Y, beta_last):
    """|beta_k · y_k[-1]| of every Ritz pair of the k-step tridiagonal."""
    return np.abs(float(beta_last) * np.asarray(Y)[-1, :])


class _Run:
    pass


STORE_MIN_DIM = 1 << 22     # below this the replay of a cycle costs less than a stored basis
RITZ_ROWS = 2048            # coefficients per ryd_ritz_stored call
CHUNK_BYTES = 32 << 30      # stored-basis block: one size for every element type, so torch's
                            # cache reuses the blocks of an FP32 run for the FP64 refinement


CHECK_STEPS = 10            # steps between convergence checks (the budget stays m·cycles)


def _fits(plan, nbytes, dt, cached_only=False):
    """True when ``nbytes`` more device memory leave the working vectors room
    (blocks cached but unused by torch's allocator count as free).
    ``cached_only``: the bytes must come from those cached blocks.  The
    automatic mode stores only there: on MI355X a run that allocated
    hundreds of GB of new blocks (130 GB requests after releasing ~200 GB
    took 1.9-3.2 s each) lost more than the replay of its Ritz vectors
    costs, while blocks held in torch's cache come for free."""
    import torch
    itemsize = torch.tensor([], dtype=dt).element_size()
    cached = torch.cuda.memory_reserved(plan.device) - torch.cuda.memory_allocated(plan.device)
    if cached_only and nbytes > cached:
        return False
    free, _ = torch.cuda.mem_get_info(plan.device)
    return nbytes + 8 * plan.dim * itemsize + (4 << 30) <= free + cached


def lanczos_run(plan, v0, m, max_cycles, tol, want=2, tiny=None, store='auto', top=False,
                check_steps=None):
    """Cycles of ``m`` device-side Lanczos steps from ``v0`` until the ``want``
    lowest genuine Ritz values pass the bound test or the budget is spent.

    ``store``: keep every Lanczos vector in HBM (chunks of rows: a whole cycle
    when it is small, else fixed 32 GiB blocks whose spare rows the next cycle
    uses) so that Ritz vectors are one streaming combination instead of a
    replay of the recurrence.  True: allocate as memory allows; 'auto'
    (dim >= 2^22): only into blocks torch's allocator already holds unused
    (a repeated solve reuses the blocks of the previous one), and always for
    the one-launch runs of small full-basis plans (``Plan.lanczos_one_launch``:
    a few MB, and a replay would repeat every step); False: three rotating
    vectors.  When storage stops, the rest of the run continues with three
    vectors and only that tail is replayed.

    Convergence is tested every ``check_steps`` (default ``CHECK_STEPS`` from
    dim = STORE_MIN_DIM on, twice that below; at most ``m``) steps; the
    step budget is ``m·max_cycles``.  ``top``: the highest Ritz value must
    pass the same bound test as the ``want`` lowest (spectral bounds)."""
    import ctypes
    import torch
    L = nat.lib()
    dim = plan.dim
    dt = v0.dtype
    itemsize = torch.tensor([], dtype=dt).element_size()
    row_bytes = dim * itemsize
    kmax = int(min(max(1, m) * max(1, max_cycles), dim))
    R = _Run()
    R.v0 = v0
    R.alpha = torch.zeros(kmax + 1, dtype=torch.float64, device=plan.device)
    R.nrm = torch.zeros(kmax + 2, dtype=torch.float64, device=plan.device)
    R.state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=plan.device)
    R.chunks = []                 # (jbase, rows tensor) of a stored basis
    R.bufs = None
    R.split = None                # first step run without storage (None: all stored)
    # one-launch runs (small full-basis plans) store their few MB: a Ritz
    # vector is then one combination instead of a replay of every step
    one = bool(getattr(plan, 'lanczos_one_launch', None) and plan.lanczos_one_launch(dt))
    storing = (dim >= STORE_MIN_DIM or one) if store == 'auto' else bool(store)
    cached_only = store == 'auto' and not one
    # small problems are bound by the host round trip of a check, large ones
    # by their steps: check less often below STORE_MIN_DIM
    default_check = CHECK_STEPS if dim >= STORE_MIN_DIM else 2 * CHECK_STEPS
    check = max(1, min(int(m), int(check_steps or default_check)))
    if not (float(tol) > 0.0):
        check = max(1, int(m))   # nothing can pass a zero tolerance: test once per cycle
    code = nat.dtype_code(dt)
    ws = plan.workspace()
    st = nat.stream_ptr()
    if tiny is None:
        tiny = 1e-13 if dt in (torch.float64, torch.complex128) else 1e-6
    # ghost copies of a value agree to the recurrence's own rounding
    merge_rel = 1e-10 if dt in (torch.float64, torch.complex128) else 1e-6

    def three_buffers(j):
        """Switch to (or start) the rotating three-vector recurrence at step j;
        the stored chunks stay (they hold u_0 .. u_j for the Ritz vectors)."""
        if R.bufs is None:
            R.bufs = [torch.empty(dim, dtype=dt, device=plan.device) for _ in range(3)]
        if R.chunks:
            jb, rows = R.chunks[-1]
            R.bufs[j % 3].copy_(rows[j - jb])
            if j > 0:
                R.bufs[(j - 1) % 3].copy_(rows[j - 1 - jb])
        elif j == 0:
            R.bufs[0].copy_(v0)

    def new_chunk(j, want_rows):
        """A chunk starting at row jb = j-1 (0 at j = 0) holding u_{jb}, u_j;
        None when memory does not allow it."""
        jb = j - 1 if j > 0 else 0
        nbytes = want_rows * row_bytes
        if nbytes > CHUNK_BYTES:             # one block size for every element type
            nbytes = CHUNK_BYTES if 3 * row_bytes <= CHUNK_BYTES else 3 * row_bytes
        if not _fits(plan, nbytes, dt, cached_only):
            return None
        try:
            buf = torch.empty(nbytes, dtype=torch.uint8, device=plan.device)
        except torch.OutOfMemoryError:
            return None
        cap = (nbytes // row_bytes)
        rows = buf[:cap * row_bytes].view(dt).view(cap, dim)
        if j == 0:
            rows[0].copy_(v0)
        else:
            pjb, prev = R.chunks[-1]
            rows[0].copy_(prev[j - 1 - pjb])
            rows[1].copy_(prev[j - pjb])
        R.chunks.append((jb, rows))
        return jb, rows

    j, cycles = 0, 0
    stopped = False
    R.stored = storing
    while True:
        steps = int(min(check, kmax - j))
        if storing:
            left, jj = steps, j
            while left > 0:
                cur = R.chunks[-1] if R.chunks else None
                room = cur[0] + cur[1].shape[0] - 1 - jj if cur is not None else 0
                if room <= 0:
                    cur = new_chunk(jj, min(left, kmax - jj) + 2)
                    if cur is None:
                        break
                    room = cur[0] + cur[1].shape[0] - 1 - jj
                k = min(left, room)
                jb, rows = cur
                nat.check(L.ryd_lanczos_stored(plan.handle, code, nat.ptr(rows), dim, jb, jj, k,
                                               nat.ptr(R.alpha), nat.ptr(R.nrm), nat.ptr(R.state),
                                               float(tiny), nat.ptr(ws), st))
                jj += k
                left -= k
            if left > 0:
                # memory ran out inside the cycle: keep what is stored, go on with three vectors
                storing = False
                R.split = jj
                three_buffers(jj)
                nat.check(L.ryd_lanczos(plan.handle, code, nat.ptr(R.bufs[0]), nat.ptr(R.bufs[1]),
                                        nat.ptr(R.bufs[2]), jj, left, nat.ptr(R.alpha),
                                        nat.ptr(R.nrm), nat.ptr(R.state), float(tiny), None,
                                        None, None, None, 0, nat.ptr(ws), st))
        else:
            if R.bufs is None:
                three_buffers(j)
            nat.check(L.ryd_lanczos(plan.handle, code, nat.ptr(R.bufs[0]), nat.ptr(R.bufs[1]),
                                    nat.ptr(R.bufs[2]), j, steps, nat.ptr(R.alpha), nat.ptr(R.nrm),
                                    nat.ptr(R.state), float(tiny), None, None, None, None, 0,
                                    nat.ptr(ws), st))
        cycles += 1
        sflag, nvalid = ctypes.c_int(0), ctypes.c_int64(0)
        nat.check(L.ryd_lanczos_status(nat.ptr(R.state), ctypes.byref(sflag),
                                       ctypes.byref(nvalid), st))
        if sflag.value:
            stopped = True
            j = int(nvalid.value)
        else:
            j += steps
        if j == 0:
            raise ValueError('Lanczos start vector is zero')
        a = R.alpha[:j].cpu().numpy()
        nr = R.nrm[:j + 1].cpu().numpy()
        b = nr[1:j]
        theta, Y = _tridiag(a, b)
        blast = 0.0 if stopped else float(nr[j])
        picks = _genuine_lowest(a, b, theta, want, _ritz_bounds(Y, blast),
                                max(float(tol), 1e-12), merge_rel=merge_rel)
        est = [abs(blast * Y[-1, i]) for i in picks]
        ok = all(e <= tol * max(1.0, abs(theta[i])) for e, i in zip(est, picks))
        if top:
            hi = int(np.argmax(theta))
            ok = ok and abs(blast * Y[-1, hi]) <= tol * max(1.0, abs(theta[hi]))
        if ok or stopped or j >= kmax:
            break
    R.j, R.theta, R.Y, R.picks, R.est, R.stopped, R.cycles = j, theta, Y, picks, est, stopped, cycles
    R.a, R.b = a, b
    return R


def lanczos_vector(R, j):
    """u_j (unnormalised, |u_j| = R.nrm[j]) when the run still holds it, else None:
    from the stored chunks, or from the three rotating vectors for the last
    steps of a run that was not (or no longer) stored."""
    if R.bufs is not None and (not R.chunks or (R.split is not None and j >= R.split)):
        return R.bufs[j % 3] if j >= R.j - 1 else None
    for jb, rows in reversed(R.chunks):
        if jb <= j < jb + rows.shape[0] and (R.split is None or j <= R.split):
            return rows[j - jb]
    return None


def ritz_vectors(plan, R, which=None):
    """Up to two Ritz vectors: one streaming combination of the stored
    basis, or a replay of the recurrence of ``R`` that accumulates them."""
    import torch
    L = nat.lib()
    which = R.picks if which is None else which
    which = list(which)[:2]
    dt = R.v0.dtype
    out = [torch.zeros(plan.dim, dtype=dt, device=plan.device) for _ in which]
    if R.stored and R.chunks:
        ys = [np.ascontiguousarray(R.Y[:, i], dtype=np.float64) for i in which]
        first = True
        jend = R.j if R.split is None else min(R.split, R.j)   # stored rows used: j < jend
        for jb, rows in R.chunks:
            j0 = jb if jb == 0 else jb + 1          # rows 0/1 repeat the previous chunk
            j1 = min(jb + rows.shape[0] - 1, jend)
            for a in range(j0, j1, RITZ_ROWS):      # the kernel keeps <= RITZ_ROWS coefficients
                b = min(a + RITZ_ROWS, j1)
                base = rows[a - jb]
                y0 = np.ascontiguousarray(ys[0][a:b])
                y1 = np.ascontiguousarray(ys[1][a:b]) if len(ys) > 1 else None
                nat.check(L.ryd_ritz_stored(nat.dtype_code(dt), nat.ptr(base), plan.dim, plan.dim,
                                            b - a, y0.ctypes.data,
                                            y1.ctypes.data if y1 is not None else None,
                                            nat.ptr(R.nrm), a, 0 if first else 1, nat.ptr(out[0]),
                                            nat.ptr(out[1]) if len(out) > 1 else None,
                                            nat.ptr(plan.workspace()), nat.stream_ptr()))
                first = False
        if R.split is not None and R.split < R.j:
            # the steps run without storage: replay them from the stored u_{split-1}, u_split
            sp = R.split
            jb, rows = [c for c in R.chunks if c[0] <= sp < c[0] + c[1].shape[0]][-1]
            B = [torch.empty(plan.dim, dtype=dt, device=plan.device) for _ in range(3)]
            B[sp % 3].copy_(rows[sp - jb])
            if sp > 0:
                B[(sp - 1) % 3].copy_(rows[sp - 1 - jb])
            yt = [torch.as_tensor(y, dtype=torch.float64, device=plan.device) for y in ys]
            nat.check(L.ryd_lanczos(plan.handle, nat.dtype_code(dt), nat.ptr(B[0]), nat.ptr(B[1]),
                                    nat.ptr(B[2]), sp, R.j - sp, nat.ptr(R.alpha), nat.ptr(R.nrm),
                                    nat.ptr(R.state), 0.0,
                                    nat.ptr(out[0]) if len(out) > 0 else None,
                                    nat.ptr(out[1]) if len(out) > 1 else None,
                                    nat.ptr(yt[0]) if len(yt) > 0 else None,
                                    nat.ptr(yt[1]) if len(yt) > 1 else None, 1,
                                    nat.ptr(plan.workspace()), nat.stream_ptr()))
        return out
    if R.bufs is None:
        R.bufs = [torch.empty(plan.dim, dtype=dt, device=plan.device) for _ in range(3)]
    R.bufs[0].copy_(R.v0)
    ys = [torch.as_tensor(np.ascontiguousarray(R.Y[:, i]), dtype=torch.float64,
                          device=plan.device) for i in which]
    nat.check(L.ryd_lanczos(plan.handle, nat.dtype_code(dt), nat.ptr(R.bufs[0]),
                            nat.ptr(R.bufs[1]), nat.ptr(R.bufs[2]), 0, R.j, nat.ptr(R.alpha),
                            nat.ptr(R.nrm), nat.ptr(R.state), 0.0,
                            nat.ptr(out[0]) if len(out) > 0 else None,
                            nat.ptr(out[1]) if len(out) > 1 else None,
                            nat.ptr(ys[0]) if len(ys) > 0 else None,
                            nat.ptr(ys[1]) if len(ys) > 1 else None, 1,
                            nat.ptr(plan.workspace()), nat.stream_ptr()))
    return out


def rayleigh_residual(plan, x, theta=0.0):
    """(E, |H x̂ − E x̂|) for x̂ = x/|x|, E the Rayleigh quotient, in x's precision
    with FP64 sums."""
    import torch
    sums = torch.zeros(4, dtype=torch.float64, device=plan.device)
    plan.dots(None, x, sums)
    nx = float(np.sqrt(sums[1].item()))
    if not nx > 0:
        raise ValueError('zero vector')
    y = torch.empty_like(x)
    plan.apply(x, y, s=1.0 / nx, ch=1.0, cx=-float(theta), sums=sums)
    r0, r1, r2 = (float(v) for v in sums[:3].cpu().numpy())
    shift = r0 / r2
    res = float(np.sqrt(max(r1 / r2 - shift * shift, 0.0)))
    return float(theta) + shift, res


def site_occupations(plan, x):
    """<n_i> per site for a full-basis state (normalised inside)."""
    import torch
    out = torch.zeros(plan.n + 1, dtype=torch.float64, device=plan.device)
    nat.check(nat.lib().ryd_site_occupation(plan.handle, nat.dtype_code(x.dtype), nat.ptr(x),
                                            nat.ptr(out), nat.ptr(plan.workspace()),
                                            nat.stream_ptr()))
    o = out.cpu().numpy()
    return o[:plan.n] / o[plan.n]


def _combine(vecs, dtype):
    """x0/|x0| + x1/|x1| in ``dtype`` (the restart / refinement seed)."""
    import torch
    start = None
    for v in vecs:
        v = v if v.dtype == dtype else cast(v, dtype)
        nv = torch.linalg.vector_norm(v)
        start = v / nv if start is None else start + v / nv
    return start


def _certify(P, vecs, R, dtype64):
    evals, res = [], []
    for v, i in zip(vecs, R.picks):
        vv = v if v.dtype == dtype64 else cast(v, dtype64)
        e, r = rayleigh_residual(P, vv, R.theta[i])
        evals.append(e)
        res.append(r)
    return np.asarray(evals), np.asarray(res)


def _solve(P, start, m, cycles, tol, want, stats, key, dtype64, check_steps=None, attempts=None):
    """Lanczos runs with explicit restarts from the Ritz pair until the explicit
    residuals of the ``want`` lowest Ritz pairs pass ``tol``·max(1,|E|) (at
    most ``attempts`` runs, default ``cycles``)."""
    for attempt in range(int(attempts or cycles)):
        R = lanczos_run(P, start, m, cycles, tol, want=want, check_steps=check_steps)
        vecs = ritz_vectors(P, R)
        R.chunks, R.bufs = [], None          # the stored basis is not needed past here
        stats[key] += R.j
        stats['restarts'] += max(0, -(-R.j // m) - 1) + (1 if attempt else 0)   # m-step cycles
        evals, res = _certify(P, vecs, R, dtype64)
        if len(evals) == want and np.all(res <= tol * np.maximum(1.0, np.abs(evals))):
            break
        start = _combine(vecs, start.dtype)
    return vecs, R, evals, res


def ground_is_unique(P):
    """True when the lowest eigenvalue is simple by Perron-Frobenius: a real
    model without pair flips and every Ω_i nonzero, on the full basis, the
    chain sector or a blockade-graph basis (single flips connect every basis
    state to the empty one, and a diagonal ±1 gauge makes every off-diagonal
    element -|Ω_i|/2 <= 0)."""
    m = getattr(P, 'model', None)
    if m is None or getattr(P, 'basis', None) not in ('full', 'chain_fib', 'graph'):
        return False
    om = np.asarray(m.Om)
    return bool(m.is_real and not m.has_pair_flips and om.size and np.all(om.real != 0))


LOCK_CHECK = 25            # steps between convergence tests of a locked run
LOCK_MAX_STEPS = 3000      # step budget of a locked run


def locked_lowest(P, x0, tol, seed=1, budget_bytes=None, max_steps=None):
    """Lowest eigenpair of H on the orthogonal complement of x0: the device
    three-term recurrence locked against x0 (``ryd_lanczos_locked``: each
    step removes the x0 component of the finished residual, so the run stays
    in the complement at rounding level), tested every LOCK_CHECK steps with
    the ghost filter of ``lanczos_run`` until its lowest genuine Ritz pair
    passes the bound test ``tol``·max(1,|θ|) or ``max_steps`` are spent, then
    that pair's Ritz vector (replayed) with its FP64 Rayleigh quotient and
    explicit residual.  With x0 the ground vector this is the second level
    counted with multiplicity: E0 again when the ground state is degenerate,
    as the reference's two-vector block Krylov (matrix_free.py:219-266) and
    eigh (equilibrium.py:65-75, :272) report.  Three vectors and O(steps)
    device memory; a fully reorthogonalised locked run
    (``ryd_lanczos_reortho_locked``) costs O(steps²·dim) and took 3.9 s for
    the kagome N=24 half-filled sector (0.1 s this way).  Returns
    (E1, residual, x1), or None when H has no complement (dim 1).
    ``budget_bytes`` is accepted for compatibility."""
    import ctypes
    import torch
    L = nat.lib()
    dim = P.dim
    if dim < 2:
        return None
    dt = x0.dtype
    code = nat.dtype_code(dt)
    g = (x0 / torch.linalg.vector_norm(x0)).contiguous()
    gen = torch.Generator(device=P.device)
    gen.manual_seed(int(seed))
    phi = torch.randn(dim, dtype=torch.float64, device=P.device, generator=gen).to(dt)
    if dt.is_complex:
        phi = phi + 1j * torch.randn(dim, dtype=torch.float64, device=P.device,
                                     generator=gen).to(dt)
    phi = phi - g * torch.vdot(g, phi)
    kmax = int(min(int(max_steps or LOCK_MAX_STEPS), dim - 1))
    alpha = torch.zeros(kmax + 1, dtype=torch.float64, device=P.device)
    nrm = torch.zeros(kmax + 2, dtype=torch.float64, device=P.device)
    gam = torch.zeros(2 * kmax + 2, dtype=torch.float64, device=P.device)
    state = torch.zeros(int(L.ryd_lanczos_state_bytes()), dtype=torch.uint8, device=P.device)
    ws = P.workspace()
    st = nat.stream_ptr()
    tiny = 1e-13 if dt in (torch.float64, torch.complex128) else 1e-6
    merge_rel = 1e-10 if dt in (torch.float64, torch.complex128) else 1e-6
    B = [phi.clone(), torch.empty_like(phi), torch.empty_like(phi)]
    j, stopped = 0, False
    while True:
        k = min(LOCK_CHECK, kmax - j)
        nat.check(L.ryd_lanczos_locked(P.handle, code, nat.ptr(B[0]), nat.ptr(B[1]), nat.ptr(B[2]),
                                       nat.ptr(g), j, k, nat.ptr(alpha), nat.ptr(nrm), nat.ptr(gam),
                                       nat.ptr(state), float(tiny), None, None, None, None, 0,
                                       nat.ptr(ws), st))
        sflag, nvalid = ctypes.c_int(0), ctypes.c_int64(0)
        nat.check(L.ryd_lanczos_status(nat.ptr(state), ctypes.byref(sflag), ctypes.byref(nvalid),
                                       st))
        stopped = bool(sflag.value)
        j = int(nvalid.value) if stopped else j + k
        if j < 1:
            return None
        a = alpha[:j].cpu().numpy()
        nr = nrm[:j + 1].cpu().numpy()
        theta, Y = _tridiag(a, nr[1:j])
        blast = 0.0 if stopped else float(nr[j])
        i = _genuine_lowest(a, nr[1:j], theta, 1, _ritz_bounds(Y, blast), max(float(tol), 1e-12),
                            merge_rel=merge_rel)[0]
        bound = abs(blast * Y[-1, i])
        if bound <= tol * max(1.0, abs(theta[i])) or stopped or j >= kmax:
            break
    # the Ritz vector by a replay of the same recurrence from the same start
    # (the replay reads the coefficients on the device)
    yt = torch.as_tensor(np.ascontiguousarray(Y[:, i]), dtype=torch.float64, device=P.device)
    x1 = torch.zeros(dim, dtype=dt, device=P.device)
    Rb = [phi, torch.empty_like(phi), torch.empty_like(phi)]
    nat.check(L.ryd_lanczos_locked(P.handle, code, nat.ptr(Rb[0]), nat.ptr(Rb[1]), nat.ptr(Rb[2]),
                                   nat.ptr(g), 0, j, nat.ptr(alpha), nat.ptr(nrm), nat.ptr(gam),
                                   nat.ptr(state), float(tiny), nat.ptr(x1), None, nat.ptr(yt),
                                   None, 1, nat.ptr(ws), st))
    E1, r1 = rayleigh_residual(P, x1, theta[i])
    return E1, r1, x1


def _start_vector(dim, seed, complex_kind, device):
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    v = torch.randn(dim, dtype=torch.float64, device=device, generator=gen)
    if complex_kind:
        v = v.to(torch.complex128) + 1j * torch.randn(dim, dtype=torch.float64, device=device,
                                                       generator=gen).to(torch.complex128)
    return v



def restricted_plan(model, basis_indices=None, blockade_edges=None, device='cuda'):
    """(Plan, order) for a restricted basis.

    ``blockade_edges``: the independent sets of that graph, enumerated on the
    device (chain edges without pair flips take the chain-sector kernels);
    with ``basis_indices`` as well, the graph basis is used only when it holds
    exactly those codes.  Otherwise ``basis_indices`` (unique codes in any
    order): the chain sector when it is exactly that, else a code-list plan
    (closed-form ranks for a complete number sector, binary search
    otherwise).  ``order`` maps the plan's ascending order back to the order
    of ``basis_indices`` (None when the caller gave no list)."""
    import torch
    from .basis import is_chain_blockade_basis, validate_edges, is_chain_edges, sector_codes
    idx = None
    if basis_indices is not None:
        if isinstance(basis_indices, torch.Tensor):
            basis_indices = basis_indices.cpu().numpy()
        idx = np.asarray(basis_indices, np.int64).reshape(-1)
        if idx.size == 0:
            raise ValueError('basis_indices is empty')
        if np.any(idx < 0) or np.any(idx >= (1 << model.n)):
            raise ValueError('basis_indices must lie in [0, 2^n)')
    if blockade_edges is not None:
        edges = validate_edges(model.n, blockade_edges)
        if is_chain_edges(model.n, edges) and not model.has_pair_flips:
            if idx is None:
                return Plan(model, basis='chain_fib', device=device), None
        else:
            cand = Plan(model, basis='graph', edges=edges, device=device)
            if idx is None:
                return cand, None
            if idx.size == cand.dim and np.array_equal(np.sort(idx),
                                                       sector_codes(cand).cpu().numpy()):
                return cand, np.argsort(idx, kind='stable')
            cand.close()
    if idx is None:
        raise ValueError('a restricted basis needs basis_indices or blockade_edges')
    order = np.argsort(idx, kind='stable')
    codes = idx[order]
    if np.any(np.diff(codes) == 0):
        raise ValueError('basis_indices contains duplicate codes')
    if not model.has_pair_flips and is_chain_blockade_basis(model.n, codes):
        return Plan(model, basis='chain_fib', device=device), order
    return Plan(model, basis='codes', codes=codes, device=device), order

def hip_lanczos_ground_state(calib, Om=None, delta=None, gamma=None, terms=None,
                             basis_indices=None, m=48, seed=0, dtype='float64', device='cuda',
                             tol=None, max_restarts=8, return_state=False,
                             memory_budget_mb=None, refine=True, plan=None, basis='full',
                             blockade_edges=None, check_degeneracy=None, bulk_tol=None,
                             refine_check_steps=None):
    """Ground state and gap with residual certification (see module docstring).

    Keys as ``torch_lanczos_ground_state``: E0, gap, evals, residuals,
    converged, backend, dim, krylov_dim, restarts, tolerance[, warning][, vec];
    plus ``precision`` and ``n_atom`` (<n_i> of the ground state).

    ``basis``: ``'full'`` or ``'chain_blockade'`` (the nearest-neighbour
    blockade sector of a chain, ranked on the device; no code list is built).
    ``basis_indices`` (any order, unique codes) restricts to those codes: the
    chain blockade sector takes the sector kernels, any other list (number
    sectors, blockade graphs, J/A models) the generic channel kernel with
    closed-form or binary-search ranks; ``vec`` comes back in the order of
    ``basis_indices`` (as matrix_free.py:118-286 with basis_indices).

    ``blockade_edges``: the basis is the independent sets of this graph,
    enumerated and ranked on the device (``RYD_BASIS_GRAPH``; chain edges
    take the chain sector kernels).  Given together with ``basis_indices``,
    the graph basis is used when it holds exactly those codes (a plan from
    ``build_plan``), else the codes themselves.

    ``check_degeneracy``: a single Lanczos vector sees each eigenvalue once,
    while the reference's two-vector block Krylov reports a degenerate
    ground level twice (gap 0).  Unless the level is simple by
    Perron-Frobenius (``ground_is_unique``), E1 is taken again over the
    orthogonal complement of the ground vector (``locked_lowest``) and
    replaces the first run's E1 when it lies lower; ``None`` (default) does
    this when needed, ``False`` never.  ``multiplicity_check`` in the result
    says which case applied.

    ``bulk_tol``: residual tolerance of the FP32 stage (default 1e-6 before
    an FP64 refinement, else ``tol``); ``refine_check_steps``: steps between
    the refinement's convergence checks (default ``CHECK_STEPS``).
    """
    import torch
    if str(device).startswith('cpu'):
        raise RuntimeError('HIP device unavailable')
    nat.require_gpu()
    if dtype not in ('float64', 'complex128', 'float32', 'complex64'):
        raise ValueError('dtype must be float32/float64/complex64/complex128')
    if int(m) < 2 or int(max_restarts) < 0:
        raise ValueError('m must be >= 2 and max_restarts >= 0')
    # a given plan (a model plan or an assembled-matrix CsrPlan) carries H
    model = from_calib(calib, Om, delta, gamma, terms) if plan is None else None
    order = None
    if plan is not None:
        P = plan
    elif basis_indices is not None or blockade_edges is not None:
        P, order = restricted_plan(model, basis_indices, blockade_edges, device=device)
    else:
        if basis not in ('full', 'chain_blockade'):
            raise ValueError('basis must be full or chain_blockade (or given by basis_indices '
                             'or blockade_edges)')
        P = Plan(model, basis='full' if basis == 'full' else 'chain_fib', device=device)
    dim = P.dim
    low = dtype in ('float32', 'complex64')
    complex_kind = (not P.is_real) or dtype.startswith('complex')
    itemsize = (8 if low else 16) if complex_kind else (4 if low else 8)
    tolerance = float(tol if tol is not None else (1e-9 if not low or refine else 2e-5))
    if tolerance <= 0 or not np.isfinite(tolerance):
        raise ValueError('tol must be positive and finite')
    required = dim * itemsize * 7 + (dim * 16 * 7 if low and refine else 0)
    if memory_budget_mb is not None and required > float(memory_budget_mb) * 1024 ** 2:
        raise MemoryError('HIP Lanczos memory estimate exceeds memory_budget_mb')
    want = 2 if dim > 1 else 1
    tdt64 = torch.complex128 if complex_kind else torch.float64
    tdt32 = torch.complex64 if complex_kind else torch.float32
    work = tdt32 if low else tdt64
    v0 = _start_vector(dim, seed, complex_kind, P.device)
    cycles = int(max_restarts) + 1
    # the FP32 stage before a refinement runs to 1e-6 (C3: 160 FP32 + 70 FP64
    # steps, 1.29 s, against 140 + 120 steps, 1.97 s, at the reference's
    # FP32 tolerance 2e-5; plain FP64 200 steps, 2.51 s) and gets one run:
    # a model whose FP32 residuals stall above it hands over to FP64 as is
    if low:
        bulk_tol = float(bulk_tol) if bulk_tol is not None else (1e-6 if refine else tolerance)
    else:
        bulk_tol = tolerance
    stats = dict(steps_bulk=0, steps_fp64=0, restarts=0)
    vecs, R, evals, res = _solve(P, v0.to(work), int(m), cycles, bulk_tol, want, stats,
                                 'steps_bulk', tdt64, attempts=1 if (low and refine) else None)
    if low and refine:
        # FP64 refinement seeded by the FP32 Ritz pair
        start = _combine(vecs, tdt64)
        vecs, R, evals, res = _solve(P, start, int(m), cycles, tolerance, want, stats,
                                     'steps_fp64', tdt64, check_steps=refine_check_steps)
    steps_bulk, steps64, restarts = stats['steps_bulk'], stats['steps_fp64'], stats['restarts']
    evals = np.array(evals, dtype=float)
    res = np.array(res, dtype=float)
    count = len(evals)
    unique = ground_is_unique(P)
    mult = 'simple (Perron-Frobenius)' if unique else 'not checked'
    if want == 2 and not unique and check_degeneracy is not False and len(vecs):
        x0 = vecs[0] if vecs[0].dtype == tdt64 else cast(vecs[0], tdt64)
        lk = locked_lowest(P, x0, tolerance, seed=int(seed) + 1)
        if lk is None:
            mult = 'not checked (the locked run does not fit device memory)'
        else:
            E1, r1, _ = lk
            mult = 'checked (second level over the complement of the ground vector)'
            if count < 2:
                evals = np.append(evals, E1)
                res = np.append(res, r1)
                count = 2
            elif E1 < evals[1] - tolerance * max(1.0, abs(evals[1])):
                evals[1], res[1] = E1, r1
    flags = res <= tolerance * np.maximum(1.0, np.abs(evals))
    converged = bool(count == want and np.all(flags))
    if R.stopped and count == want:
        converged = converged or bool(np.all(res <= tolerance * np.maximum(1.0, np.abs(evals))))
    out = dict(E0=float(evals[0]),
               gap=float(evals[1] - evals[0]) if converged and count > 1 else float('nan'),
               evals=evals, residuals=res, converged=converged,
               backend='hip-matrixfree-lanczos', dim=dim, krylov_dim=int(steps_bulk + steps64),
               restarts=int(restarts), tolerance=tolerance,
               multiplicity_check=mult,
               precision=dict(bulk='float32' if low else 'float64',
                              alpha_beta='float64', tridiagonal='float64',
                              refinement='float64' if (not low or refine) else None,
                              steps_bulk=int(steps_bulk), steps_fp64=int(steps64)))
    if not converged:
        out['warning'] = 'lowest eigenpairs did not all reach the residual tolerance; gap not reported'
    x0 = vecs[0]
    if getattr(P, 'basis', None) != 'csr':       # a matrix plan has no sites
        out['n_atom'] = site_occupations(P, x0)
    if return_state:
        x = x0 / torch.linalg.vector_norm(x0)
        v = x.cpu().numpy().astype(np.complex128)
        if order is not None:
            back = np.empty_like(v)
            back[order] = v
            v = back
        out['vec'] = v
    return out


def hamiltonian_linear_operator_hip(calib, Om=None, delta=None, gamma=None, terms=None,
                                    basis_indices=None, device='cuda'):
    """``hamiltonian_linear_operator`` (matrix_free.py:51-62) on the device: a
    scipy ``LinearOperator`` (dim × dim, complex128) whose matvec is the
    plan's H·v, plus the loss diagonal −(i/2)·γ_i·1e-3/CONV·n_i of
    channel_spec (channels.py:58-77) when ``gamma`` is not zero
    (``ryd_loss_axpy``).  ``basis_indices``: the codes of a restricted basis,
    vectors in their order, as ChannelKernel maps codes to basis indices.
    Each matvec copies the vector to the device and back."""
    import torch
    from scipy.sparse.linalg import LinearOperator
    from .evolve import CONV
    n = len(calib.get('atoms') or [])
    g = np.zeros(n) if gamma is None else np.asarray(gamma, float).reshape(-1)
    if g.shape != (n,) or not np.all(np.isfinite(g)):
        raise ValueError('gamma must be a finite array of length N')
    model = from_calib(calib, Om, delta, None, terms)
    if basis_indices is None:
        P, order = Plan(model, device=device), None
    else:
        P, order = restricted_plan(model, basis_indices, None, device=device)
    loss = np.ascontiguousarray(g * 1e-3 / CONV)
    lossy = bool(np.any(g != 0.0))

    def matvec(v):
        v = np.asarray(v, complex).reshape(-1)
        if v.size != P.dim:
            raise ValueError('vector of length %d for a %d-dimensional operator' % (v.size, P.dim))
        if order is not None:
            v = v[order]
        x = torch.as_tensor(np.ascontiguousarray(v), device=P.device)
        y = P.hv(x)
        if lossy:
            nat.check(nat.lib().ryd_loss_axpy(P.handle, nat.RYD_C128, nat.ptr(x), nat.ptr(y),
                                              loss.ctypes.data, nat.stream_ptr()))
        out = y.cpu().numpy()
        if order is not None:
            back = np.empty_like(out)
            back[order] = out
            out = back
        return out

    op = LinearOperator((P.dim, P.dim), matvec=matvec, dtype=np.complex128)
    op.plan = P
    return op
