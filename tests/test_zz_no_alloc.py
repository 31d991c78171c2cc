"""The allocation-free KPM moments entry point (ryd_cheb_moments_noalloc)
against ryd_cheb_moments with a caller workspace, through the C ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(pkg, dev, n):
    calib = pkg.from_positions(pkg.chain_positions(n), pkg.c6_coupling(10.0), terms=('I1',))
    return pkg.Plan(pkg.from_calib(calib, np.ones(n), -1.2 * np.ones(n)), device=dev)


def _moments(pkg, P, dtype, nsteps, c, r, noalloc):
    """sums[(nsteps+1)*4] of one reference-stream probe, through the C ABI."""
    import torch
    from energy_level_amd import _native as nat
    L = nat.lib()
    tdt = {'float64': torch.float64, 'float32': torch.float32}[dtype]
    probe = torch.from_numpy(next(pkg.probe_bits(P.dim, 1, 7))).to(P.device)
    bufs = [torch.empty(P.dim, dtype=tdt, device=P.device) for _ in range(3)]
    sums = torch.zeros(4 * (nsteps + 1), dtype=torch.float64, device=P.device)
    code = nat.dtype_code(tdt)
    if noalloc:
        rc = L.ryd_cheb_moments_noalloc(P.handle, code, nat.ptr(probe), nat.ptr(bufs[0]),
                                        nat.ptr(bufs[1]), nat.ptr(bufs[2]), nsteps, float(c),
                                        float(r), nat.ptr(sums), nat.stream_ptr())
    else:
        rc = L.ryd_cheb_moments(P.handle, code, nat.ptr(probe), nat.ptr(bufs[0]),
                                nat.ptr(bufs[1]), nat.ptr(bufs[2]), nsteps, float(c), float(r),
                                nat.ptr(sums), nat.ptr(P.workspace()), nat.stream_ptr())
    nat.check(rc)
    torch.cuda.synchronize()
    return sums.cpu().numpy(), bufs[nsteps % 3].cpu().numpy()


def test_noalloc_moments_equal_workspace_entry_point(pkg, dev):
    """Two calls in a row, the second on a smaller plan (the array keeps the
    first call's partials), FP64 and FP32: sums and the last Chebyshev vector
    equal those of the workspace entry point bit for bit (same kernels, same
    launch shapes, same fixed-order reductions)."""
    for n, nsteps in ((22, 24), (13, 17)):
        P = _plan(pkg, dev, n)
        _, _, c, r = pkg.spectral_bounds(P, margin=0.01)
        for dtype in ('float64', 'float32'):
            ref_s, ref_t = _moments(pkg, P, dtype, nsteps, c, r, noalloc=False)
            got_s, got_t = _moments(pkg, P, dtype, nsteps, c, r, noalloc=True)
            assert np.array_equal(ref_s, got_s), (n, dtype, np.max(np.abs(ref_s - got_s)))
            assert np.array_equal(ref_t, got_t), (n, dtype)
            assert np.all(np.isfinite(got_s)) and abs(got_s[4 * 1 + 2] - P.dim) < 1e-6 * P.dim
        P.close()


def test_noalloc_refuses_plans_beyond_its_array(pkg, dev):
    """A plan whose reductions need more than the 8 MiB array is refused by
    the host before any device work (the vector arguments are never read)."""
    from energy_level_amd import _native as nat
    L = nat.lib()
    P = _plan(pkg, dev, 32)
    assert int(L.ryd_workspace_bytes(P.handle)) > (8 << 20)
    fake = ctypes.c_void_p(1 << 20)
    rc = L.ryd_cheb_moments_noalloc(P.handle, nat.RYD_F64, fake, fake, fake, fake, 4, 0.0, 1.0,
                                    fake, nat.stream_ptr())
    assert rc == nat.RYD_ERR_INVALID
    assert 'bytes of scratch' in nat.last_error()
    P.close()
