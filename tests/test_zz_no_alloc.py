"""zq_eigvals_noalloc, the entry point that allocates nothing, against the
existing entry point whose kernels it shares (zq_zquatev_device with
jobz = 'N' on a context with the built-in panel widths): two calls in a row
through the C ABI, the second with a smaller input, each equal to the
existing entry point's result bit for bit and within the eigenvalue bound
against LAPACK.  Its scratch array is shared by every call on the device, so
the second call also shows that nothing it needs is left over from the
first."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import zquatev_amd as zq  # noqa: E402
from zquatev_amd import gen  # noqa: E402
from tests import invariants as inv  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc():
    f = zq.lib().zq_eigvals_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return f


def _solve_both(n, seed, stream, ctx):
    import torch
    import scipy.linalg as sl
    D = gen.uniform_case(n, seed)
    Dt = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()   # row-major D^T == column-major D
    keep = Dt.clone()
    eig_a = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    info = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = _noalloc()(2 * n, Dt.data_ptr(), 2 * n, eig_a.data_ptr(), info.data_ptr(), stream.cuda_stream)
    assert rc == 0
    stream.synchronize()
    assert info.item() == 0
    assert torch.equal(Dt, keep), "the input was modified"
    eig_b = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert ctx.zquatev_device(2 * n, Dt.data_ptr(), 2 * n, eig_b.data_ptr(), "N") == 0
    assert ctx.last_info() == 0
    assert torch.equal(eig_a, eig_b), "differs from zq_zquatev_device(jobz='N')"
    w = sl.eigh(gen.full_hermitian(D, n), eigvals_only=True, driver="evd")
    nh = max(abs(w[0]), abs(w[-1]))
    e = eig_a.cpu().numpy()
    assert np.abs(w[0::2] - e).max() <= inv.EIG_TOL * nh and np.abs(w[1::2] - e).max() <= inv.EIG_TOL * nh


def test_noalloc_matches_device_entry_twice_smaller_second():
    import torch
    stream = torch.cuda.Stream()
    with zq.Context(0) as ctx:          # built-in panel widths, as the no-allocation entry uses
        _say("case: noalloc 2n=2048 (the maximum) then 2n=600")
        _solve_both(1024, 7100, stream, ctx)
        _solve_both(300, 7101, stream, ctx)


def test_noalloc_rejects_and_reports():
    import torch
    f = _noalloc()
    Dt = torch.zeros(2050, 2050, dtype=torch.complex128, device="cuda")
    e = torch.zeros(1025, dtype=torch.float64, device="cuda")
    assert f(2050, Dt.data_ptr(), 2050, e.data_ptr(), None, None) == -5   # above the maximum order
    assert f(7, Dt.data_ptr(), 8, e.data_ptr(), None, None) == -1
    # a NaN input: status ZQ_ERR_NONFINITE through dinfo, eigenvalues NaN
    n = 40
    _say("case: noalloc non-finite input")
    D = gen.uniform_case(n, 7200)
    D[7, 3] = np.nan
    Dn = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert f(2 * n, Dn.data_ptr(), 2 * n, e.data_ptr(), info.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert info.item() == zq.ERR_NONFINITE
    assert torch.isnan(e[:n]).all().item()
