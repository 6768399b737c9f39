"""GPU parity across panel widths and larger sizes (C ABI calls)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from zquatev_amd import gen  # noqa: E402
from oracle import zq_oracle  # noqa: E402
from tests import invariants as inv  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tridiag_gpu(ctx, D, n):
    import torch
    Dt = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
    dd = torch.zeros(n, dtype=torch.float64, device="cuda")
    de = torch.zeros(max(n - 1, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.tridiagonalize_device(2 * n, Dt.data_ptr(), 2 * n, dd.data_ptr(), de.data_ptr())
    ctx.synchronize()
    assert rc == 0
    return dd.cpu().numpy(), de.cpu().numpy()[: n - 1]


@pytest.mark.parametrize("nb", [1, 7, 16, 32, 63, 64, 100, 128])
def test_tridiagonal_panel_widths(gpu_ctx, nb):
    n = 300
    _say(f"case: tridiagonalize n={n} nb={nb}")
    D = gen.uniform_case(n, 4000)
    d0, e0, _, _ = zq_oracle.reduce_to_tridiagonal(D[:n, :n], D[n:, :n])
    gpu_ctx.set_panel_width(nb)
    try:
        d, e = _tridiag_gpu(gpu_ctx, D, n)
    finally:
        gpu_ctx.set_panel_width(0)
    scale = np.linalg.norm(gen.full_hermitian(D, n), 2)
    assert np.abs(d - d0).max() <= 1e-12 * scale
    assert np.abs(e - np.abs(e0)).max() <= 1e-12 * scale


@pytest.mark.parametrize("n", [1000, 2048, 4096])
def test_tridiag_eig_large(gpu_ctx, n):
    import torch
    import scipy.linalg as sl
    _say(f"case: tridiag_eig n={n}")
    rng = np.random.default_rng(n)
    d = rng.normal(size=n)
    e = np.abs(rng.normal(size=n - 1))
    dd = torch.from_numpy(d.copy()).cuda()
    de = torch.from_numpy(np.concatenate([e, [0.0]])).cuda()
    dw = torch.zeros(n, dtype=torch.float64, device="cuda")
    dz = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.tridiag_eig_device(n, dd.data_ptr(), de.data_ptr(), dw.data_ptr(), dz.data_ptr())
    gpu_ctx.synchronize()
    assert rc == 0
    w = dw.cpu().numpy()
    ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
    tn = np.abs(d).max() + 2 * np.abs(e).max()
    assert np.abs(w - ref).max() <= 1e-13 * tn * np.sqrt(n)
    Zt = dz.reshape(n, n)   # row-major of Z^T
    Zg = Zt.T
    T = torch.from_numpy(np.diag(d) + np.diag(e, 1) + np.diag(e, -1)).cuda()
    res = (T @ Zg - Zg * dw[None, :]).abs().max().item()
    orth = (Zg.T @ Zg - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max().item()
    assert res <= 1e-13 * tn * np.sqrt(n), res
    assert orth <= 1e-13, orth


@pytest.mark.parametrize("n2,nb", [(1024, 64), (1024, 32), (2048, 64), (2048, 128)])
def test_zquatev_large(gpu_ctx, n2, nb):
    _say(f"case: zquatev host 2n={n2} nb={nb}")
    n = n2 // 2
    D = gen.uniform_case(n, 5000 + n2)
    H = inv.full_h(D, n)
    X = np.array(D, order="F", copy=True)
    gpu_ctx.set_panel_width(nb)
    try:
        info, eig = gpu_ctx.zquatev(X)
    finally:
        gpu_ctx.set_panel_width(0)
    assert info == 0
    inv.assert_solution(H, X, eig)
