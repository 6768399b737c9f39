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


@pytest.mark.parametrize("tsm,n", [(0, 300), (150, 300), (0, 97), (0, 130)])
def test_tridiagonal_matvec_switch(gpu_ctx, tsm, n):
    """The matvec tile shapes against the oracle.  Switch 0 runs the wide
    (64 x 32, branch-free) tiles at every trailing size, including the partly
    empty last row groups of small trailing sizes (97 and 130 leave 32 and 1
    rows in the last group at the start); 150 changes shape mid-solve."""
    _say(f"case: tridiagonalize n={n} matvec switch {tsm}")
    D = gen.uniform_case(n, 4100 + n + tsm)
    d0, e0, _, _ = zq_oracle.reduce_to_tridiagonal(D[:n, :n], D[n:, :n])
    gpu_ctx.set_tile_shape(tsm, 0)
    try:
        d, e = _tridiag_gpu(gpu_ctx, D, n)
    finally:
        gpu_ctx.set_tile_shape(-1, 0)
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


@pytest.mark.parametrize("n2,pad", [(4096, 6), (3000, 1)])
def test_zquatev_host_chunked(gpu_ctx, n2, pad):
    """Host entry at sizes whose copies span several pinned chunks, with and
    without rows of padding below the matrix (ld2 > n2): the padding rows and
    the columns beyond n2 are left untouched."""
    _say(f"case: zquatev host chunked 2n={n2} pad={pad}")
    n = n2 // 2
    D = gen.uniform_case(n, 7000 + n2 + pad)
    H = inv.full_h(D, n)
    X = np.zeros((n2 + pad, n2 + 1), dtype=np.complex128, order="F")
    X[:n2, :n2] = D
    X[n2:, :n2] = 7.0 - 3.0j
    X[:, n2] = -1.5 + 2.5j
    info, eig = gpu_ctx.zquatev(X, n2)
    assert info == 0
    assert np.all(X[n2:, :n2] == 7.0 - 3.0j)
    assert np.all(X[:, n2] == -1.5 + 2.5j)
    inv.assert_solution(H, np.asfortranarray(X[:n2, :n2]), eig)


def _device_checks(Dcols, n, eig, H, ncols=64):
    """Size-independent checks of a device solve: trace identity, residual and
    orthonormality of sampled eigenpairs, Kramers structure of the sampled columns."""
    import torch
    n2 = 2 * n
    X = Dcols.T                                        # column-major view of the result
    cols = torch.arange(0, n, max(1, n // ncols), device=H.device)
    cols = torch.cat([cols, cols + n])
    Xs = X[:, cols].contiguous()
    lam = torch.cat([eig, eig])[cols].to(H.dtype)
    g = torch.Generator(device=H.device).manual_seed(3)
    v = torch.randn(n2, dtype=torch.float64, device=H.device, generator=g).to(H.dtype)
    for _ in range(40):                                # power iteration: ||H||_2 from below
        v = H @ v
        v = v / torch.linalg.vector_norm(v)
    hn = torch.linalg.vector_norm(H @ v).item()
    res = (torch.linalg.norm(H @ Xs - Xs * lam[None, :]) / (n * hn)).item()
    orth = (Xs.conj().T @ Xs - torch.eye(len(cols), dtype=Xs.dtype, device=H.device)).abs().max().item()
    trace = abs(2.0 * eig.sum().item() - torch.diagonal(H).real.sum().item()) / (n2 * hn)
    half = len(cols) // 2
    U = Xs[:n, :half]
    V = Xs[n:, :half]
    struct = torch.equal(Xs[:n, half:], -V.conj()) and torch.equal(Xs[n:, half:], U.conj())
    return res, orth, trace, struct, bool(torch.all(eig[1:] >= eig[:-1]).item())


_EIG_1300 = {}


@pytest.mark.parametrize("nb", [16, 32, 64, 128])
def test_zquatev_wide_column_kernel(gpu_ctx, nb):
    """n = 1300: trailing sizes above 1024 run the wide column kernel (one
    instantiation per panel width class 16/32/64/128 columns), and the switch
    to the light one (trailing size 1024, column 275) falls inside a panel
    for every nb.  Device checks of the solve, and its eigenvalues against
    LAPACK's on the full 2600 x 2600 H."""
    import torch
    n = 1300
    _say(f"case: device solve 2n={2 * n} nb={nb}")
    D = gen.uniform_case(n, 6100)
    if "ref" not in _EIG_1300:
        H = gen.full_hermitian(D, n)
        _EIG_1300["ref"] = np.linalg.eigvalsh(H)[::2]      # each eigenvalue twice (Kramers)
        _EIG_1300["norm"] = np.abs(_EIG_1300["ref"]).max()
    Dcols = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
    H = gen.full_hermitian_t(Dcols, n)
    eig = torch.zeros(n, dtype=torch.float64, device="cuda")
    gpu_ctx.set_panel_width(nb)
    try:
        torch.cuda.synchronize()
        rc = gpu_ctx.zquatev_device(2 * n, Dcols.data_ptr(), 2 * n, eig.data_ptr(), "V")
        gpu_ctx.synchronize()
    finally:
        gpu_ctx.set_panel_width(0)
    assert rc == 0
    res, orth, trace, struct, ascending = _device_checks(Dcols, n, eig, H)
    _say(f"  residual {res:.3e} orth {orth:.3e} trace {trace:.3e}")
    assert res <= inv.RES_TOL and orth <= inv.ORTH_TOL and trace <= 1e-13 and struct and ascending
    err = np.abs(eig.cpu().numpy() - _EIG_1300["ref"]).max()
    assert err <= inv.EIG_TOL * _EIG_1300["norm"], err


def test_zquatev_2n32768_large_solve(gpu_ctx):
    """The large single solve at full size (2n = 32768, ~17 GB matrix resident
    in HBM, uniform entries with diag(A) + 10): trace identity, residual and
    orthonormality of 128 sampled eigenvectors, Kramers structure, ascending
    eigenvalues."""
    import torch
    n = 16384
    _say(f"case: large solve 2n={2 * n}")
    Dcols = torch.zeros(2 * n, 2 * n, dtype=torch.complex128, device="cuda")
    gen.uniform_case_device(n, 1004, Dcols, diag_shift=10.0)
    H = gen.full_hermitian_t(Dcols, n)
    eig = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.zquatev_device(2 * n, Dcols.data_ptr(), 2 * n, eig.data_ptr(), "V")
    gpu_ctx.synchronize()
    assert rc == 0
    res, orth, trace, struct, ascending = _device_checks(Dcols, n, eig, H)
    _say(f"  residual {res:.3e} orth {orth:.3e} trace {trace:.3e}")
    assert res <= inv.RES_TOL and orth <= inv.ORTH_TOL and trace <= 1e-13 and struct and ascending
    del H, Dcols
    torch.cuda.empty_cache()
