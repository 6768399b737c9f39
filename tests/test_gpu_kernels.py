"""Per-kernel GPU test of the quaternion GEMM on FP64 MFMA (zq_qgemm, the
kernel of the trailing update and of the back-transform) against a plain
FP64 product: quaternion matrices are multiplied through their 2 x 2
complex-block embedding H = [[A, -conj B], [B, conj A]] with numpy, a
representation independent of the kernel's 4-vector L(q) expansion."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rand_q(rng, r, c):
    """A random quaternion matrix as its complex components (a, b)."""
    return (rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c)),
            rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c)))


def _embed(a, b):
    return np.block([[a, -b.conj()], [b, a.conj()]])


def _unembed(H, r, c):
    return H[:r, :c], H[r:2 * r, :c]


def _pack(a, b, ld):
    """Column-major device layout: element (i, j) at 4 doubles (Re a, Im a, Re b, Im b), leading dimension ld."""
    r, c = a.shape
    out = np.zeros((c, ld, 4))
    out[:, :r, 0] = a.real.T
    out[:, :r, 1] = a.imag.T
    out[:, :r, 2] = b.real.T
    out[:, :r, 3] = b.imag.T
    return out


def _unpack(buf, r, c):
    a = buf[:c, :r, 0].T + 1j * buf[:c, :r, 1].T
    b = buf[:c, :r, 2].T + 1j * buf[:c, :r, 3].T
    return a, b


def _ct(a, b):
    """Quaternion conjugate transpose of (a, b): elements (conj a, -b), transposed."""
    return a.conj().T, -b.T


# alpha = -1, beta = 1 are the trailing update's and the back-transform's.
@pytest.mark.parametrize("m,n,k,ta,tb,lower,alpha,beta", [
    (1, 1, 1, "N", "N", False, -0.75, 1.25), (37, 53, 19, "N", "N", False, -0.75, 1.25),
    (64, 64, 64, "N", "C", True, -0.75, 1.25), (100, 80, 70, "C", "N", False, -0.75, 1.25),
    (129, 200, 64, "C", "C", False, -0.75, 1.25), (300, 300, 64, "N", "C", True, -0.75, 1.25),
    (513, 129, 257, "N", "N", False, -0.75, 1.25), (70, 65, 0, "N", "N", False, -0.75, 1.25),
    (300, 300, 64, "N", "C", True, -1.0, 1.0), (513, 129, 128, "N", "N", False, -1.0, 1.0),
    (200, 100, 24, "N", "N", False, 1.0, -0.5), (129, 200, 40, "C", "N", False, 1.0, 2.0),
    (64, 64, 33, "N", "C", True, -1.0, 1.0), (1100, 70, 128, "N", "N", False, -1.0, 1.0),
    (90, 90, 256, "C", "C", True, 1.0, 0.0), (1030, 1040, 128, "N", "N", False, -1.0, 1.0),
    (1500, 1500, 64, "N", "C", True, -1.0, 1.0)])
@pytest.mark.parametrize("t64", [0, 1])
def test_qgemm_against_complex_embedding(gpu_ctx, m, n, k, ta, tb, lower, alpha, beta, t64):
    """t64 = 1: 64-row tiles for every launch of at least one 64 x 64 tile per
    compute unit (0: the context's default threshold, 32-row tiles here)."""
    import torch
    _say(f"case: qgemm m={m} n={n} k={k} {ta}{tb} lower={lower} alpha={alpha} beta={beta} t64={t64}")
    rng = np.random.default_rng(m * 1000 + n + k)
    Ar, Ac = (k, m) if ta == "C" else (m, k)
    Br, Bc = (n, k) if tb == "C" else (k, n)
    A = _rand_q(rng, Ar, Ac)
    B = _rand_q(rng, Br, Bc)
    C = _rand_q(rng, m, n)
    opA = _ct(*A) if ta == "C" else A
    opB = _ct(*B) if tb == "C" else B
    HA, HB, HC = _embed(*opA), _embed(*opB), _embed(*C)
    ref = _unembed(alpha * (HA @ HB) + beta * HC if k > 0 else beta * HC, m, n)
    lda, ldb, ldc = Ar + 3, Br + 1, m + 2
    dA = torch.from_numpy(_pack(*A, lda)).cuda()
    dB = torch.from_numpy(_pack(*B, ldb)).cuda()
    dC = torch.from_numpy(_pack(*C, ldc)).cuda()
    before = dC.clone()
    torch.cuda.synchronize()
    gpu_ctx.set_tile_shape(-1, t64)
    try:
        rc = gpu_ctx.qgemm_device(ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(),
                                  ldc, lower)
        gpu_ctx.synchronize()
    finally:
        gpu_ctx.set_tile_shape(-1, 0)
    assert rc == 0
    got = _unpack(dC.cpu().numpy(), m, n)
    scale = max(1.0, np.sqrt(k)) * 8.0
    mask = np.tril(np.ones((m, n), dtype=bool)) if lower else np.ones((m, n), dtype=bool)
    for g, r in zip(got, ref):
        assert np.abs(g - r)[mask].max() <= 1e-14 * scale, np.abs(g - r)[mask].max()
    if lower:   # the strict upper triangle is left as it was
        old = _unpack(before.cpu().numpy(), m, n)
        for g, o in zip(got, old):
            assert np.array_equal(g[~mask], o[~mask])
    # padding rows beyond m of every column are untouched
    assert torch.equal(dC[:, m:, :], before[:, m:, :])


@pytest.mark.parametrize("m,n,k,ta,tb,lower", [(1030, 1040, 128, "N", "N", False), (1500, 1500, 64, "N", "C", True),
                                               (1100, 1024, 96, "C", "N", False)])
def test_qgemm_tile_rows_change_no_result(gpu_ctx, m, n, k, ta, tb, lower):
    """The same product with 32-row and with 64-row tiles is bitwise equal
    (a batch and a solo solve may pick different tile rows for one problem)."""
    import torch
    _say(f"case: qgemm tile rows m={m} n={n} k={k} {ta}{tb} lower={lower}")
    rng = np.random.default_rng(m + 7 * k)
    Ar, Ac = (k, m) if ta == "C" else (m, k)
    Br, Bc = (n, k) if tb == "C" else (k, n)
    dA = torch.from_numpy(_pack(*_rand_q(rng, Ar, Ac), Ar)).cuda()
    dB = torch.from_numpy(_pack(*_rand_q(rng, Br, Bc), Br)).cuda()
    C0 = torch.from_numpy(_pack(*_rand_q(rng, m, n), m)).cuda()
    outs = []
    for t64 in (0, 1):
        dC = C0.clone()
        torch.cuda.synchronize()
        gpu_ctx.set_tile_shape(-1, t64)
        try:
            assert gpu_ctx.qgemm_device(ta, tb, m, n, k, -1.0, dA.data_ptr(), Ar, dB.data_ptr(), Br, 1.0,
                                        dC.data_ptr(), m, lower) == 0
            gpu_ctx.synchronize()
        finally:
            gpu_ctx.set_tile_shape(-1, 0)
        outs.append(dC)
    assert torch.equal(outs[0], outs[1])


def test_qgemm_argument_errors(gpu_ctx):
    assert gpu_ctx.qgemm_device("X", "N", 2, 2, 2, 1.0, 0, 2, 0, 2, 0.0, 0, 2) == -2
    assert gpu_ctx.qgemm_device("N", "N", -1, 2, 2, 1.0, 0, 2, 0, 2, 0.0, 0, 2) == -4
    assert gpu_ctx.qgemm_device("N", "N", 2, 2, 2, 1.0, 0, 2, 0, 2, 0.0, 0, 2) == -5
