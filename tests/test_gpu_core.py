"""GPU parity tests: every call goes through the C ABI of libzquatev_amd.so."""
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


def test_mfma_register_layout(gpu_ctx):
    """v_mfma_f64_16x16x4_f64: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
    register r of lane l holds D[4r + l // 16][l % 16] (the layout gemm.hip assumes)."""
    import torch
    _say("case: mfma layout probe")
    g = torch.Generator().manual_seed(5)
    a = torch.randn(64, dtype=torch.float64, generator=g)
    b = torch.randn(64, dtype=torch.float64, generator=g)
    da, db = a.cuda(), b.cuda()
    out = torch.zeros(256, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.selftest_mfma_device(da.data_ptr(), db.data_ptr(), out.data_ptr())
    gpu_ctx.synchronize()
    assert rc == 0
    A = a.numpy().reshape(4, 16).T
    B = b.numpy().reshape(4, 16)
    C = A @ B
    got = out.cpu().numpy().reshape(64, 4)
    expect = np.array([[C[4 * r + l // 16][l % 16] for r in range(4)] for l in range(64)])
    assert np.allclose(got, expect, rtol=1e-14, atol=1e-14), np.abs(got - expect).max()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 11, 21, 22, 33, 64, 65, 100])
def test_tridiagonal_matches_oracle(gpu_ctx, n):
    """The reduced real tridiagonal (diag, |subdiag|) equals the reference
    algorithm's (unique by the implicit-Q theorem), to rounding."""
    import torch
    _say(f"case: tridiagonalize n={n}")
    D = gen.uniform_case(n, 2000 + n)
    d0, e0, _, _ = zq_oracle.reduce_to_tridiagonal(D[:n, :n], D[n:, :n])
    Dt = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()   # row-major of D^T == column-major D
    dd = torch.zeros(n, dtype=torch.float64, device="cuda")
    de = torch.zeros(max(n - 1, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.tridiagonalize_device(2 * n, Dt.data_ptr(), 2 * n, dd.data_ptr(), de.data_ptr())
    gpu_ctx.synchronize()
    assert rc == 0
    d = dd.cpu().numpy()
    e = de.cpu().numpy()[: n - 1]
    scale = np.linalg.norm(gen.full_hermitian(D, n), 2)
    assert np.abs(d - d0).max() <= 1e-13 * scale * max(n, 1)
    if n > 1:
        assert np.abs(e - np.abs(e0)).max() <= 1e-13 * scale * n


# mirror: a palindromic tridiagonal (d and e read the same backwards), whose
# two halves have the same spectrum after the middle tear, so the top merge
# pairs every value of one half with an equal one of the other: deflation
# rotations across the halves, the kept columns mixing both (the merge
# product's both-halves group).  glued: copies of Wilkinson's W21+ glued by
# 1e-8, clusters of nearly equal eigenvalues spread over the tree.
@pytest.mark.parametrize("n,kind", [(1, "rand"), (2, "rand"), (7, "rand"), (64, "rand"), (65, "rand"),
                                    (130, "rand"), (300, "rand"), (257, "wilk"), (200, "clust"),
                                    (190, "split"), (128, "zero"), (256, "mirror"), (512, "mirror"),
                                    (336, "glued")])
def test_tridiag_eig(gpu_ctx, n, kind):
    import torch
    import scipy.linalg.lapack as L
    _say(f"case: tridiag_eig n={n} kind={kind}")
    rng = np.random.default_rng(n)
    if kind == "rand":
        d = rng.normal(size=n); e = np.abs(rng.normal(size=max(n - 1, 0)))
    elif kind == "wilk":
        d = np.abs(np.arange(n) - (n - 1) / 2.0); e = np.ones(n - 1)
    elif kind == "clust":
        d = 1.0 + 1e-12 * rng.normal(size=n); e = 1e-10 * np.abs(rng.normal(size=n - 1))
    elif kind == "split":
        d = rng.normal(size=n); e = np.abs(rng.normal(size=n - 1)); e[::5] = 0.0
    elif kind == "mirror":
        h = n // 2
        d1 = rng.normal(size=h); e1 = np.abs(rng.normal(size=h - 1))
        d = np.concatenate([d1, d1[::-1]]); e = np.concatenate([e1, [0.5], e1[::-1]])
    elif kind == "glued":
        w = np.abs(np.arange(21) - 10.0)
        d = np.tile(w, n // 21 + 1)[:n]; e = np.ones(n - 1)
        e[20::21] = 1e-8
    else:
        d = np.zeros(n); e = np.zeros(n - 1)
    dd = torch.from_numpy(d.copy()).cuda()
    de = torch.from_numpy(np.concatenate([e, [0.0]])).cuda()
    dw = torch.zeros(n, dtype=torch.float64, device="cuda")
    dz = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.tridiag_eig_device(n, dd.data_ptr(), de.data_ptr(), dw.data_ptr(), dz.data_ptr())
    gpu_ctx.synchronize()
    assert rc == 0
    w = dw.cpu().numpy()
    Z = dz.cpu().numpy().reshape(n, n).T   # column-major
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref = np.linalg.eigvalsh(T)
    tn = np.abs(T).max() if np.abs(T).max() > 0 else 1.0
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - ref).max() <= 1e-13 * tn * max(1, np.sqrt(n))
    assert np.abs(T @ Z - Z * w[None, :]).max() <= 1e-13 * tn * max(1, np.sqrt(n))
    assert np.abs(Z.T @ Z - np.eye(n)).max() <= 1e-13
    # eigenvalues-only path
    dw2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = gpu_ctx.tridiag_eig_device(n, dd.data_ptr(), de.data_ptr(), dw2.data_ptr(), 0)
    gpu_ctx.synchronize()
    assert rc == 0
    assert np.abs(dw2.cpu().numpy() - ref).max() <= 1e-13 * tn * max(1, np.sqrt(n))


@pytest.mark.parametrize("n2", [2, 4, 6, 8, 16, 22, 42, 44, 64, 128, 130, 400, 510, 512])
def test_zquatev_host(gpu_ctx, n2):
    _say(f"case: zquatev host 2n={n2}")
    n = n2 // 2
    D = gen.uniform_case(n, 3000 + n2)
    H = inv.full_h(D, n)
    X = np.array(D, order="F", copy=True)
    info, eig = gpu_ctx.zquatev(X)
    assert info == 0
    inv.assert_solution(H, X, eig)
    # parity with the oracle: eigenvalues and the gauge-free Kramers-pair projectors
    _, Xo, eo = zq_oracle.zquatev(D)
    nh = np.linalg.norm(H, 2)
    assert np.abs(eig - eo).max() <= 1e-12 * nh
    assert inv.projector_distance(X, Xo, eo, n) <= 1e-10
