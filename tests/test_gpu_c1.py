"""The single-solve workload at full size: 2n = 8192, uniform_case(4096, 1001),
eigenvalues and eigenvectors, panel width chosen by the autotuner.

The reference checks every solve against zheev_ on the full 2n matrix
(test.cc:83-118); so does this test: all 4096 distinct eigenvalues against
LAPACK on the full H (each twice in its spectrum), then the residual and
orthonormality of 256 sampled eigenpairs (128 Kramers pairs, both columns),
the exact Kramers structure of the whole result, and ascending order."""
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


@pytest.mark.timeout(900)
def test_single_solve_2n8192():
    import torch
    import scipy.linalg as sl
    n = 4096
    _say("case: 2n=8192 single solve (autotuned panel width)")
    D = gen.uniform_case(n, 1001)
    Dt = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()       # row-major D^T == column-major D
    eig_d = torch.zeros(n, dtype=torch.float64, device="cuda")
    with zq.Context(0) as ctx:
        chosen, _ = ctx.autotune([n])
        _say(f"autotune: panel width {chosen[0]}")
        torch.cuda.synchronize()
        rc = ctx.zquatev_device(2 * n, Dt.data_ptr(), 2 * n, eig_d.data_ptr(), "V")
        assert rc == 0
        assert ctx.last_info() == 0
        # the values-only path on the same input agrees
        Dn = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
        eig_n = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert ctx.zquatev_device(2 * n, Dn.data_ptr(), 2 * n, eig_n.data_ptr(), "N") == 0
        assert ctx.last_info() == 0
    eig = eig_d.cpu().numpy()
    X = Dt.cpu().numpy().T                                        # column-major result, (2n, 2n)
    del Dn
    assert np.all(np.diff(eig) >= 0), "eigenvalues not ascending"
    H = gen.full_hermitian(D, n)
    _say("LAPACK eigenvalues of the full 8192 x 8192 H")
    w = sl.eigh(H, eigvals_only=True, driver="evd")
    nh = max(abs(w[0]), abs(w[-1]))                               # ||H||_2 of a Hermitian matrix
    e0 = np.abs(w[0::2] - eig).max() / nh
    e1 = np.abs(w[1::2] - eig).max() / nh
    _say(f"eigenvalue error {e0:.2e} / {e1:.2e}")
    assert e0 <= inv.EIG_TOL and e1 <= inv.EIG_TOL
    assert np.abs(eig_n.cpu().numpy() - eig).max() <= inv.EIG_TOL * nh
    # exact Kramers structure of the whole right half
    U, V = X[:n, :n], X[n:, :n]
    assert np.array_equal(X[:n, n:], -V.conj()) and np.array_equal(X[n:, n:], U.conj())
    # residual and orthonormality of 128 sampled pairs (columns i and i + n)
    pick = np.unique(np.linspace(0, n - 1, 128).astype(int))
    cols = np.concatenate([pick, pick + n])
    Xs = np.ascontiguousarray(X[:, cols])
    lam = np.concatenate([eig[pick], eig[pick]])
    res = np.linalg.norm(H @ Xs - Xs * lam[None, :]) / (n * nh)
    orth = np.abs(Xs.conj().T @ Xs - np.eye(len(cols))).max()
    _say(f"sampled residual {res:.2e} orthonormality {orth:.2e}")
    assert res <= inv.RES_TOL
    assert orth <= inv.ORTH_TOL
    # orthogonality of the sampled columns to all others, through X^H X = I on the device
    Xd = torch.from_numpy(X).cuda()
    G = Xd[:, torch.from_numpy(cols).cuda()].conj().T @ Xd
    G[torch.arange(len(cols)), torch.from_numpy(cols).cuda()] -= 1.0
    orth_all = G.abs().max().item()
    _say(f"sampled rows of X^H X - I: {orth_all:.2e}")
    assert orth_all <= inv.ORTH_TOL
