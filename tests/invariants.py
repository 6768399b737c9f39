"""Gauge-free checks of a quaternion-Hermitian eigen-decomposition.

Bounds (stated once, used by every test):
  eigenvalues     max |lambda - lambda_zheev| <= 1e-12 ||H||_2
  residual        ||H X - X Lambda||_F / (n ||H||_2) <= 1e-13
  orthonormality  max |X^H X - I| <= 1e-13
  structure       right half == [[-conj V], [conj U]] bit for bit
"""
import numpy as np

EIG_TOL = 1e-12
RES_TOL = 1e-13
ORTH_TOL = 1e-13


def full_h(D, n):
    A = D[:n, :n]
    B = D[n:2 * n, :n]
    return np.block([[A, -B.conj()], [B, A.conj()]])


def metrics(H, X, eig):
    """H (2n x 2n), X (2n x 2n) eigenvectors, eig (n,) ascending."""
    n2 = H.shape[0]
    n = n2 // 2
    nh = max(np.linalg.norm(H, 2), 1e-300)
    lam = np.concatenate([eig, eig])
    res = np.linalg.norm(H @ X - X * lam[None, :]) / (max(n, 1) * nh)
    orth = np.abs(X.conj().T @ X - np.eye(n2)).max()
    U = X[:n, :n]
    V = X[n:, :n]
    struct = (np.array_equal(X[:n, n:], -V.conj()) and np.array_equal(X[n:, n:], U.conj()))
    return dict(res=res, orth=orth, struct=struct, hnorm=nh)


def eig_error(H, eig):
    w = np.linalg.eigvalsh(H)
    nh = max(np.linalg.norm(H, 2), 1e-300)
    return np.abs(w[::2] - eig).max() / nh, np.abs(w[1::2] - eig).max() / nh


def assert_solution(H, X, eig, eig_tol=EIG_TOL, res_tol=RES_TOL, orth_tol=ORTH_TOL):
    e0, e1 = eig_error(H, eig)
    m = metrics(H, X, eig)
    assert e0 <= eig_tol and e1 <= eig_tol, f"eigenvalue error {e0:.3e}/{e1:.3e} > {eig_tol}"
    assert m["res"] <= res_tol, f"residual {m['res']:.3e} > {res_tol}"
    assert m["orth"] <= orth_tol, f"orthonormality {m['orth']:.3e} > {orth_tol}"
    assert m["struct"], "Kramers structure of the right half broken"
    return dict(eig=max(e0, e1), **m)


def pair_projectors(X, n):
    """2n x 2n projector onto each Kramers pair: columns i and i+n."""
    P = []
    for i in range(n):
        Y = X[:, [i, i + n]]
        P.append(Y @ Y.conj().T)
    return P


def projector_distance(X1, X2, eig, n, gap_tol=1e-6):
    """max over well-separated eigenvalues of ||P1_i - P2_i||_max (gauge-free)."""
    worst = 0.0
    for i in range(n):
        gap = min(abs(eig[i] - eig[i - 1]) if i > 0 else np.inf,
                  abs(eig[i + 1] - eig[i]) if i + 1 < n else np.inf)
        if gap < gap_tol:
            continue
        Y1 = X1[:, [i, i + n]]
        Y2 = X2[:, [i, i + n]]
        d = np.abs(Y1 @ Y1.conj().T - Y2 @ Y2.conj().T).max()
        worst = max(worst, d)
    return worst
