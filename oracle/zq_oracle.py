"""CPU oracle for ts::zquatev -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the HIP solver.  Only tests/, the smoke check
in __graft_entry__.py and the cpu_baseline leg of bench.py may import it; the
library itself never calls it and has no CPU fallback.

It restates, in numpy, the algorithm of the reference implementation:

* ``zquatev``         -- the driver, zquatev.cc:42-100
* ``unblocked_sweep`` -- the per-column symplectic reduction, unblocked.cc:44-131
                         (B-column Householder, symplectic Givens, A-column
                         Householder; blocked.cc:45-546 applies the same
                         transforms in compact-WY form, so the reduced matrix
                         is the same up to rounding)
* ``zlarfg`` / ``zlartg`` -- the two LAPACK generators those files call
                         (LAPACK 3.x documented algorithms; pinned against
                         scipy.linalg.lapack's own zlarfg/zlartg)
* ``tridiag_eig``     -- zhbev("V","L",kd=1) as called at zquatev.cc:74-84:
                         zhbtrd's diagonal phase scaling to a real symmetric
                         tridiagonal, then QL/QR (LAPACK dstev = dsteqr,
                         through scipy) and Z = diag(phase) * Z_real.

Pinning (see tests/test_oracle.py): scipy zlarfg/zlartg known answers; the
compiled reference (oracle/_ref, built from /root/reference by
oracle/Makefile) on the same inputs -- tridiagonal T, eigenvalues, Kramers
pair projectors; LAPACK zheev on the full 2n x 2n matrix; closed forms
(B = 0, n = 1, n = 2); mpmath at 50 digits for 2n <= 8.
"""

import numpy as np
import scipy.linalg.lapack as _lapack

_EPS = np.finfo(np.float64).eps * 0.5        # dlamch('E')
_SAFMIN = np.finfo(np.float64).tiny          # dlamch('S')


def _dlapy3(x, y, z):
    w = max(abs(x), abs(y), abs(z))
    if w == 0.0:
        return abs(x) + abs(y) + abs(z)
    return w * np.sqrt((x / w) ** 2 + (y / w) ** 2 + (z / w) ** 2)


def _dznrm2(x):
    """2-norm without underflow or overflow (LAPACK dznrm2 semantics)."""
    if x.size == 0:
        return 0.0
    scale = float(np.max(np.abs(np.concatenate([x.real, x.imag]))))
    if scale == 0.0:
        return 0.0
    y = x / scale
    return scale * float(np.sqrt(np.sum(y.real * y.real + y.imag * y.imag)))


def zlarfg(alpha, x):
    """Elementary reflector H = I - tau [1; v][1; v]^H with H^H [alpha; x] = [beta; 0],
    beta real (LAPACK zlarfg).  Returns (beta, v, tau); x is not modified.

    Called at unblocked.cc:58, :104 and blocked.cc:176, :345."""
    x = np.asarray(x, dtype=np.complex128)
    alpha = complex(alpha)
    if x.size == 0:
        return alpha, x.copy(), 0.0 + 0.0j
    xnorm = _dznrm2(x)
    alphr, alphi = alpha.real, alpha.imag
    if xnorm == 0.0 and alphi == 0.0:
        return alpha, x.copy(), 0.0 + 0.0j
    beta = -np.copysign(_dlapy3(alphr, alphi, xnorm), alphr)
    rsafmn = 1.0 / (_SAFMIN / _EPS)
    safmin = _SAFMIN / _EPS
    knt = 0
    xs = x.copy()
    if abs(beta) < safmin:
        while True:
            knt += 1
            xs *= rsafmn
            beta *= rsafmn
            alphi *= rsafmn
            alphr *= rsafmn
            if not (abs(beta) < safmin and knt < 20):
                break
        xnorm = _dznrm2(xs)
        alpha = complex(alphr, alphi)
        beta = -np.copysign(_dlapy3(alphr, alphi, xnorm), alphr)
    tau = complex((beta - alphr) / beta, -alphi / beta)
    scal = 1.0 / (complex(alphr, alphi) - beta)
    v = xs * scal
    for _ in range(knt):
        beta *= safmin
    return beta, v, tau


def zlartg(f, g):
    """Plane rotation [c s; -conj(s) c] [f; g] = [r; 0] with c real >= 0 (LAPACK zlartg).

    Called at unblocked.cc:86 and blocked.cc:237."""
    f = complex(f)
    g = complex(g)
    if g == 0:
        return 1.0, 0.0 + 0.0j, f
    if f == 0:
        d = abs(g)
        return 0.0, np.conj(g) / d, complex(d)
    f1 = abs(f)
    g1 = abs(g)
    # scale-safe form of c = |f|/sqrt(|f|^2+|g|^2), s = f conj(g) / (|f| sqrt(..)), r = f/|f| sqrt(..)
    scale = max(f1, g1)
    fs = f / scale
    gs = g / scale
    f2 = abs(fs) ** 2
    g2 = abs(gs) ** 2
    h2 = f2 + g2
    rt = np.sqrt(h2)
    c = np.sqrt(f2 / h2)
    ph = fs / abs(fs)
    s = ph * np.conj(gs) / rt
    r = ph * rt * scale
    return float(c), complex(s), complex(r)


def _zrot(x, y, c, s):
    """zrot: x <- c x + s y ; y <- c y - conj(s) x (in place on views)."""
    t = c * x + s * y
    y[...] = c * y - np.conj(s) * x
    x[...] = t


def unblocked_sweep(D0, D1, Q0, Q1, k_begin=0):
    """Reduce (D0 = A, D1 = B) to (T, 0) with the symplectic transforms of
    unblocked.cc:52-129, accumulating them into (Q0, Q1).  All four arrays are
    n x n complex, modified in place.  Columns k_begin.. are processed."""
    n = D0.shape[0]
    for k in range(k_begin, n - 1):
        ln = n - k - 1
        if ln > 1:
            # Householder from the B column (unblocked.cc:53-81)
            beta, v1, tau = zlarfg(D1[k + 1, k], D1[k + 2:, k])
            tau = np.conj(tau)
            vec = np.concatenate(([1.0 + 0j], v1))
            cvec = np.conj(vec)
            # "00": two-sided update of D0 (unblocked.cc:64-67)
            tmp = D0[k + 1:, k:].conj().T @ cvec                    # zgemv("C", len, len+1)
            tmp[1:] += -np.conj(tau) * 0.5 * np.vdot(tmp[1:], cvec) * cvec
            D0[k + 1:, k:] += -np.conj(tau) * np.outer(cvec, tmp.conj())   # zgerc
            D0[k:, k + 1:] += -tau * np.outer(tmp, vec)                    # zgeru
            # "10": two-sided update of D1 (unblocked.cc:70-72)
            tmp = D1[k:, k + 1:] @ cvec                              # zgemv("N", len+1, len)
            D1[k + 1:, k:] += tau * np.outer(vec, tmp)               # zgeru
            D1[k:, k + 1:] += -tau * np.outer(tmp, vec)              # zgeru
            # accumulate into Q (unblocked.cc:75-80)
            tmp = Q0[:, k + 1:] @ cvec
            Q0[:, k + 1:] += -tau * np.outer(tmp, vec)
            tmp = Q1[:, k + 1:] @ cvec
            Q1[:, k + 1:] += -tau * np.outer(tmp, vec)

        # symplectic Givens rotation clearing D1[k+1, k] (unblocked.cc:83-96)
        c, s, _ = zlartg(D0[k + 1, k], D1[k + 1, k])
        x = D0[k + 1, k:].copy()
        y = D1[k + 1, k:].copy()
        _zrot(x, y, c, s)
        D0[k + 1, k:] = x
        D1[k + 1, k:] = y
        x = np.conj(D1[k:, k + 1])
        y = D0[k:, k + 1].copy()
        _zrot(x, y, c, s)
        D1[k:, k + 1] = np.conj(x)
        D0[k:, k + 1] = y
        x = np.conj(Q1[:, k + 1])
        y = Q0[:, k + 1].copy()
        _zrot(x, y, c, s)
        Q1[:, k + 1] = np.conj(x)
        Q0[:, k + 1] = y

        if ln > 1:
            # Householder from the A column (unblocked.cc:98-127)
            beta, v1, tau = zlarfg(D0[k + 1, k], D0[k + 2:, k])
            tau = np.conj(tau)
            vec = np.concatenate(([1.0 + 0j], v1))
            cvec = np.conj(vec)
            tmp = D0[k + 1:, k:].conj().T @ vec                      # zgemv("C")
            tmp[1:] += -tau * 0.5 * np.vdot(vec, tmp[1:]) * vec
            D0[k + 1:, k:] += -tau * np.outer(vec, tmp.conj())       # zgerc
            D0[k:, k + 1:] += -np.conj(tau) * np.outer(tmp, vec.conj())   # zgerc
            tmp = D1[k + 1:, k:].T @ vec                             # zgemv("T")
            D1[k + 1:, k:] += -np.conj(tau) * np.outer(cvec, tmp)    # zgeru
            D1[k:, k + 1:] += np.conj(tau) * np.outer(tmp, vec.conj())    # zgerc
            tmp = Q0[:, k + 1:] @ vec
            Q0[:, k + 1:] += -np.conj(tau) * np.outer(tmp, vec.conj())
            tmp = -(Q1[:, k + 1:] @ vec)
            Q1[:, k + 1:] += np.conj(tau) * np.outer(tmp, vec.conj())


def reduce_to_tridiagonal(A, B):
    """Symplectic reduction of the left half [A; B] (zquatev.cc:47-72 with the
    unblocked transforms).  Returns (d, e, Q0, Q1): real diagonal d (n), complex
    subdiagonal e (n-1) of T = Q^H H Q restricted to the top-left block, and the
    n x n blocks of Q = [[Q0, -conj(Q1)], [Q1, conj(Q0)]]."""
    n = A.shape[0]
    D0 = np.array(A, dtype=np.complex128, copy=True)
    D1 = np.array(B, dtype=np.complex128, copy=True)
    Q0 = np.eye(n, dtype=np.complex128)
    Q1 = np.zeros((n, n), dtype=np.complex128)
    unblocked_sweep(D0, D1, Q0, Q1)
    d = np.real(np.diag(D0)).copy()
    e = np.array([D0[i + 1, i] for i in range(n - 1)], dtype=np.complex128)
    return d, e, Q0, Q1


def tridiag_phase(e):
    """zhbtrd with kd = 1 (the first half of zhbev, zquatev.cc:84): unit phases p
    with diag(p)^H T diag(p) real, p[0] = 1.  Returns (|e|, p)."""
    n = e.size + 1
    p = np.ones(n, dtype=np.complex128)
    ae = np.abs(e)
    for i in range(n - 1):
        t = e[i] * p[i]
        p[i + 1] = t / abs(t) if abs(t) != 0.0 else 1.0
    return ae, p


def tridiag_eig(d, e, vectors=True):
    """Eigenpairs of the Hermitian tridiagonal (d real, e complex subdiagonal),
    ascending, as zhbev("V","L",n,1,...) returns them (zquatev.cc:76-84)."""
    ae, p = tridiag_phase(e)
    if d.size == 1:
        return d.copy(), (np.ones((1, 1), dtype=np.complex128) if vectors else None)
    if vectors:
        w, z, info = _lapack.dstev(d, ae, compute_v=1)
        if info != 0:
            raise RuntimeError(f"dstev info={info}")
        return w, p[:, None] * z
    w, _, info = _lapack.dstev(d, ae, compute_v=0)
    if info != 0:
        raise RuntimeError(f"dstev info={info}")
    return w, None


def zquatev(D, n2=None):
    """Oracle of ts::zquatev (zquatev.cc:42-100) for the ld2 == n2 layout.

    ``D`` is a (2n x 2n) complex128 array whose left half holds [A; B].
    Returns (info, X, eig): X = [[U, -conj V], [V, conj U]] (2n x 2n),
    eig (n,) ascending; D is not modified."""
    D = np.asarray(D)
    if n2 is None:
        n2 = D.shape[0]
    if n2 % 2:
        return -1, None, None
    n = n2 // 2
    A = D[:n, :n]
    B = D[n:n2, :n]
    d, e, Q0, Q1 = reduce_to_tridiagonal(A, B)
    w, Z = tridiag_eig(d, e, vectors=True)
    U = Q0 @ Z                                      # zquatev.cc:87
    V = Q1 @ Z                                      # zquatev.cc:90
    X = np.empty((n2, n2), dtype=np.complex128, order="F")
    X[:n, :n] = U
    X[n:, :n] = V
    X[:n, n:] = -np.conj(V)                         # zquatev.cc:92-98
    X[n:, n:] = np.conj(U)
    return 0, X, w


def eigvals(D, n2=None):
    """Eigenvalues only, through the same reduction."""
    D = np.asarray(D)
    if n2 is None:
        n2 = D.shape[0]
    n = n2 // 2
    d, e, _, _ = reduce_to_tridiagonal(D[:n, :n], D[n:n2, :n])
    w, _ = tridiag_eig(d, e, vectors=False)
    return w
