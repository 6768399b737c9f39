"""Seeded synthetic inputs for the quaternion-Hermitian eigensolver.

Every matrix is returned in the layout that ``ts::zquatev`` reads
(zquatev.h:41-52): a column-major ``2n x 2n`` complex128 array whose left
half holds ``[A; B]`` with A Hermitian and B complex antisymmetric
(``B^T = -B``, zero diagonal; the only reference generator builds B that way,
test.cc:66-67).  The right half is left zero: the solver uses it as output.

The random stream is counter-based SplitMix64, so any element can be produced
independently of the others and the same (generator, seed) gives the same
matrix on every machine:

    U(k) = mix64(seed + (k + 1) * 0x9E3779B97F4A7C15) >> 11  scaled to [0, 1)

Element (i, j) of the strictly lower triangle (i > j) owns stream slots
``4*(j*n + i) + c`` for c = 0..3 (Re A, Im A, Re B, Im B); the diagonal A_ii
owns slot ``4*(i*n + i)``.
"""

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed, idx):
    """U[0,1) doubles (53-bit) for integer stream positions ``idx`` (any shape)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _lower_slots(n, j0, j1):
    """Stream base slots for all (i, j) with j in [j0, j1), i in [0, n)."""
    i = np.arange(n, dtype=np.uint64)[:, None]
    j = np.arange(j0, j1, dtype=np.uint64)[None, :]
    return np.uint64(4) * (j * np.uint64(n) + i)


def _normal(seed, k0, k1):
    """Box-Muller normal from two stream slots."""
    u1 = uniform(seed, k0)
    u2 = uniform(seed, k1)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def _assemble(n, fill_cols, chunk=512):
    """Build D (2n x 2n, F-order) by calling fill_cols(j0, j1) -> (A_cols, B_cols)
    giving the strictly-lower parts plus diagonal of columns j0..j1 (n x w each)."""
    D = np.zeros((2 * n, 2 * n), dtype=np.complex128, order="F")
    A = D[:n, :n]
    B = D[n:, :n]
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        a, b = fill_cols(j0, j1)
        A[:, j0:j1] = a
        B[:, j0:j1] = b
    # mirror the strict lower triangles: A Hermitian, B antisymmetric
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        # upper part of columns j0..j1 = rows 0..j1 of those columns, from rows j0..j1 of the lower part
        blkA = A[j0:j1, :j1]  # rows j0..j1 (lower entries live where row > col)
        blkB = B[j0:j1, :j1]
        ii = np.arange(j0, j1)[:, None]
        jj = np.arange(0, j1)[None, :]
        low = ii > jj
        A[:j1, j0:j1] = np.where(low.T, blkA.T.conj(), A[:j1, j0:j1])
        B[:j1, j0:j1] = np.where(low.T, -blkB.T, B[:j1, j0:j1])
    return D


def uniform_case(n, seed, diag_shift=0.0):
    """Uniform workload: Re/Im of off-diagonal A and B entries ~ U(-1,1), A_ii ~ U(-1,1) (+ shift).

    This is the distribution of the 2n = 512 CPU case, the 2n = 8192 single
    solve and (with diag_shift = 10) the 2n = 32768 large solve."""

    def fill(j0, j1):
        s = _lower_slots(n, j0, j1)
        i = np.arange(n)[:, None]
        j = np.arange(j0, j1)[None, :]
        low = i > j
        a = (2.0 * uniform(seed, s) - 1.0) + 1j * (2.0 * uniform(seed, s + np.uint64(1)) - 1.0)
        b = (2.0 * uniform(seed, s + np.uint64(2)) - 1.0) + 1j * (2.0 * uniform(seed, s + np.uint64(3)) - 1.0)
        a = np.where(low, a, 0.0)
        b = np.where(low, b, 0.0)
        diag = (i == j)
        dvals = 2.0 * uniform(seed, s) - 1.0 + diag_shift
        a = np.where(diag, dvals, a)
        return a, b

    return _assemble(n, fill)


def fock_case(n, seed, decay=40.0, b_sigma=0.1):
    """Fock-like workload: A_ij = g (1+i)/sqrt(2) exp(-|i-j|/decay), g ~ N(0,1) real,
    diag(A) = linspace(-20, 5, n); B_ij = N(0, b_sigma) exp(-|i-j|/decay), real, antisymmetric.

    (Batched and serving workloads; sigma = 0.1 is read as a standard deviation.)"""
    diag = np.linspace(-20.0, 5.0, n)

    def fill(j0, j1):
        s = _lower_slots(n, j0, j1)
        i = np.arange(n)[:, None]
        j = np.arange(j0, j1)[None, :]
        low = i > j
        env = np.exp(-np.abs(i - j) / decay)
        g = _normal(seed, s, s + np.uint64(1))
        h = _normal(seed, s + np.uint64(2), s + np.uint64(3))
        a = np.where(low, g * (1.0 + 1.0j) / np.sqrt(2.0) * env, 0.0)
        b = np.where(low, b_sigma * h * env, 0.0).astype(np.complex128)
        a = np.where(i == j, diag[np.clip(i, 0, n - 1)] + 0.0 * j, a)
        return a, b

    return _assemble(n, fill)


def full_hermitian(D, n):
    """The full 2n x 2n matrix H = [[A, -conj(B)], [B, conj(A)]] from the left half of D."""
    A = D[:n, :n]
    B = D[n:2 * n, :n]
    return np.block([[A, -B.conj()], [B, A.conj()]])


def batched_sizes(count, seed, lo=64, hi=1024):
    """Even 2n values drawn uniformly from [lo, hi] (batched workload)."""
    u = uniform(seed, np.arange(count))
    k = (lo // 2 + np.floor(u * ((hi - lo) // 2 + 1))).astype(np.int64)
    return 2 * k


def flops_vectors(n):
    """Algorithmic flop count of one solve with eigenvectors, n = half-dimension:
    (64/3) n^3 for the reduction plus 32 n^3 for the back-transform."""
    return 160.0 / 3.0 * float(n) ** 3


def flops_values(n):
    """Algorithmic flop count of one eigenvalues-only solve."""
    return 64.0 / 3.0 * float(n) ** 3
