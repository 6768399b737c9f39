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


# ---------------------------------------------------------------------------
# Device-side generation (torch) of the same distributions, for workloads too
# large to build on the host (batched and serving streams).  Same stream
# positions as the numpy versions; the float transforms (log1p, cos) come from
# the device math library, so values agree with the host ones to rounding.
# Only what the solver reads is written: the lower triangles (with diagonal)
# of A and B in the left half of a column-major 2n x 2n matrix.
# ---------------------------------------------------------------------------
def _i64(u):
    return u - (1 << 64) if u >= (1 << 63) else u


_T_GOLDEN = _i64(0x9E3779B97F4A7C15)
_T_M1 = _i64(0xBF58476D1CE4E5B9)
_T_M2 = _i64(0x94D049BB133111EB)


def _lsr(z, k):
    return (z >> k) & ((1 << (64 - k)) - 1)


def uniform_t(seed, idx):
    """torch version of :func:`uniform` on an int64 index tensor (wrapping arithmetic)."""
    z = idx.new_full((), _i64(seed % (1 << 64))) + (idx + 1) * _T_GOLDEN
    z = (z ^ _lsr(z, 30)) * _T_M1
    z = (z ^ _lsr(z, 27)) * _T_M2
    z = z ^ _lsr(z, 31)
    return _lsr(z, 11).to(dtype=__import__("torch").float64) * (1.0 / 9007199254740992.0)


def fock_case_device(n, seed, out, decay=40.0, b_sigma=0.1, chunk=256):
    """:func:`fock_case` written into ``out``: a torch complex128 tensor of
    shape (>= n columns, 2n + pad rows) holding column j of D in out[j] (i.e.
    the row-major view of a column-major D with ld2 = out.shape[1]).  Only
    the lower triangles of A and B are written."""
    import torch
    dev = out.device
    diag = torch.linspace(-20.0, 5.0, n, dtype=torch.float64, device=dev)
    i = torch.arange(n, dtype=torch.int64, device=dev)[None, :]
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        j = torch.arange(j0, j1, dtype=torch.int64, device=dev)[:, None]
        s = 4 * (j * n + i)                       # (w, n): column j, row i
        low = i > j
        env = torch.exp(-(i - j).abs().to(torch.float64) / decay)

        def normal(k0, k1):
            u1 = uniform_t(seed, k0)
            u2 = uniform_t(seed, k1)
            return torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * torch.pi * u2)

        g = normal(s, s + 1)
        h = normal(s + 2, s + 3)
        a = torch.where(low, g * env / (2.0 ** 0.5), torch.zeros((), dtype=torch.float64, device=dev))
        a = torch.complex(a, a)
        a = torch.where(i == j, torch.complex(diag[None, :].expand_as(g), torch.zeros_like(g)), a)
        b = torch.where(low, b_sigma * h * env, torch.zeros((), dtype=torch.float64, device=dev))
        out[j0:j1, :n] = a
        out[j0:j1, n:2 * n] = torch.complex(b, torch.zeros_like(b))
    return out


def full_hermitian_t(Dcols, n):
    """H (2n x 2n, torch) from the lower triangles written by fock_case_device
    (Dcols[j, i] = D[i, j])."""
    import torch
    A = Dcols[:n, :n].T
    B = Dcols[:n, n:2 * n].T
    L = torch.tril(A, -1)
    A = L + L.conj().T + torch.diag(torch.diagonal(A).real.to(A.dtype))
    Bl = torch.tril(B, -1)
    B = Bl - Bl.T
    return torch.cat([torch.cat([A, -B.conj()], 1), torch.cat([B, A.conj()], 1)], 0)


def serving_sizes(count, seed, lo=128, hi=4096):
    """Even 2n values log-uniform over [lo, hi] (serving workload)."""
    u = uniform(seed, np.arange(count))
    v = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    return (2 * np.clip(np.round(v / 2.0), lo // 2, hi // 2)).astype(np.int64)


def uniform_case_device(n, seed, out, diag_shift=0.0, chunk=256):
    """:func:`uniform_case` written into ``out`` (torch complex128, out[j] =
    column j of D, ld2 = out.shape[1]); only the lower triangles of A and B."""
    import torch
    dev = out.device
    i = torch.arange(n, dtype=torch.int64, device=dev)[None, :]
    zero = torch.zeros((), dtype=torch.float64, device=dev)
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        j = torch.arange(j0, j1, dtype=torch.int64, device=dev)[:, None]
        s = 4 * (j * n + i)
        low = i > j
        u0 = 2.0 * uniform_t(seed, s) - 1.0
        a_re = torch.where(low, u0, zero)
        a_im = torch.where(low, 2.0 * uniform_t(seed, s + 1) - 1.0, zero)
        b_re = torch.where(low, 2.0 * uniform_t(seed, s + 2) - 1.0, zero)
        b_im = torch.where(low, 2.0 * uniform_t(seed, s + 3) - 1.0, zero)
        a_re = torch.where(i == j, u0 + diag_shift, a_re)
        out[j0:j1, :n] = torch.complex(a_re, a_im)
        out[j0:j1, n:2 * n] = torch.complex(b_re, b_im)
    return out
