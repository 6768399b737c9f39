"""ctypes access to the compiled reference (oracle/_ref/libzquatev_ref.so) --
TEST INFRASTRUCTURE ONLY, used to pin the oracle and as the CPU baseline.

The library is built by oracle/Makefile from the reference's own sources.  It
links MKL (intel_thread); the thread count follows MKL_NUM_THREADS.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libzquatev_ref.so")

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        # RTLD_GLOBAL: MKL's dispatch libraries resolve mkl_core symbols globally
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        f = _lib._ZN2ts7zquatevEiPSt7complexIdEiPd
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        p = _lib._ZN2ts4impl12panel_updateEiiPSt7complexIdES3_S3_S3_iiS3_
        p.restype = None
        p.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        u = _lib._ZN2ts4impl16unblocked_updateEiPSt7complexIdES3_S3_S3_iiS3_
        u.restype = None
        u.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return _lib


def zquatev(D):
    """Call ts::zquatev (zquatev.cc:42) on a copy of D (2n x 2n, ld = 2n).
    Returns (info, X, eig)."""
    lib = _load()
    X = np.array(D, dtype=np.complex128, order="F", copy=True)
    n2 = X.shape[0]
    eig = np.zeros(max(n2 // 2, 1), dtype=np.float64)
    info = lib._ZN2ts7zquatevEiPSt7complexIdEiPd(n2, X.ctypes.data, n2, eig.ctypes.data)
    return info, X, eig[: n2 // 2]


def reduce_to_tridiagonal(A, B, nb=20):
    """Run the reference's reduction loop (the panel/unblocked dispatch of
    zquatev.cc:68-72, calling its exported impl::panel_update and
    impl::unblocked_update) and return (d, e, Q0, Q1) as the oracle does."""
    lib = _load()
    n = A.shape[0]
    ld = n
    D0 = np.array(A, dtype=np.complex128, order="F", copy=True)
    D1 = np.array(B, dtype=np.complex128, order="F", copy=True)
    Q0 = np.asfortranarray(np.eye(n, dtype=np.complex128))
    Q1 = np.zeros((n, n), dtype=np.complex128, order="F")
    work = np.zeros(nb * 5 + nb * nb * 15 + (max(4, nb) + 10 * nb + 3) * n, dtype=np.complex128)
    z = 16  # bytes per complex
    pu = lib._ZN2ts4impl12panel_updateEiiPSt7complexIdES3_S3_S3_iiS3_
    uu = lib._ZN2ts4impl16unblocked_updateEiPSt7complexIdES3_S3_S3_iiS3_
    for p in range(0, n, nb):
        off = (p * ld + p) * z
        qoff = p * ld * z
        if n - p > nb + 1:
            pu(n - p, nb, D0.ctypes.data + off, D1.ctypes.data + off,
               Q0.ctypes.data + qoff, Q1.ctypes.data + qoff, ld, n, work.ctypes.data)
        else:
            uu(n - p, D0.ctypes.data + off, D1.ctypes.data + off,
               Q0.ctypes.data + qoff, Q1.ctypes.data + qoff, ld, n, work.ctypes.data)
    d = np.real(np.diag(D0)).copy()
    e = np.array([D0[i + 1, i] for i in range(n - 1)], dtype=np.complex128)
    return d, e, Q0, Q1
