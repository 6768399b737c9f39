"""MI355X (gfx950) quaternion-Hermitian eigensolver -- Python binding of the C ABI.

``zquatev(D)`` mirrors the reference entry point ``ts::zquatev(n2, D, ld, eig)``
(zquatev.h:54): D is a column-major (Fortran-order) complex128 array of shape
(2n, 2n) whose left half holds [A; B]; on return D holds [[U, -conj V],
[V, conj U]] and the function returns (info, eig) with eig ascending.

All arithmetic runs in libzquatev_amd.so (HIP kernels for gfx950).  There is
no CPU fallback: every entry point raises ImportError if the library is missing.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzquatev_amd.so")

_c_int = ctypes.c_int
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every C-ABI entry point in include/zq.h
_SIGNATURES = {
    "zq_version": (_c_int, []),
    "zq_error_string": (ctypes.c_char_p, [_c_int]),
    "zq_create": (_c_int, [ctypes.POINTER(_vp), _c_int]),
    "zq_destroy": (_c_int, [_vp]),
    "zq_set_stream": (_c_int, [_vp, _vp]),
    "zq_use_own_stream": (_c_int, [_vp]),
    "zq_set_panel_width": (_c_int, [_vp, _c_int]),
    "zq_tuned_panel_width": (_c_int, [_c_int]),
    "zq_autotune": (_c_int, [_vp, _vp, _c_int, _vp, _vp]),
    "zq_panel_width": (_c_int, [_vp, _c_int]),
    "zq_set_tile_shape": (_c_int, [_vp, _c_int, _c_int]),
    "zq_tile_shape": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "zq_graph_reserve": (_c_int, [_vp, ctypes.c_size_t]),
    "zq_graph_release": (_c_int, [_vp]),
    "zq_zquatev": (_c_int, [_c_int, _vp, _c_int, _vp]),
    "zq_zquatev_ctx": (_c_int, [_vp, _c_int, _vp, _c_int, _vp]),
    "zq_zquatev_device": (_c_int, [_vp, _c_int, _vp, _c_int, _vp, _c_int]),
    "zq_zquatev_batched_device": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp]),
    "zq_zquatev_batched": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "zq_server_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, ctypes.c_double]),
    "zq_server_submit": (_c_int, [_vp, _c_int, _vp, _c_int, _vp, _c_int, ctypes.POINTER(ctypes.c_ulonglong)]),
    "zq_server_poll": (_c_int, [_vp, _vp, _c_int]),
    "zq_server_drain": (_c_int, [_vp]),
    "zq_server_clock_us": (ctypes.c_double, [_vp]),
    "zq_server_destroy": (_c_int, [_vp]),
    "zq_tridiagonalize": (_c_int, [_vp, _c_int, _vp, _c_int, _vp, _vp]),
    "zq_tridiag_eig": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp]),
    "zq_qgemm": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_double, _vp, _c_int, _vp, _c_int,
                          ctypes.c_double, _vp, _c_int, _c_int]),
    "zq_last_timings": (_c_int, [_vp, ctypes.POINTER(ctypes.c_double), _c_int]),
    "zq_selftest_mfma": (_c_int, [_vp, _vp, _vp, _vp]),
    "zq_synchronize": (_c_int, [_vp]),
    "zq_last_info": (_c_int, [_vp]),
    "zq_set_workspace_limit": (_c_int, [_vp, ctypes.c_size_t]),
    "zq_set_iteration_caps": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "zq_batch_workspace": (_c_int, [_vp, _c_int, _vp, _c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _vp]),
    "zq_eigvals_noalloc": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp]),
    "zq_server_set_budget": (_c_int, [_vp, ctypes.c_double]),
    "zq_server_budget_gb": (ctypes.c_double, [_vp]),
    "zq_set_instrument": (_c_int, [_vp, _c_int]),
    "zq_instrument_report": (_c_int, [_vp, _c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_c_int),
                                      ctypes.POINTER(ctypes.c_double)]),
}

_LIB = None


class _Fns:
    pass


_F = _Fns()


def lib():
    """Load libzquatev_amd.so (once).  Raises ImportError if it was not built:
    there is no CPU fallback."""
    global _LIB
    if _LIB is None:
        # PyTorch ships its own HIP runtime under the same soname; when torch is
        # present it must be loaded first so that the process has one runtime
        # and torch's device pointers are valid here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C zquatev_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
            setattr(_F, name, f)
        _LIB = L
    return _LIB


def _fn(name):
    lib()
    return getattr(_F, name)


ERR_HIP = -1001
ERR_NOMEM = -1002
ERR_ARG = -1003
ERR_GRAPH = -1004
ERR_NONFINITE = -1005
ERR_BUDGET = -1006


class ZQError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        msg = _fn("zq_error_string")(code).decode()
        super().__init__(f"{where}: {msg} (code {code})" if where else f"{msg} (code {code})")


def version():
    return _fn("zq_version")()


def error_string(code):
    return _fn("zq_error_string")(code).decode()


def tuned_panel_width(n):
    return _fn("zq_tuned_panel_width")(n)


def _check(rc, where):
    if rc != 0:
        raise ZQError(rc, where)


def _fortran_c128(D):
    if not isinstance(D, np.ndarray) or D.dtype != np.complex128 or D.ndim != 2:
        raise TypeError("D must be a 2-D complex128 numpy array")
    if not D.flags.f_contiguous:
        raise ValueError("D must be column-major (Fortran order)")
    return D


def _host_n2(D, n2):
    """The order n2 solved in D: all of D's columns by default; an explicit n2
    must fit D (the C entry writes n2 columns of D.shape[0] rows)."""
    if n2 is None:
        return D.shape[1]
    n2 = int(n2)
    if n2 > D.shape[1] or n2 > D.shape[0]:
        raise ValueError(f"n2={n2} does not fit D of shape {D.shape}")
    return n2


def batch_workspace(sizes, jobz="V", budget=0, ctx=None):
    """Device workspace (bytes) a batched call of these orders 2n needs, and
    each problem's run when the batch is split to fit `budget` bytes (0: no
    split).  Host arithmetic only (no device call).  Returns (bytes, runs,
    run index per problem)."""
    sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32).reshape(-1))
    run = np.zeros(max(sz.size, 1), dtype=np.int32)
    nbytes = ctypes.c_size_t()
    h = ctx.handle if ctx is not None else None
    k = _fn("zq_batch_workspace")(h, int(sz.size), sz.ctypes.data, ord(jobz), int(budget), ctypes.byref(nbytes),
                                  run.ctypes.data)
    if k < 0:
        raise ZQError(k, "zq_batch_workspace")
    return nbytes.value, k, [int(v) for v in run[: sz.size]]


def zquatev(D, n2=None):
    """In-place drop-in for ts::zquatev on host memory (zquatev.cc:42).

    D: (ld2, >= n2) complex128, Fortran order; the left half of the leading
    n2 x n2 block holds [A; B].  Returns (info, eig[n2/2])."""
    D = _fortran_c128(D)
    n2 = _host_n2(D, n2)
    ld2 = D.shape[0]
    eig = np.zeros(max(n2 // 2, 1), dtype=np.float64)
    info = _fn("zq_zquatev")(n2, D.ctypes.data, ld2, eig.ctypes.data)
    return info, eig[: n2 // 2]


class Context:
    """A device context: one HIP stream and a cached device workspace."""

    def __init__(self, device=0, panel_width=0):
        h = _vp()
        _check(_fn("zq_create")(ctypes.byref(h), device), "zq_create")
        self._h = h
        self.device = device
        self._stream = None     # the stream set with set_stream (None: the context's own)
        if panel_width:
            self.set_panel_width(panel_width)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _fn("zq_destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_ptr):
        """Run on the given HIP stream handle (0 = the null stream, which is
        torch's default stream)."""
        _check(_fn("zq_set_stream")(self._h, _vp(stream_ptr or 0)), "zq_set_stream")
        self._stream = int(stream_ptr or 0)

    def use_own_stream(self):
        _check(_fn("zq_use_own_stream")(self._h), "zq_use_own_stream")
        self._stream = None

    def set_panel_width(self, nb):
        _check(_fn("zq_set_panel_width")(self._h, int(nb)), "zq_set_panel_width")

    def synchronize(self):
        _check(_fn("zq_synchronize")(self._h), "zq_synchronize")

    def last_info(self):
        """Status of the last asynchronous solve on this context (waits for
        its stream): 0, > 0 (no convergence), ERR_NONFINITE, ..."""
        return int(_fn("zq_last_info")(self._h))

    def set_iteration_caps(self, leaf=0, secular=0, bisect=0):
        """Iteration caps of the tridiagonal solver (0: default 60 / 200 / 64);
        reaching one is reported as non-convergence (info > 0)."""
        _check(_fn("zq_set_iteration_caps")(self._h, int(leaf), int(secular), int(bisect)),
               "zq_set_iteration_caps")

    def set_workspace_limit(self, nbytes):
        """Cap the device workspace this context may hold (0: no cap)."""
        _check(_fn("zq_set_workspace_limit")(self._h, int(nbytes)), "zq_set_workspace_limit")

    def autotune(self, sizes):
        """Time the reduction with panel widths 16, 32, 64 and 128 for each
        half-dimension in ``sizes`` and keep the fastest for that size's
        power-of-two bucket on this context.  Returns ``(chosen, ms)``:
        the pick per size and an array (len(sizes), 4) of the timings in ms
        (-1 where a width was skipped as wider than the matrix)."""
        sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32).reshape(-1))
        chosen = np.zeros(max(sz.size, 1), dtype=np.int32)
        ms = np.zeros((max(sz.size, 1), 4), dtype=np.float64)
        _check(_fn("zq_autotune")(self._h, sz.ctypes.data, int(sz.size), chosen.ctypes.data, ms.ctypes.data),
               "zq_autotune")
        return [int(v) for v in chosen[: sz.size]], ms[: sz.size]

    def graph_reserve(self, nbytes):
        """Size the pinned pool that uploads recorded into captured graphs
        come from (only while no captured graph holds it)."""
        _check(_fn("zq_graph_reserve")(self._h, int(nbytes)), "zq_graph_reserve")

    def graph_release(self):
        """Every graph captured on this context has been destroyed: free the
        pool's contents and let the workspace be reallocated again."""
        _check(_fn("zq_graph_release")(self._h), "zq_graph_release")

    def set_tile_shape(self, matvec_switch=-1, gemm64_tiles_per_cu=0):
        """Override the matvec strip switch (trailing rows up to which narrow
        strips are used, 0..2048) and the GEMM tile-rows threshold (64-row
        tiles from this many 64 x 64 tiles per CU); -1 / 0 restore the tuned
        (or default) values."""
        _check(_fn("zq_set_tile_shape")(self._h, int(matvec_switch), int(gemm64_tiles_per_cu)),
               "zq_set_tile_shape")

    def tile_shape(self, n):
        """(matvec strip switch, GEMM 64-row threshold) a solve of half-dimension n uses."""
        a, b = _c_int(), _c_int()
        _check(_fn("zq_tile_shape")(self._h, int(n), ctypes.byref(a), ctypes.byref(b)), "zq_tile_shape")
        return a.value, b.value

    def panel_width(self, n):
        """Panel width a solve of half-dimension n on this context uses."""
        return int(_fn("zq_panel_width")(self._h, int(n)))

    def zquatev(self, D, n2=None):
        """Host-memory solve on this context; same contract as :func:`zquatev`."""
        D = _fortran_c128(D)
        n2 = _host_n2(D, n2)
        eig = np.zeros(max(n2 // 2, 1), dtype=np.float64)
        info = _fn("zq_zquatev_ctx")(self._h, n2, D.ctypes.data, D.shape[0], eig.ctypes.data)
        return info, eig[: n2 // 2]

    # device-pointer entry points: pointers are integers (e.g. torch .data_ptr())
    def zquatev_device(self, n2, d_ptr, ld2, eig_ptr, jobz="V"):
        return _fn("zq_zquatev_device")(self._h, int(n2), _vp(d_ptr), int(ld2), _vp(eig_ptr), ord(jobz))

    def zquatev_batched(self, Ds):
        """Host-memory batch: Ds is a list of column-major complex128 arrays
        (each solved in place, as :meth:`zquatev`).  Returns (rc, infos, eigs)."""
        Ds = [_fortran_c128(D) for D in Ds]
        for D in Ds:
            if D.shape[1] > D.shape[0]:
                raise ValueError(f"D of shape {D.shape}: the order (columns) exceeds the leading dimension (rows)")
        cnt = len(Ds)
        n2 = (_c_int * max(cnt, 1))(*[D.shape[1] for D in Ds])
        ld2 = (_c_int * max(cnt, 1))(*[D.shape[0] for D in Ds])
        eigs = [np.zeros(max(D.shape[1] // 2, 1), dtype=np.float64) for D in Ds]
        dp = (_vp * max(cnt, 1))(*[D.ctypes.data for D in Ds])
        ep = (_vp * max(cnt, 1))(*[e.ctypes.data for e in eigs])
        info = (_c_int * max(cnt, 1))()
        rc = _fn("zq_zquatev_batched")(self._h, cnt, n2, dp, ld2, ep, info)
        return rc, list(info)[:cnt], [e[: D.shape[1] // 2] for e, D in zip(eigs, Ds)]

    def zquatev_batched_device(self, n2s, d_ptrs, ld2s, eig_ptrs, jobz="V", want_info=False):
        """Device-pointer batch (integers, e.g. torch .data_ptr()).  Returns rc,
        or (rc, infos) with want_info (which waits for completion)."""
        cnt = len(n2s)
        n2 = (_c_int * max(cnt, 1))(*[int(v) for v in n2s])
        ld2 = (_c_int * max(cnt, 1))(*[int(v) for v in ld2s])
        dp = (_vp * max(cnt, 1))(*[int(v) for v in d_ptrs])
        ep = (_vp * max(cnt, 1))(*[int(v) for v in eig_ptrs])
        info = (_c_int * max(cnt, 1))() if want_info else None
        rc = _fn("zq_zquatev_batched_device")(self._h, cnt, n2, dp, ld2, ep, ord(jobz), info)
        return (rc, list(info)[:cnt]) if want_info else rc

    def tridiagonalize_device(self, n2, d_ptr, ld2, dd_ptr, de_ptr):
        return _fn("zq_tridiagonalize")(self._h, int(n2), _vp(d_ptr), int(ld2), _vp(dd_ptr), _vp(de_ptr))

    def tridiag_eig_device(self, n, dd_ptr, de_ptr, dw_ptr, dz_ptr=0):
        return _fn("zq_tridiag_eig")(self._h, int(n), _vp(dd_ptr), _vp(de_ptr), _vp(dw_ptr), _vp(dz_ptr or 0))

    def qgemm_device(self, transa, transb, m, n, k, alpha, a_ptr, lda, b_ptr, ldb, beta, c_ptr, ldc,
                     lower_only=False):
        """Quaternion GEMM on device buffers of 4 doubles per quaternion:
        C = alpha op(A) op(B) + beta C, op = 'N' or 'C' (conjugate transpose)."""
        return _fn("zq_qgemm")(self._h, ord(transa), ord(transb), int(m), int(n), int(k), float(alpha),
                               _vp(a_ptr), int(lda), _vp(b_ptr), int(ldb), float(beta), _vp(c_ptr), int(ldc),
                               1 if lower_only else 0)

    def selftest_mfma_device(self, a_ptr, b_ptr, out_ptr):
        return _fn("zq_selftest_mfma")(self._h, _vp(a_ptr), _vp(b_ptr), _vp(out_ptr))

    def set_instrument(self, on=True):
        _check(_fn("zq_set_instrument")(self._h, 1 if on else 0), "zq_set_instrument")

    KERNEL_CLASSES = ("symv", "trail", "backtr", "merge")

    def instrument_report(self, cls):
        """(total_ms, launches, algorithmic units) of one kernel class in the last solve."""
        if isinstance(cls, str):
            cls = self.KERNEL_CLASSES.index(cls)
        t = ctypes.c_double()
        k = _c_int()
        u = ctypes.c_double()
        _check(_fn("zq_instrument_report")(self._h, cls, ctypes.byref(t), ctypes.byref(k), ctypes.byref(u)),
               "zq_instrument_report")
        return t.value, k.value, u.value

    def last_timings(self):
        """Device phase times of the last solve (ms): reduce, tridiagonal, back-transform, layout, total."""
        buf = (ctypes.c_double * 5)()
        k = _fn("zq_last_timings")(self._h, buf, 5)
        keys = ("reduce", "tridiag", "backtransform", "layout", "total")
        return {keys[i]: buf[i] for i in range(k)}


class Completion(ctypes.Structure):
    """zq_completion: one served request."""
    _fields_ = [("ticket", ctypes.c_ulonglong), ("info", _c_int), ("batch_size", _c_int), ("lane", _c_int),
                ("t_submit_us", ctypes.c_double), ("t_start_us", ctypes.c_double), ("t_done_us", ctypes.c_double)]


class Server:
    """Continuous-batching server (zq_server_*): submit device-pointer requests,
    poll completion records."""

    def __init__(self, device=0, lanes=3, max_batch=128, max_batch_gb=0.0):
        h = _vp()
        _check(_fn("zq_server_create")(ctypes.byref(h), device, lanes, max_batch, float(max_batch_gb)),
               "zq_server_create")
        self._h = h
        self._buf = (Completion * 1024)()

    def submit(self, n2, d_ptr, ld2, eig_ptr, jobz="V"):
        t = ctypes.c_ulonglong()
        _check(_fn("zq_server_submit")(self._h, int(n2), _vp(d_ptr), int(ld2), _vp(eig_ptr), ord(jobz),
                                       ctypes.byref(t)), "zq_server_submit")
        return t.value

    def poll(self):
        """Completed requests since the last poll, as dicts."""
        out = []
        while True:
            k = _fn("zq_server_poll")(self._h, self._buf, len(self._buf))
            if k < 0:
                raise ZQError(k, "zq_server_poll")
            for c in self._buf[:k]:
                out.append({f: getattr(c, f) for f, _ in Completion._fields_})
            if k < len(self._buf):
                return out

    def drain(self):
        _check(_fn("zq_server_drain")(self._h), "zq_server_drain")

    def set_budget(self, gib):
        """Device-memory budget of the lanes' workspaces in GiB (0: default,
        80% of the memory free at creation)."""
        _check(_fn("zq_server_set_budget")(self._h, float(gib)), "zq_server_set_budget")

    def budget_gb(self):
        return float(_fn("zq_server_budget_gb")(self._h))

    def clock_us(self):
        return _fn("zq_server_clock_us")(self._h)

    def close(self):
        if self._h:
            _fn("zq_server_destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def latency_stats(records, arrivals=None):
    """Serving statistics from completion records (dicts with t_submit_us /
    t_done_us, or (arrival, done) pairs in seconds when `arrivals` is given
    as the arrival times): p50 / p99 / mean latency in ms and solves/s =
    completed / (last completion - first arrival)."""
    if arrivals is not None:
        a = np.asarray(arrivals, dtype=np.float64)
        d = np.asarray(records, dtype=np.float64)
    else:
        a = np.array([r["t_submit_us"] for r in records]) * 1e-6
        d = np.array([r["t_done_us"] for r in records]) * 1e-6
    if len(a) == 0:
        return {"completed": 0}
    lat = (d - a) * 1000.0
    span = d.max() - a.min()
    return {"completed": int(len(a)), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()), "solves_per_s": float(len(a) / span) if span > 0 else float("inf")}


# torch helpers (torch is plumbing for device memory and streams only)
def _torch():
    import torch
    return torch


def solve_torch(ctx, D, jobz="V"):
    """Solve on a device tensor D (2n x 2n complex128, column-major, i.e. the
    transpose of a contiguous row-major tensor is accepted as D.T), in order
    on torch's current stream: the solve runs after the work already queued
    there (the producers of D) and later work there sees its results.
    Returns (rc, eig tensor); D is overwritten with the eigenvectors for
    jobz='V'.  rc is the launch status (argument checks, allocation); the
    solve's own status (convergence, non-finite input) is ctx.last_info(),
    which waits for the stream."""
    torch = _torch()
    n2 = D.shape[0]
    if D.dtype != torch.complex128 or not D.is_cuda or D.dim() != 2 or D.shape[1] != n2:
        raise ValueError("D must be a square complex128 device tensor")
    if D.stride(0) != 1:
        raise ValueError("D must be column-major (use a transposed contiguous tensor)")
    ld2 = D.stride(1)
    stream = torch.cuda.current_stream(D.device)
    eig = torch.zeros(max(n2 // 2, 1), dtype=torch.float64, device=D.device)   # queued on `stream`
    previous = ctx._stream
    ctx.set_stream(stream.cuda_stream)
    try:
        rc = ctx.zquatev_device(n2, D.data_ptr(), ld2, eig.data_ptr(), jobz)
    finally:
        if previous is None:
            ctx.use_own_stream()
        else:
            ctx.set_stream(previous)
    return rc, eig[: n2 // 2]
