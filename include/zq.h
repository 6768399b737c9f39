/*
 * zq.h -- C ABI of libzquatev_amd, the MI355X (gfx950) quaternion-Hermitian
 * eigensolver.  Plain pointers and sizes only.
 *
 * The problem (zquatev.h:41-52 of the reference): diagonalise the 2n x 2n
 * Hermitian matrix
 *
 *        H = [ A  -conj(B) ]      A Hermitian, B complex antisymmetric (B^T = -B)
 *            [ B   conj(A) ]
 *
 * given its left half, and return the n distinct eigenvalues (each is doubly
 * degenerate in H) and the Kramers-paired eigenvectors [[U, -conj V], [V, conj U]].
 *
 * Matrix arguments are complex128 arrays passed as double* (interleaved real,
 * imaginary), column-major with a leading dimension counted in complex
 * elements, exactly as std::complex<double>* in the reference.
 *
 * Return codes: 0 success; -i argument i is invalid (LAPACK style); > 0 the
 * tridiagonal eigensolver did not converge (as zhbev's info at zquatev.cc:84,
 * :99: a leaf QL iteration, a secular-equation root or an eigenvalue bracket
 * that did not converge; the value is 1 + a local row or root index);
 * ZQ_ERR_* below for runtime errors and for a non-finite input.
 */
#ifndef ZQ_H
#define ZQ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZQ_OK 0
#define ZQ_ERR_HIP (-1001)    /* a HIP runtime call failed */
#define ZQ_ERR_NOMEM (-1002)  /* device memory exhausted */
#define ZQ_ERR_ARG (-1003)    /* unsupported internal configuration */
#define ZQ_ERR_GRAPH (-1004)  /* graph capture: a step that cannot be captured, or resources held by graphs */
#define ZQ_ERR_NONFINITE (-1005)  /* the input holds a NaN or an infinity (or its magnitude overflowed the
                                     reduction); eigenvalues are returned as NaN, vectors are unspecified.  An
                                     addition of this build: the reference does not check its input. */
#define ZQ_ERR_BUDGET (-1006) /* serving: the request alone needs more device memory than the server's budget */

typedef struct zq_context zq_context;

/* Library version as 10000*major + 100*minor + patch. */
int zq_version(void);

/* Human-readable text for a return code. */
const char* zq_error_string(int code);

/* Create a context bound to HIP device `device`: its own stream and a device
 * workspace cached across calls.  The stream is created when work first runs
 * on it, so a context that is given a caller's stream (zq_set_stream) never
 * holds one: each stream takes one of the process's GPU_MAX_HW_QUEUES
 * hardware queues (4 by default), and streams beyond those share queues and
 * serialise.  No counterpart in the reference, whose entry point allocates
 * per call (zquatev.cc:63-66). */
int zq_create(zq_context** ctx, int device);
int zq_destroy(zq_context* ctx);

/* Use the caller's HIP stream (a hipStream_t passed as void*; NULL is the HIP
 * null stream).  The device entry points are asynchronous on it. */
int zq_set_stream(zq_context* ctx, void* stream);

/* Go back to the stream the context created (the default after zq_create). */
int zq_use_own_stream(zq_context* ctx);

/* Panel width (quaternion columns per panel, 1..128).  0 selects the tuned
 * table.  The reference hard-codes nb = 20 (zquatev.cc:64). */
int zq_set_panel_width(zq_context* ctx, int nb);

/* Panel width the built-in table picks for half-dimension n. */
int zq_tuned_panel_width(int n);

/* Panel-width and tile-shape autotuner (no counterpart in the reference,
 * which hard-codes nb = 20 at zquatev.cc:64).  For each half-dimension
 * sizes[k] it times the reduction, the phase the panel width changes, with
 * nb = 16, 32, 64 and 128
 * on a synthetic U(-1,1) matrix generated on the device, and records the
 * fastest for the size bucket floor(log2 n) of this context; later solves of
 * any n in that bucket use it unless zq_set_panel_width set an override.
 * chosen[k] (may be NULL) receives the pick; ms[4k + t] (may be NULL) the
 * best of two timed runs in milliseconds for candidate t, -1 where the
 * candidate was skipped (wider than the matrix).  With the pick it then
 * tunes the bucket's tile shapes (zq_set_tile_shape below).  Synchronous;
 * uses the context's workspace.  Returns 0, or -2 (sizes NULL or an entry < 1), -3
 * (count < 0), or a ZQ_ERR_* code. */
int zq_autotune(zq_context* ctx, const int* sizes, int count, int* chosen, double* ms);

/* Panel width a solve of half-dimension n on this context uses now. */
int zq_panel_width(zq_context* ctx, int n);

/* Tile shapes (no counterpart in the reference, whose blocking is fixed at
 * nb = 20, zquatev.cc:64).  zq_autotune also times, per size bucket and
 * after the panel width, the panel matvec's strip switch -- narrow column
 * strips (16 quaternions) for trailing sizes up to this many rows, wide (32)
 * above; candidates 512, 768, 1024, 1536; default 768 -- and the quaternion
 * GEMM's tile rows -- 64-row tiles for launches of at least this many
 * 64 x 64 tiles per compute unit, else 32-row tiles; candidates 4, 16, 64;
 * default 16.  zq_set_tile_shape overrides both (switch -1 and threshold 0
 * restore the tuned values; switch 0..2048, -2 / -3 when out of range).
 * The switch changes the matvec's summation order, so a batch groups its
 * problems by it (as by panel width); the tile rows change no result.
 * zq_tile_shape reports what a solve of half-dimension n uses now. */
int zq_set_tile_shape(zq_context* ctx, int matvec_switch_rows, int gemm64_tiles_per_cu);
int zq_tile_shape(zq_context* ctx, int n, int* matvec_switch_rows, int* gemm64_tiles_per_cu);

/* HIP graph capture (no counterpart in the reference, a host library with no
 * streams).  zq_zquatev_device and zq_zquatev_batched_device (with info =
 * NULL) may be called while the context's stream is being captured
 * (hipStreamBeginCapture on the stream set with zq_set_stream): the solve is
 * recorded, nothing runs, and replays of the graph solve the same buffers
 * again (the caller refreshes the inputs in place between replays).  For
 * that, the call must find its workspace already allocated (run the same
 * call once outside capture first), descriptor uploads are recorded from a
 * pinned pool the context owns (8 MiB at zq_create), and from the first
 * capture on the workspace is held: a later call needing more workspace
 * returns ZQ_ERR_GRAPH until zq_graph_release.  Calls that must wait
 * (info != NULL, zq_autotune) return ZQ_ERR_GRAPH during a capture.
 *
 * zq_graph_reserve: size the pinned upload pool to `bytes` (only while no
 * captured graph holds it).  zq_graph_release: the caller has destroyed
 * every graph captured on this context; frees the pool's contents and the
 * workspace for reuse and reallocation. */
int zq_graph_reserve(zq_context* ctx, size_t bytes);
int zq_graph_release(zq_context* ctx);

/* Drop-in semantics of ts::zquatev (zquatev.h:54, zquatev.cc:42-100) on host
 * memory.  D: nld2 x n2 complex128 (the full buffer is used; only the left
 * half is read on input).  eig: n2/2 doubles, ascending on output.  On return
 * D holds [[U, -conj V], [V, conj U]]; columns i and i + n2/2 are the Kramers
 * pair of eig[i].  Synchronous.  Uses a process-wide context on device 0. */
int zq_zquatev(int n2, double* D, int nld2, double* eig);

/* Same on the given context (host memory, synchronous). */
int zq_zquatev_ctx(zq_context* ctx, int n2, double* D, int nld2, double* eig);

/* Device-memory variant: dD (device, nld2 x n2 complex128), deig (device,
 * n2/2 doubles).  jobz = 'V' computes eigenvectors into dD as above; jobz =
 * 'N' computes eigenvalues only and leaves dD unchanged.  Asynchronous on the
 * context's stream: the return value covers argument checks and launches,
 * and zq_last_info gives the solve's status (0, > 0, ZQ_ERR_NONFINITE). */
int zq_zquatev_device(zq_context* ctx, int n2, double* dD, int nld2, double* deig, int jobz);

/* Batched solves of independent problems of uneven sizes in one call.  The
 * reference solves one matrix per call (zquatev.cc:42-72); a caller with many
 * matrices loops over ts::zquatev.  Problem i is (n2[i], D[i], ld2[i],
 * eig[i]) with the meaning of zq_zquatev_device's arguments; n2[i] = 0 is
 * skipped.  Problems are grouped by panel width and each group runs its
 * columns in lockstep, so every problem's result is bitwise the result of its
 * own zq_zquatev_device call on the same context.
 * Device-memory variant: asynchronous on the context's stream unless info is
 * given (host array of count ints), in which case it waits and stores each
 * problem's status there (0, or > 0: tridiagonal solver did not converge).
 * Returns 0; ZQ_ERR_ARG for a null context; -2 count < 0; -3..-6 a null n2,
 * D, ld2 or eig array; -7 bad jobz; -8 some problem's own arguments are
 * invalid (nothing is run; info[i] holds that problem's zq_zquatev code,
 * -1..-4); ZQ_ERR_NONFINITE if some problem's input held a NaN or an
 * infinity, else 1 if some problem did not converge (info[i] holds each
 * problem's own code: 0, > 0 or ZQ_ERR_NONFINITE; the other problems are
 * solved normally); ZQ_ERR_HIP / ZQ_ERR_NOMEM.  Without info the call is
 * asynchronous and zq_last_info reports the aggregate. */
int zq_zquatev_batched_device(zq_context* ctx, int count, const int* n2, double* const* D, const int* ld2,
                              double* const* eig, int jobz, int* info);

/* ---- Serving loop ----------------------------------------------------------
 * No counterpart in the reference (one synchronous ts::zquatev call per
 * matrix, zquatev.cc:42-100).  Requests are queued by zq_server_submit and
 * served by continuous batching: a scheduler thread batches whatever has
 * arrived (the oldest request and those after it of the same jobz and size
 * class -- 2n up to 384, up to 1280, larger -- at most max_batch problems and
 * max_batch_gb of workspace) onto one of `lanes` contexts (own stream,
 * resident workspace); batches on different lanes run concurrently.  Each request's result is
 * bitwise that of its own zq_zquatev_device call.  Device pointers stay
 * owned by the caller and must stay valid until the request's completion
 * record has been returned by zq_server_poll.  Times are microseconds of a
 * steady clock started by zq_server_create. */
typedef struct zq_server zq_server;
typedef struct {
  unsigned long long ticket;
  int info;         /* 0; > 0 no convergence; ZQ_ERR_NONFINITE, ZQ_ERR_BUDGET, or another error code */
  int batch_size;   /* problems in the batch that served the request */
  int lane;
  double t_submit_us, t_start_us, t_done_us;
} zq_completion;

/* lanes 1..8 (-3), max_batch >= 1 (-4), max_batch_gb >= 0, 0 = no limit (-5).
 * Each lane is a stream: for its batches to overlap, the process's streams in
 * use (the lanes plus the caller's own) must not exceed GPU_MAX_HW_QUEUES
 * (4 by default, so up to 3 lanes beside one caller stream).
 * The lanes' workspaces together stay within a device-memory budget, by
 * default 80% of the device memory free at creation (zq_server_set_budget).
 * A lane's workspace grows only while no other lane has work in flight (a
 * reallocation waits for the whole device): a batch is cut to what its lane
 * holds while others run, and a request that needs more waits for them.  A
 * request that alone needs more than the budget completes at once with info
 * ZQ_ERR_BUDGET; a request whose input holds a NaN or an infinity completes
 * with ZQ_ERR_NONFINITE.  Neither affects the other requests. */
int zq_server_create(zq_server** out, int device, int lanes, int max_batch, double max_batch_gb);
/* Device-memory budget of the lanes' workspaces in GiB (> 0), or 0 for the
 * default; applies to batches scheduled from now on (-2: negative). */
int zq_server_set_budget(zq_server* srv, double budget_gb);
double zq_server_budget_gb(zq_server* srv);
/* Queue one request; arguments as zq_zquatev_device (errors -2..-6 for n2,
 * dD, ld2, deig, jobz).  Returns at once with its ticket. */
int zq_server_submit(zq_server* srv, int n2, double* dD, int ld2, double* deig, int jobz,
                     unsigned long long* ticket);
/* Move up to max completion records into out; returns how many (never waits). */
int zq_server_poll(zq_server* srv, zq_completion* out, int max);
/* Wait until every submitted request has completed. */
int zq_server_drain(zq_server* srv);
/* The server's clock (microseconds since creation). */
double zq_server_clock_us(zq_server* srv);
/* Drain, stop the threads, free the lanes. */
int zq_server_destroy(zq_server* srv);

/* Host-memory batch with eigenvectors (ts::zquatev semantics per problem),
 * synchronous.  Same return codes; info optional. */
int zq_zquatev_batched(zq_context* ctx, int count, const int* n2, double* const* D, const int* ld2,
                       double* const* eig, int* info);

/* Component entry points (device memory, asynchronous).
 *
 * zq_tridiagonalize: the symplectic reduction of the reference's panel loop
 * (zquatev.cc:68-72 over impl::panel_update, blocked.cc:45, and
 * impl::unblocked_update, unblocked.cc:44).  Reads the left half of dD and
 * writes the real tridiagonal it reduces to: dd (n doubles, diagonal) and
 * de (n-1 doubles, subdiagonal magnitudes) -- the (phase-scaled) matrix the
 * reference passes to zhbev at zquatev.cc:79-84.
 */
int zq_tridiagonalize(zq_context* ctx, int n2, const double* dD, int nld2, double* dd, double* de);

/* zq_tridiag_eig: eigenpairs of the real symmetric tridiagonal (dd, de) of
 * order n, the tridiagonal step of zquatev.cc:74-84 (zhbev_ with kd = 1, after
 * its phase scaling).  dw: n eigenvalues ascending; dZ (n x n, ld n, may be
 * NULL for eigenvalues only): orthonormal eigenvectors. */
int zq_tridiag_eig(zq_context* ctx, int n, const double* dd, const double* de, double* dw, double* dZ);

/* zq_qgemm: the quaternion GEMM on FP64 MFMA that runs the trailing update
 * (blocked.cc:416-475's zgemm3m_ calls) and the back-transform
 * (zquatev.cc:86-90): C (m x n) = alpha op(A) op(B) + beta C with op(X) = X
 * (trans 'N') or its quaternion conjugate transpose (trans 'C').  A
 * quaternion q = (a, b), i.e. the 2 x 2 complex block [[a, -conj b], [b,
 * conj a]], is stored as the 4 doubles (Re a, Im a, Re b, Im b); leading
 * dimensions count quaternions; lower_only != 0 writes only C[i][j] with
 * i >= j.  Device memory, asynchronous.  Tile rows follow the context's
 * GEMM threshold for trailing size m (zq_set_tile_shape); they change no
 * result.  Returns 0, -2 / -3 bad trans, -4 a negative size, -5 / -6 a null
 * operand, -7 a leading dimension too small. */
int zq_qgemm(zq_context* ctx, int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
             const double* B, int ldb, double beta, double* C, int ldc, int lower_only);

/* Device-side timings of the last solve on ctx, in milliseconds:
 * ms[0] reduction, ms[1] tridiagonal eigensolve, ms[2] back-transform,
 * ms[3] layout in/out, ms[4] total.  Returns the number of values written. */
int zq_last_timings(zq_context* ctx, double* ms, int count);

/* Kernel-class timing for measurement harnesses: with instrumentation on,
 * each launch of an instrumented kernel class is bracketed by HIP events on
 * the context's stream.  Classes: 0 panel Hermitian matvec (units: bytes of
 * the trailing lower triangle read), 1 trailing rank-2nb update (units:
 * flops), 2 back-transform products (flops), 3 divide & conquer merge GEMMs.
 * zq_instrument_report sums the last solve's launches of one class. */
int zq_set_instrument(zq_context* ctx, int on);
int zq_instrument_report(zq_context* ctx, int cls, double* total_ms, int* count, double* units);

/* Self-test of the MFMA register layout the kernels rely on: out (64 x 4
 * doubles, device) = raw accumulator of one v_mfma_f64_16x16x4_f64 with
 * A and B (64 doubles each, device) as the per-lane operands. */
int zq_selftest_mfma(zq_context* ctx, const double* dA, const double* dB, double* dOut);

/* ---- An entry point that allocates nothing ----------------------------------
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation.  Eigenvalues only -- the kernels of
 * zq_zquatev_device with jobz = 'N' and the built-in panel width, so the
 * result equals that call's on a context with no panel-width override or
 * autotuning -- of one problem of order n2 <= ZQ_NOALLOC_MAX_N2 (dD, nld2,
 * deig as there; dD is only read), on the current HIP device and the given
 * stream (a hipStream_t as void*, NULL for the null stream; asynchronous).
 * Every scratch buffer lies in one fixed device array of
 * ZQ_NOALLOC_SCRATCH_BYTES compiled into the library (each device the
 * library runs on holds it until the process ends); the call allocates
 * nothing and copies nothing from the host.  All calls on one device share
 * that array and leave their intermediate data in it: calls that overlap on
 * one device race -- serialise them (one stream, or events).  dinfo (device
 * int, may be NULL) receives the status zq_last_info would report (0, > 0
 * no convergence, ZQ_ERR_NONFINITE).  Returns 0, -1..-4 as zq_zquatev, -5 for
 * n2 above the maximum, ZQ_ERR_HIP. */
#define ZQ_NOALLOC_MAX_N2 2048
#define ZQ_NOALLOC_SCRATCH_BYTES (48u << 20)
int zq_eigvals_noalloc(int n2, const double* dD, int nld2, double* deig, int* dinfo, void* stream);

/* Stream synchronisation helper for callers without their own HIP runtime binding. */
int zq_synchronize(zq_context* ctx);

/* Status of the last solve issued on ctx by an asynchronous entry point
 * (zq_zquatev_device, zq_tridiagonalize, zq_tridiag_eig, or a batched call):
 * waits for the context's stream, then returns 0, ZQ_ERR_NONFINITE, or > 0
 * for no convergence -- the code zhbev's info carries in the reference
 * (zquatev.cc:84, :99), which the host entry points return directly.  For a
 * batch, the aggregate its batched call returns. */
int zq_last_info(zq_context* ctx);

/* Iteration caps of the tridiagonal eigensolver on ctx (0 = default): QL
 * sweeps per eigenvalue of a divide & conquer leaf (60), safeguarded Newton
 * steps per secular-equation root (200), multisection steps per eigenvalue
 * of the eigenvalues-only path (64).  A cap reached is reported as
 * non-convergence (info > 0), as zhbev reports a QL failure (zquatev.cc:84).
 * Meant for tests of that path; -2 for a negative cap. */
int zq_set_iteration_caps(zq_context* ctx, int leaf, int secular, int bisect);

/* Cap the device workspace ctx may hold (bytes; 0 = no cap but free device
 * memory).  A batch whose workspace exceeds the cap, or what the device can
 * give, is solved as consecutive runs of its size-sorted problems, each run
 * fitting (results unchanged: each member still equals its solo solve); a
 * single problem over the cap returns ZQ_ERR_NOMEM.  No counterpart in the
 * reference, which allocates per call (zquatev.cc:63-66). */
int zq_set_workspace_limit(zq_context* ctx, size_t bytes);

/* Device workspace bytes a batched call of these sizes (jobz 'V' or 'N')
 * needs on ctx (NULL: the built-in panel-width table), split into runs that
 * fit `budget` bytes (0: one run per panel-width group).  *bytes receives
 * the largest run's workspace, run[i] (optional, count ints) problem i's run
 * (-1 for an empty problem).  Returns the number of runs, or -2 count < 0,
 * -3 bad sizes, -4 bad jobz.  Host arithmetic only: no device call. */
int zq_batch_workspace(zq_context* ctx, int count, const int* n2, int jobz, size_t budget, size_t* bytes, int* run);

#ifdef __cplusplus
}
#endif

#endif /* ZQ_H */
