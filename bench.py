#!/usr/bin/env python3
"""Benchmark of the quaternion-Hermitian eigensolver on one MI355X.

Workload: one solve of 2n = 8192 (complex128, A Hermitian / B antisymmetric
with U(-1,1) entries, eigenvalues + Kramers-paired eigenvectors).  A step is
one call of the device entry point zq_zquatev_device on a matrix already in
HBM (the left half is restored from a pristine device copy first: an O(n^2)
device copy inside the step, < 0.2% of its time).

Prints one JSON line: FP64 GFLOP/s on the algorithmic count (160/3) n^3,
the roofline of the dominant kernel class measured with HIP events on the
library's stream, and a CPU baseline (the reference compiled from its own
sources, timed on this host's cores on a bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 GFLOP/s at 2n=8192; batched solves/s + p50/p99 latency vs plain loop"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n2", type=int, default=8192)
    ap.add_argument("--nb", type=int, default=0,
                    help="panel width (0: autotuned over 16/32/64/128 before the warm-up; the built-in table "
                         "with --profile-run)")
    ap.add_argument("--tsm", type=int, default=-1,
                    help="matvec strip switch (trailing rows with narrow strips; -1: autotuned, or the default "
                         "with --profile-run)")
    ap.add_argument("--tile64", type=int, default=0,
                    help="GEMM 64-row tiles from this many 64x64 tiles per CU (0: autotuned / default)")
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--diag-shift", type=float, default=0.0,
                    help="single: added to the diagonal of A (the 2n = 32768 large-solve case uses 10)")
    ap.add_argument("--cpu-n2", type=int, default=4096,
                    help="size of the bounded CPU-baseline sample (median of --cpu-reps runs on the host cores)")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--cpu-small-reps", type=int, default=20,
                    help="runs of the 2n=512 CPU workload (median), on all cores and on one")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving summary of the single workload")
    ap.add_argument("--summary-serve-seconds", type=float, default=15.0)
    ap.add_argument("--summary-loop-seconds", type=float, default=5.0)
    ap.add_argument("--summary-capacity", type=int, default=1500)
    ap.add_argument("--summary-pool", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-values-only", action="store_true", help="skip the eigenvalues-only (jobz='N') summary")
    ap.add_argument("--no-host-entry", action="store_true",
                    help="skip the host-memory (drop-in, PCIe-inclusive) summary")
    ap.add_argument("--workload", choices=("single", "batched", "serving", "qgemm"), default="single",
                    help="single: the 2n=8192 solve (plus a batched summary); batched: the batched workload only; "
                         "serving: the Poisson request stream through the server and the plain loop; qgemm: the "
                         "quaternion GEMM alone (zq_qgemm, C -= A B), for its MFMA roofline and counters")
    ap.add_argument("--qgemm-shape", default="4096,4096,1024", metavar="M,N,K",
                    help="qgemm: quaternion sizes of C (M x N) and the inner dimension K")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--loop-seconds", type=float, default=10.0, help="arrival window replayed through the plain loop")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--pool", type=int, default=1024)
    ap.add_argument("--capacity-requests", type=int, default=3000,
                    help="serving: requests of the same mix submitted back to back (saturation capacity; 0: skip)")
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--max-batch", type=int, default=128,
                    help="serving: most problems per batch (128: saturated capacity 509 solves/s against "
                         "~497 at 64; batches at the offered rate average ~1.2 problems either way)")
    ap.add_argument("--max-batch-gb", type=float, default=24.0)
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched summary of the single workload")
    ap.add_argument("--profile-run", action="store_true",
                    help="only warmup + timed steps (for rocprofv3), no instrumentation pass or baseline (with "
                         "--workload batched: no plain-loop baseline either)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="with --profile-run: also run one elementwise copy (torch.neg) of MIB MiB, "
                         "the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--cpu-child", nargs=3, type=int, metavar=("N2", "THREADS", "REPS"), help=argparse.SUPPRESS)
    return ap.parse_args()


def cpu_child(n2, threads, reps):
    """Runs in a subprocess with MKL_NUM_THREADS set: `reps` timed reference
    solves (wall clock around ts::zquatev only) of the same input."""
    import numpy as np
    from zquatev_amd import gen
    n = n2 // 2
    D = gen.uniform_case(n, 1000 if n2 == 512 else 1001)   # the CPU workload's seed at 2n = 512
    kind = "reference"
    try:
        from oracle import ref
        if not ref.available():
            raise ImportError("oracle/_ref not built")
        ref.zquatev(gen.uniform_case(8, 1))      # load MKL outside the timed calls
        solve = ref.zquatev
    except (ImportError, OSError):
        from oracle import zq_oracle
        kind = "port"
        solve = zq_oracle.zquatev
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        info, X, eig = solve(D)
        ts.append(time.perf_counter() - t0)
        assert info == 0
    print(json.dumps({"seconds": ts, "kind": kind, "n2": n2, "threads": threads,
                      "eig0": float(np.asarray(eig)[0])}))


def _host_description():
    """nproc / lscpu as the host reports them (the box may expose more cores
    than its share)."""
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)"):
                info["lscpu_" + k.strip().replace(" ", "_").replace("(s)", "s")] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _cpu_run(n2, threads, reps):
    env = dict(os.environ)
    env["MKL_NUM_THREADS"] = str(threads)
    env["OMP_NUM_THREADS"] = str(threads)
    env["HIP_VISIBLE_DEVICES"] = ""
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", str(n2), str(threads), str(reps)],
                         env=env, capture_output=True, text=True, timeout=1200)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-300:])
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(n2, reps, small_reps):
    """The reference compiled from its own sources (oracle/_ref, MKL), timed on
    this host: the median of `reps` solves at 2n = n2 on all the cores this
    process may use (at most the 16 of one GPU's share of the box), plus the
    2n = 512 CPU workload (median of `small_reps`) on those cores and on one."""
    import numpy as np
    from zquatev_amd import gen
    host = _host_description()
    cores = host.get("affinity_cpus") or os.cpu_count() or 1
    threads = max(1, min(16, cores))
    try:
        big = _cpu_run(n2, threads, reps)
        small = _cpu_run(512, threads, small_reps)
        small1 = _cpu_run(512, 1, small_reps)
    except (RuntimeError, subprocess.SubprocessError, ValueError) as e:
        return {"value": None, "unit": "GFLOP/s", "cores": threads, "kind": "reference",
                "sample": f"failed: {e}", "host": host}
    n = n2 // 2
    med = float(np.median(big["seconds"]))
    m512 = float(np.median(small["seconds"]))
    m512_1 = float(np.median(small1["seconds"]))
    f512 = gen.flops_vectors(256)
    return {"value": round(gen.flops_vectors(n) / med / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": big["kind"],
            "sample": (f"median of {reps} ts::zquatev calls (eigenvalues+eigenvectors) at 2n={n2} on {threads} MKL "
                       f"threads, same U(-1,1) generator, wall clock around the call only "
                       f"({', '.join(f'{t:.2f}' for t in big['seconds'])} s); GFLOP/s on (160/3) n^3"),
            "host": host,
            "runs_s": big["seconds"],
            "cpu_workload_2n512": {
                "seed": 1000, "runs": small_reps,
                "median_ms_all_cores": round(1000 * m512, 3), "gflops_all_cores": round(f512 / m512 / 1e9, 2),
                "median_ms_1_core": round(1000 * m512_1, 3), "gflops_1_core": round(f512 / m512_1 / 1e9, 2)}}


def batched_workload(count, seed, dev):
    """The batched workload: `count` Fock-like problems, 2n uniform over even
    [64, 1024] (sizes from seed, matrix i from seed + i), built on the device
    in one buffer.  Returns (n2 list, buffer, per-problem column views, pristine copy)."""
    import torch
    from zquatev_amd import gen
    n2s = [int(v) for v in gen.batched_sizes(count, seed)]
    offs, o = [], 0
    for v in n2s:
        offs.append(o)
        o += v * v
    buf = torch.zeros(o, dtype=torch.complex128, device=dev)
    views = []
    for k, v in enumerate(n2s):
        Dk = buf[offs[k]: offs[k] + v * v].view(v, v)   # Dk[j, i] = D[i, j], ld2 = 2n
        gen.fock_case_device(v // 2, seed + k, Dk)
        views.append(Dk)
    return n2s, buf, views, buf.clone()


def run_batched(ctx, stream, count, seed, steps, warmup, check=True, profile=False):
    """Batched solves/s of the batched workload against the plain loop of
    single device solves on the same stream.  profile=True (for rocprofv3):
    only the warm-up and timed batched calls, without the loop baseline, the
    instrumented batch and the checks."""
    import torch
    from zquatev_amd import gen
    dev = torch.device("cuda:0")
    n2s, buf, views, pristine = batched_workload(count, seed, dev)
    eigs = [torch.zeros(v // 2, dtype=torch.float64, device=dev) for v in n2s]
    offs = [t.data_ptr() - buf.data_ptr() for t in views]
    eptr = [e.data_ptr() for e in eigs]
    flops = sum(gen.flops_vectors(v // 2) for v in n2s)
    # The solves overwrite their inputs, so every warm-up and timed step gets
    # its own copy of the whole batch, staged in HBM before the timed region
    # (buffer 0 is buf, which the views and the checks read).
    staged = [buf] + [pristine.clone() for _ in range(max(steps + warmup, 2) - 1)]
    ptrs = [[b.data_ptr() + o for o in offs] for b in staged]
    nxt = [0]

    def refill():                        # untimed: fresh inputs in every staged copy, next step uses buf
        for b in staged:
            b.copy_(pristine)
        nxt[0] = 0

    def take():
        k = nxt[0]
        nxt[0] = (k + 1) % len(staged)
        return ptrs[k]

    def batch_step():
        rc = ctx.zquatev_batched_device(n2s, take(), n2s, eptr, "V")
        if rc != 0:
            raise RuntimeError(f"zq_zquatev_batched_device failed ({rc})")

    def loop_step():
        for v, d, e in zip(n2s, take(), eptr):
            rc = ctx.zquatev_device(v, d, v, e, "V")
            if rc != 0:
                raise RuntimeError(f"zq_zquatev_device failed ({rc})")

    def timed(fn, k, w):
        refill()
        torch.cuda.synchronize()
        for _ in range(w):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    if profile:
        t_batch = timed(batch_step, steps, warmup)
        return {"problems": count, "batched_ms": round(1000 * t_batch, 2),
                "batched_solves_per_s": round(count / t_batch, 1)}
    t_loop = timed(loop_step, 1, 1)
    res_loop = [e.clone() for e in eigs[:8]]
    t_batch = timed(batch_step, steps, warmup)
    # one instrumented batch: per-kernel-class times (HIP events on the library's stream)
    refill()
    ctx.set_instrument(True)
    batch_step()
    torch.cuda.synchronize()
    kernels = {}
    for cls in ctx.KERNEL_CLASSES:
        t, k, u = ctx.instrument_report(cls)
        v = {"launches": k, "total_ms": round(t, 3)}
        if k and u and t > 0:
            if cls == "symv":
                v["GB_per_s"] = round(u / (t / 1000.0) / 1e9, 1)
                v["frac_of_hbm_peak"] = round(u / (t / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
            else:
                v["TFLOP_per_s"] = round(u / (t / 1000.0) / 1e12, 2)
                v["frac_of_fp64_mfma_peak"] = round(u / (t / 1000.0) / 1e12 / FP64_MFMA_PEAK_TFS, 4)
        kernels[cls] = v
    ctx.set_instrument(False)
    out = {"problems": count, "n2_range": [min(n2s), max(n2s)], "seed": seed,
           "sum_n3": float(sum((v // 2) ** 3 for v in n2s)),
           "batched_solves_per_s": round(count / t_batch, 1), "batched_ms": round(1000 * t_batch, 2),
           "batched_gflops": round(flops / t_batch / 1e9, 1),
           "loop_solves_per_s": round(count / t_loop, 1), "loop_ms": round(1000 * t_loop, 2),
           "speedup_vs_loop": round(t_loop / t_batch, 2),
           "kernels": kernels}
    if check:
        # batch members equal the loop's solves (bitwise), and two members against H
        same = all(torch.equal(a, b) for a, b in zip(res_loop, eigs[:8]))
        worst = 0.0
        for k in (0, count // 2):
            n = n2s[k] // 2
            H = gen.full_hermitian_t(pristine[sum(v * v for v in n2s[:k]):][: n2s[k] ** 2].view(n2s[k], n2s[k]), n)
            X = views[k].T
            lam = torch.cat([eigs[k], eigs[k]]).to(H.dtype)
            hn = torch.linalg.matrix_norm(H, 2).item()
            worst = max(worst, (torch.linalg.norm(H @ X - X * lam[None, :]) / (n * hn)).item())
        out["check"] = {"loop_equal_first8": bool(same), "residual_max_2": worst,
                        "ok": bool(same and worst <= 1e-13)}
    return out


def run_qgemm(args):
    """The quaternion GEMM of the trailing update and back-transform on its
    own: C (M x N) -= A (M x K) B (K x N), FP64 MFMA, random operands resident
    in HBM, the library's stream; warm-up and timed calls back to back."""
    import torch
    import zquatev_amd as zq
    m, n, k = (int(v) for v in args.qgemm_shape.split(","))
    dev = torch.device("cuda:0")
    ctx = zq.Context(0)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    if args.tile64 > 0:
        ctx.set_tile_shape(-1, args.tile64)
    g = torch.Generator(device=dev).manual_seed(11)
    A = torch.randn(4 * m * k, dtype=torch.float64, device=dev, generator=g)
    B = torch.randn(4 * k * n, dtype=torch.float64, device=dev, generator=g)
    C = torch.zeros(4 * m * n, dtype=torch.float64, device=dev)

    def call():
        rc = ctx.qgemm_device("N", "N", m, n, k, -1.0, A.data_ptr(), m, B.data_ptr(), k, 1.0, C.data_ptr(), m)
        if rc != 0:
            raise RuntimeError(f"zq_qgemm failed ({rc})")

    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    tf = 32.0 * m * n * k / t / 1e12
    print(json.dumps({"metric": "quaternion GEMM TFLOP/s", "value": round(tf, 2), "unit": "TFLOP/s",
                      "ms_per_call": round(1000 * t, 3), "frac_of_fp64_mfma_peak": round(tf / FP64_MFMA_PEAK_TFS, 4),
                      "dtype": "f64", "config": {"workload": f"zq_qgemm C({m}x{n}) -= A({m}x{k}) B({k}x{n})",
                                                 "flops_per_call": 32.0 * m * n * k}}))
    return 0


def run_serving(args, summary=False):
    """The serving workload: Poisson arrivals (rate args.rate) for
    args.serve_seconds; 2n log-uniform over even [128, 4096]; 70 % of the
    requests want eigenvectors.  Requests cycle through a pool of distinct
    Fock-like matrices resident in HBM (restored from a pristine copy before
    each use).  Served by the continuous-batching server, then by the plain
    loop of one synchronous device solve per request (FIFO) on the same
    arrival times.  summary=True: a bounded run (the --summary-* sizes) whose
    figures are returned for the single workload's line instead of printed."""
    import numpy as np
    import torch
    import zquatev_amd as zq
    from zquatev_amd import gen
    dev = torch.device("cuda:0")
    T, lam, P = args.serve_seconds, args.rate, args.pool
    loop_seconds, capacity_requests = args.loop_seconds, args.capacity_requests
    if summary:
        T, P = args.summary_serve_seconds, args.summary_pool
        loop_seconds, capacity_requests = args.summary_loop_seconds, args.summary_capacity
    # arrivals and per-request choices (seed 1003)
    nmax_req = int(T * lam * 1.5) + 64
    gaps = -np.log1p(-gen.uniform(1003, np.arange(nmax_req))) / lam
    arr = np.cumsum(gaps)
    count = int(np.searchsorted(arr, T))
    arr = arr[:count]
    vec = gen.uniform(1003 + 7, np.arange(count)) < 0.7
    P = min(P, count)
    sizes = [int(v) for v in gen.serving_sizes(P, 1003)]
    pristine, work, eigs = [], [], []
    for s_, v in enumerate(sizes):
        buf = torch.zeros(v, v, dtype=torch.complex128, device=dev)
        gen.fock_case_device(v // 2, 2003 + s_, buf)
        pristine.append(buf[: v // 2].clone())
        work.append(buf)
        eigs.append(torch.zeros(v // 2, dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    cur = torch.cuda.current_stream(dev)

    def restore(s_):
        work[s_][: sizes[s_] // 2].copy_(pristine[s_])
        cur.synchronize()

    # ---- continuous batching server
    srv = zq.Server(0, lanes=args.lanes, max_batch=args.max_batch, max_batch_gb=args.max_batch_gb)
    # warm the lanes' workspaces with one burst of the largest items
    big = sorted(range(P), key=lambda q: -sizes[q])[: 2 * args.lanes]
    for q in big:
        restore(q)
        srv.submit(sizes[q], work[q].data_ptr(), sizes[q], eigs[q].data_ptr(), "V")
    srv.drain()
    srv.poll()
    last_ticket = {}
    done_t = {}
    recs = []
    t_base = srv.clock_us() + 1000.0
    tick_of = {}
    for r in range(count):
        s_ = r % P
        due = t_base + arr[r] * 1e6
        while srv.clock_us() < due:
            for c in srv.poll():
                done_t[c["ticket"]] = c
            dt = (due - srv.clock_us()) * 1e-6
            if dt > 0.0005:
                time.sleep(dt - 0.0003)
        if s_ in last_ticket:
            while last_ticket[s_] not in done_t:
                for c in srv.poll():
                    done_t[c["ticket"]] = c
        restore(s_)
        tk = srv.submit(sizes[s_], work[s_].data_ptr(), sizes[s_], eigs[s_].data_ptr(), "V" if vec[r] else "N")
        last_ticket[s_] = tk
        tick_of[tk] = (r, due)
    srv.drain()
    for c in srv.poll():
        done_t[c["ticket"]] = c
    lat_arr, lat_done, bsz, infos = [], [], [], []
    for tk, (r, due) in tick_of.items():
        c = done_t[tk]
        lat_arr.append(due * 1e-6)
        lat_done.append(c["t_done_us"] * 1e-6)
        bsz.append(c["batch_size"])
        infos.append(c["info"])
    serve = zq.latency_stats(lat_done, arrivals=lat_arr)
    serve["mean_batch"] = float(np.mean(bsz))
    serve["max_batch_seen"] = int(np.max(bsz))
    serve["bad_info"] = int(sum(1 for v in infos if v != 0))
    # accuracy: the last use of a few items (vectors) against H, and alone == in the mix
    chk_ok, worst = True, 0.0
    ctx = zq.Context(0)
    ctx.set_stream(cur.cuda_stream)
    checked = 0
    for r in range(count - 1, -1, -1):
        s_ = r % P
        if r != max(q for q in range(s_, count, P)) or not vec[r] or checked >= 6:
            continue
        n = sizes[s_] // 2
        X = work[s_].T.clone()
        e_srv = eigs[s_].clone()
        full = torch.zeros(sizes[s_], sizes[s_], dtype=torch.complex128, device=dev)
        full[:n] = pristine[s_]
        H = gen.full_hermitian_t(full, n)
        lamv = torch.cat([e_srv, e_srv]).to(H.dtype)
        hn = torch.linalg.matrix_norm(H, 2).item()
        worst = max(worst, (torch.linalg.norm(H @ X - X * lamv[None, :]) / (n * hn)).item())
        restore(s_)
        rc = ctx.zquatev_device(sizes[s_], work[s_].data_ptr(), sizes[s_], eigs[s_].data_ptr(), "V")
        torch.cuda.synchronize()
        chk_ok = chk_ok and rc == 0 and torch.equal(eigs[s_], e_srv) and torch.equal(work[s_].T, X)
        checked += 1
    serve["check"] = {"residual_max": worst, "alone_equals_mix": bool(chk_ok), "checked": checked,
                      "ok": bool(chk_ok and worst <= 1e-13)}

    # ---- saturation capacity: the same request mix, submitted back to back
    cap = None
    if capacity_requests > 0:
        ncap = capacity_requests
        cdone, csub, last_cap = {}, {}, {}
        for r in range(ncap):
            s_ = r % P
            if s_ in last_cap:
                while last_cap[s_] not in cdone:
                    for c in srv.poll():
                        cdone[c["ticket"]] = c
            restore(s_)
            tk = srv.submit(sizes[s_], work[s_].data_ptr(), sizes[s_], eigs[s_].data_ptr(),
                            "V" if vec[r % count] else "N")
            last_cap[s_] = tk
            csub[tk] = srv.clock_us()
        srv.drain()
        for c in srv.poll():
            cdone[c["ticket"]] = c
        cs = zq.latency_stats([cdone[t]["t_done_us"] * 1e-6 for t in csub],
                              arrivals=[csub[t] * 1e-6 for t in csub])
        cap = {"requests": ncap, "solves_per_s": cs["solves_per_s"],
               "mean_batch": float(np.mean([cdone[t]["batch_size"] for t in csub])),
               "bad_info": int(sum(1 for t in csub if cdone[t]["info"] != 0))}
    srv.close()

    # ---- plain loop: FIFO, one synchronous device solve per request, same arrivals
    loop_arr, loop_done = [], []
    t0 = time.perf_counter() + 0.001
    limit = min(count, int(loop_seconds * lam)) if loop_seconds > 0 else count
    for r in range(limit):
        s_ = r % P
        due = t0 + arr[r]
        while time.perf_counter() < due:
            pass
        restore(s_)
        rc = ctx.zquatev_device(sizes[s_], work[s_].data_ptr(), sizes[s_], eigs[s_].data_ptr(),
                                "V" if vec[r] else "N")
        cur.synchronize()
        if rc != 0:
            raise RuntimeError(f"zq_zquatev_device failed ({rc})")
        loop_arr.append(due)
        loop_done.append(time.perf_counter())
    loop = zq.latency_stats(loop_done, arrivals=loop_arr)
    loop["requests_offered"] = limit
    ctx.close()
    speedup = round(serve["solves_per_s"] / loop["solves_per_s"], 2)
    speedup_cap = round(cap["solves_per_s"] / loop["solves_per_s"], 2) if cap else None
    if summary:
        del work, pristine
        return {"workload": f"Poisson arrivals {lam}/s for {T} s, 2n log-uniform over even [128, 4096], "
                            "70% eigenvectors, Fock-like matrices",
                "requests": count, "pool": P, "lanes": args.lanes, "max_batch": args.max_batch,
                "p50_ms": round(serve["p50_ms"], 3), "p99_ms": round(serve["p99_ms"], 3),
                "solves_per_s": round(serve["solves_per_s"], 1), "mean_batch": round(serve["mean_batch"], 2),
                "check": serve["check"], "bad_info": serve["bad_info"],
                "plain_loop": {"requests": loop["completed"], "p50_ms": round(loop["p50_ms"], 3),
                               "p99_ms": round(loop["p99_ms"], 3), "solves_per_s": round(loop["solves_per_s"], 1)},
                "speedup_offered_rate": speedup,
                "capacity": cap, "speedup_capacity": speedup_cap,
                "note": ("at the offered rate the server's solves/s is the arrival rate (it keeps up), so "
                         "speedup_offered_rate is bounded by offered rate / plain-loop capacity; "
                         "speedup_capacity compares both saturated")}
    line = {"metric": "serving solves/s + p50/p99 latency vs plain loop", "value": round(serve["solves_per_s"], 1),
            "unit": "solves/s", "n_gpus": 1, "steps": count, "warmup": len(big),
            "ms_per_step": round(1000.0 / serve["solves_per_s"], 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Poisson arrivals {lam}/s for {T} s, 2n log-uniform over even [128, 4096], "
                                   "70% eigenvectors, Fock-like matrices",
                       "requests": count, "pool": P, "seed": 1003, "lanes": args.lanes,
                       "max_batch": args.max_batch},
            "serving": serve, "plain_loop": loop,
            # at the offered load the server's solves/s is the arrival rate (it keeps up), so
            # this ratio is bounded by offered / loop capacity; the capacity ratio compares the
            # two saturated (the loop is saturated at this load: its solves/s is its capacity)
            "speedup_solves_per_s": speedup,
            "capacity": cap,
            "speedup_capacity": speedup_cap}
    print(json.dumps(line))
    return 0


def measured_traffic(kernel, n2, nb, tsm):
    """HBM bytes per launch of `kernel` at this workload from the committed
    counter summary (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a known-size copy), or None when no summary covers it (another
    panel width or matvec strip switch)."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_2n{n2}_pmc.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if f"panel width {nb}" not in text:
        return None
    m = re.search(r"traffic_bytes_per_launch \(matvec switch %d\):\s*([0-9.eE+]+)" % tsm, text)
    return float(m.group(1)) if m else None


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(*args.cpu_child)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the solver is single-device; multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import zquatev_amd as zq
    from zquatev_amd import gen

    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "qgemm":
        return run_qgemm(args)
    if args.workload == "batched":
        ctx = zq.Context(0)
        stream = torch.cuda.current_stream(torch.device("cuda:0"))
        ctx.set_stream(stream.cuda_stream)
        if args.nb:
            ctx.set_panel_width(args.nb)
        if args.tsm >= 0 or args.tile64 > 0:
            ctx.set_tile_shape(args.tsm, args.tile64)
        r = run_batched(ctx, stream, args.batch_count, 1002, args.steps, args.warmup, check=not args.no_check,
                        profile=args.profile_run)
        if args.profile_run:
            print(json.dumps(r))
            return 0
        line = {"metric": "batched solves/s (batched workload)", "value": r["batched_solves_per_s"],
                "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": r["batched_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f64", "data": "synthetic",
                "config": {"workload": f"{args.batch_count} Fock-like problems, 2n uniform over even [64, 1024], "
                                       "eigenvalues + eigenvectors, one batched call",
                           "seed": 1002, "generator": "splitmix64 (device)"},
                "batched": r}
        print(json.dumps(line))
        return 0

    n2 = args.n2
    n = n2 // 2
    dev = torch.device("cuda:0")
    if args.profile_run and n2 > 8192:
        # large profile runs: the same matrix generated on the device, lower
        # triangles of A and B only (what the solver reads; the host copy and
        # the checks below are not run here)
        D = None
        Dt = torch.zeros((n2, n2), dtype=torch.complex128, device=dev)
        gen.uniform_case_device(n, args.seed, Dt, diag_shift=args.diag_shift)
    else:
        D = gen.uniform_case(n, args.seed, diag_shift=args.diag_shift)
        # row-major D^T == column-major D
        Dt = torch.from_numpy(np.ascontiguousarray(D.T)).to(dev)
    pristine = Dt[:n, :].clone()          # left half (first n columns of D)
    eig = torch.zeros(n, dtype=torch.float64, device=dev)
    ctx = zq.Context(0, panel_width=args.nb)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    tune = None
    if args.nb:
        nb = args.nb
    elif args.profile_run:
        nb = ctx.panel_width(n)
    else:
        chosen, tms = ctx.autotune([n])
        nb = chosen[0]
        tune = {str(w): round(float(t), 3) for w, t in zip((16, 32, 64, 128), tms[0]) if t >= 0}
        print(f"autotune n={n}: reduction ms {tune} -> panel width {nb}", file=sys.stderr, flush=True)
    if args.tsm >= 0 or args.tile64 > 0:
        ctx.set_tile_shape(args.tsm, args.tile64)
    tsm, t64 = ctx.tile_shape(n)
    print(f"tile shape n={n}: matvec switch {tsm}, GEMM 64-row tiles from {t64} tiles/CU", file=sys.stderr,
          flush=True)

    def solve(buf):
        rc = ctx.zquatev_device(n2, buf.data_ptr(), n2, eig.data_ptr(), "V")
        if rc != 0:
            raise RuntimeError(f"zq_zquatev_device failed: {zq.error_string(rc)} ({rc})")

    def step():                          # untimed uses: restore the input, then solve in place
        Dt[:n, :].copy_(pristine)
        solve(Dt)

    # The solve overwrites its input, so every warm-up and timed step gets its
    # own copy of the input, staged in HBM before the timed region: the timed
    # steps are the solves alone, back to back.
    staged = [Dt] + [Dt.clone() for _ in range(args.warmup + args.steps - 1)]
    for k in range(args.warmup):
        solve(staged[k])
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for k in range(args.steps):
        solve(staged[args.warmup + k])
    ev1.record(stream)
    torch.cuda.synchronize()
    if len(staged) > 1:                  # the checks below read the last timed step's result from Dt
        Dt.copy_(staged[-1])
    del staged[1:]
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    ev_ms = ev0.elapsed_time(ev1) / args.steps
    phases = ctx.last_timings()
    flops = gen.flops_vectors(n)
    value = flops / (ms_per_step / 1000.0) / 1e9

    if args.profile_run:
        out = {"ms_per_step": ms_per_step, "gflops": value, "phases_ms": phases, "panel_width": nb,
               "matvec_switch": tsm, "gemm64_tiles_per_cu": t64}
        if args.calib_copy > 0:
            nel = args.calib_copy * (1 << 20) // 8
            src = torch.ones(nel, dtype=torch.float64, device=dev)
            dst = torch.empty_like(src)
            torch.cuda.synchronize()
            torch.neg(src, out=dst)
            torch.cuda.synchronize()
            out["calib_copy_bytes_each_way"] = nel * 8
        print(json.dumps(out))
        return 0

    check = None
    if not args.no_check:
        # size-independent checks on the device result: trace identity and the
        # residual / orthonormality of 64 eigenpairs against H rebuilt from the input
        A = pristine[:, :n].T          # (n x n) = D[:n, :n]
        B = pristine[:, n:].T          # D[n:, :n]
        H = torch.cat([torch.cat([A, -B.conj()], 1), torch.cat([B, A.conj()], 1)], 0)
        X = Dt.T                       # column-major view of the result
        cols = torch.arange(0, n2, max(1, n2 // 64), device=dev)
        Xs = X[:, cols].contiguous()
        lam = torch.cat([eig, eig])[cols]
        g = torch.Generator(device=dev).manual_seed(7)
        v = torch.randn(n2, dtype=torch.float64, device=dev, generator=g).to(H.dtype)
        for _ in range(30):            # power iteration: a lower bound of ||H||_2
            v = H @ v
            v = v / torch.linalg.vector_norm(v)
        hn = torch.linalg.vector_norm(H @ v).item()
        res = (torch.linalg.norm(H @ Xs - Xs * lam[None, :]) / (n * hn)).item()
        orth = (Xs.conj().T @ Xs - torch.eye(len(cols), dtype=Xs.dtype, device=dev)).abs().max().item()
        trace = abs(2.0 * eig.sum().item() - torch.diagonal(H).real.sum().item()) / (n2 * hn)
        check = {"residual_64": res, "orth_64": orth, "trace_rel": trace,
                 "ok": bool(res <= 1e-13 and orth <= 1e-13 and trace <= 1e-13)}
        del H

    # eigenvalues only (jobz = 'N'): the same input, timed on its own count (64/3) n^3
    vals = None
    if not args.no_values_only:
        eig_v = eig.clone()
        eig_n = torch.zeros_like(eig)
        tv = []
        for r in range(3):
            Dt[:n, :].copy_(pristine)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            rc = ctx.zquatev_device(n2, Dt.data_ptr(), n2, eig_n.data_ptr(), "N")
            torch.cuda.synchronize()
            if rc != 0:
                raise RuntimeError(f"zq_zquatev_device(jobz='N') failed ({rc})")
            tv.append(time.perf_counter() - t1)
        tmin = min(tv[1:])
        vals = {"ms": round(1000.0 * tmin, 3), "gflops_on_64_3_n3": round(gen.flops_values(n) / tmin / 1e9, 1),
                "phases_ms": {k: round(v, 3) for k, v in ctx.last_timings().items()},
                "max_abs_diff_vs_vectors": float((eig_n - eig_v).abs().max().item())}

    # the drop-in host-memory entry point (ts::zquatev semantics) on the same
    # input: host -> device copy, solve, device -> host copy (PCIe-inclusive)
    host = None
    if not args.no_host_entry:
        th = []
        for r in range(2):
            Dh = np.asfortranarray(D.copy())
            t1 = time.perf_counter()
            info, eig_h = ctx.zquatev(Dh)
            th.append(time.perf_counter() - t1)
            if info != 0:
                raise RuntimeError(f"zq_zquatev_ctx failed ({info})")
        host = {"ms": round(1000.0 * th[-1], 3), "gflops_pcie_inclusive": round(flops / th[-1] / 1e9, 1),
                "bytes_h2d": int(D.nbytes) // 2, "bytes_d2h": int(D.nbytes),
                "max_abs_diff_vs_device": float(np.abs(eig_h - eig.cpu().numpy()).max())}
        del Dh

    # instrumented pass: per-kernel-class HIP event times on the library's stream
    ctx.set_instrument(True)
    step()
    torch.cuda.synchronize()
    kernels = {}
    for cls in ctx.KERNEL_CLASSES:
        t, k, u = ctx.instrument_report(cls)
        kernels[cls] = {"launches": k, "total_ms": round(t, 3), "avg_ms": (t / k if k else None), "units": u}
    ctx.set_instrument(False)
    dom = max(("symv", "trail", "backtr"), key=lambda c: kernels[c]["total_ms"])
    kd = kernels[dom]
    if dom == "symv":
        ach = kd["units"] / (kd["total_ms"] / 1000.0) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic("k_symv", n2, nb, tsm),
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                "algorithmic_bytes_per_launch": kd["units"] / kd["launches"],
                "kernel": "k_symv (panel Hermitian matvec)"}
    else:
        ach = kd["units"] / (kd["total_ms"] / 1000.0) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_MFMA_PEAK_TFS, 4), "traffic": None,
                "kernel": "k_qgemm (%s)" % ("trailing rank-2nb update" if dom == "trail" else "back-transform")}
    roof["avg_launch_ms"] = kd["avg_ms"]
    roof["launches"] = kd["launches"]
    if dom == "symv":
        # the streaming read rate this device reaches (torch's sum over 4 GiB
        # resident in HBM, best of 5), beside the 8 TB/s datasheet peak
        xs = torch.ones(1 << 29, dtype=torch.float64, device=dev)
        xs.sum()
        best = None
        for _ in range(5):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            xs.sum()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            best = dt if best is None else min(best, dt)
        rd = xs.numel() * 8 / best / 1e9
        roof["measured_stream_read_GBps"] = round(rd, 1)
        roof["frac_of_measured_stream_read"] = round(roof["achieved"] / rd, 4)
        del xs
    for c in kernels:
        v = kernels[c]
        if v["launches"] and v["units"]:
            if c == "symv":
                v["GB_per_s"] = round(v["units"] / (v["total_ms"] / 1000.0) / 1e9, 1)
            else:
                v["TFLOP_per_s"] = round(v["units"] / (v["total_ms"] / 1000.0) / 1e12, 2)

    batched = None
    if not args.no_batched:
        del Dt, pristine
        torch.cuda.empty_cache()
        batched = run_batched(ctx, stream, args.batch_count, 1002, 2, 1, check=not args.no_check)

    serving = None
    if not args.no_serving:
        torch.cuda.empty_cache()
        try:
            serving = run_serving(args, summary=True)
        except Exception as e:   # the single-solve line stands on its own; record what failed
            serving = {"error": f"{type(e).__name__}: {e}"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_n2, args.cpu_reps, args.cpu_small_reps)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single solve, 2n={n2} complex128, A Hermitian / B antisymmetric U(-1,1), "
                               "eigenvalues + Kramers-paired eigenvectors",
                   "n2": n2, "panel_width": nb, "panel_width_autotune_ms": tune,
                   "tile_shape": {"matvec_switch": tsm, "gemm64_tiles_per_cu": t64}, "seed": args.seed,
                   "generator": "splitmix64",
                   "flop_count": "(160/3) n^3, n = 2n/2"},
        "event_ms_per_step": round(ev_ms, 3),
        "phases_ms": {k: round(v, 3) for k, v in phases.items()},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "check": check,
        "batched": batched,
        "serving": serving,
        "values_only": vals,
        "host_entry": host,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
