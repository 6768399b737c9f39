"""Serving loop (zq_server_*): a request answered inside a batch equals its
solo solve bit for bit ("alone == in the mix"), every ticket completes once,
and the completion records are consistent."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import zquatev_amd as zq  # noqa: E402
from zquatev_amd import gen  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_server_alone_equals_in_the_mix(gpu_ctx):
    import torch
    sizes = [40, 300, 9, 128, 257, 1, 64, 75, 200, 33, 410, 16, 90, 2, 150, 260]
    jobz = ["V" if k % 3 else "N" for k in range(len(sizes))]
    _say(f"case: server burst sizes={sizes}")
    Ds = [gen.uniform_case(n, 9100 + k) for k, n in enumerate(sizes)]
    dev = [torch.from_numpy(np.ascontiguousarray(D.T)).cuda() for D in Ds]
    eig = [torch.zeros(max(n, 1), dtype=torch.float64, device="cuda") for n in sizes]
    torch.cuda.synchronize()
    with zq.Server(0, lanes=2, max_batch=8) as srv:
        tickets = [srv.submit(2 * n, t.data_ptr(), 2 * n, e.data_ptr(), j) for n, t, e, j in zip(sizes, dev, eig, jobz)]
        srv.drain()
        recs = srv.poll()
    assert sorted(r["ticket"] for r in recs) == sorted(tickets)
    assert all(r["info"] == 0 for r in recs)
    assert all(r["t_submit_us"] <= r["t_start_us"] <= r["t_done_us"] for r in recs)
    assert max(r["batch_size"] for r in recs) > 1, "the burst was never batched"
    assert max(r["batch_size"] for r in recs) <= 8
    for k, (n, D) in enumerate(zip(sizes, Ds)):
        Dt = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
        e1 = torch.zeros(max(n, 1), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert gpu_ctx.zquatev_device(2 * n, Dt.data_ptr(), 2 * n, e1.data_ptr(), jobz[k]) == 0
        gpu_ctx.synchronize()
        assert torch.equal(eig[k][:n], e1[:n]), f"request {k} (n={n}, jobz={jobz[k]}) eigenvalues"
        if jobz[k] == "V":
            assert torch.equal(dev[k], Dt), f"request {k} (n={n}) eigenvectors"


def test_server_ticket_order_and_empty_request():
    import torch
    _say("case: server small stream")
    D = gen.uniform_case(20, 9300)
    t = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
    e = torch.zeros(20, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with zq.Server(0, lanes=1, max_batch=4) as srv:
        t0 = srv.submit(0, 0, 1, 0, "V")            # empty request completes at once
        t1 = srv.submit(40, t.data_ptr(), 40, e.data_ptr(), "V")
        srv.drain()
        recs = {r["ticket"]: r for r in srv.poll()}
        assert set(recs) == {t0, t1} and t1 > t0
        assert recs[t0]["batch_size"] == 0 and recs[t1]["batch_size"] == 1
        assert srv.poll() == []
        with pytest.raises(zq.ZQError):
            srv.submit(41, t.data_ptr(), 40, e.data_ptr(), "V")


def _size_class(n2):
    return 0 if n2 <= 384 else (1 if n2 <= 1280 else (2 if n2 <= 2048 else (3 if n2 <= 2900 else 4)))


def _fock_request(n2, seed):
    import torch
    buf = torch.zeros(n2, n2, dtype=torch.complex128, device="cuda")
    gen.fock_case_device(n2 // 2, seed, buf)
    return buf


def _check_result(buf_in, buf_out, eig, jobz):
    """Eigenvalues against LAPACK on the full H; for 'V' the residual and
    orthonormality of all eigenpairs and the exact Kramers structure."""
    import torch
    import scipy.linalg as sl
    from tests import invariants as inv
    n2 = buf_in.shape[0]
    n = n2 // 2
    H = gen.full_hermitian_t(buf_in, n)
    w = sl.eigh(H.cpu().numpy(), eigvals_only=True, driver="evd")
    nh = max(abs(w[0]), abs(w[-1]))
    e = eig.cpu().numpy()
    assert np.all(np.diff(e) >= 0)
    assert np.abs(w[0::2] - e).max() <= inv.EIG_TOL * nh and np.abs(w[1::2] - e).max() <= inv.EIG_TOL * nh
    if jobz == "V":
        X = buf_out.T                                   # column-major result
        lam = torch.cat([eig, eig]).to(X.dtype)
        res = (torch.linalg.norm(H @ X - X * lam[None, :]) / (n * nh)).item()
        orth = (X.conj().T @ X - torch.eye(n2, dtype=X.dtype, device=X.device)).abs().max().item()
        assert res <= inv.RES_TOL, res
        assert orth <= inv.ORTH_TOL, orth
        U, V = X[:n, :n], X[n:, :n]
        assert torch.equal(X[:n, n:], -V.conj()) and torch.equal(X[n:, n:], U.conj())


@pytest.mark.timeout(900)
def test_server_serving_workload_checked(gpu_ctx):
    """40 requests of the serving workload (2n log-uniform over even
    [128, 4096], Fock-like matrices, 70% eigenvectors) arriving as a Poisson
    stream of 200/s at a 3-lane server: every result is checked against
    LAPACK (and for eigenvectors by residual and orthonormality), equals its
    solo solve bit for bit, and large-class batches never hold all lanes."""
    import time
    import torch
    count, rate, lanes = 40, 200.0, 3
    sizes = [int(v) for v in gen.serving_sizes(count, 1103)]
    vec = gen.uniform(1103 + 7, np.arange(count)) < 0.7
    jobz = ["V" if v else "N" for v in vec]
    gaps = -np.log1p(-gen.uniform(1103, np.arange(count))) / rate
    _say(f"case: serving workload sizes={sizes}")
    assert any(_size_class(v) >= 2 for v in sizes)
    pristine = [_fock_request(v, 2103 + r) for r, v in enumerate(sizes)]
    work = [p.clone() for p in pristine]
    eig = [torch.zeros(v // 2, dtype=torch.float64, device="cuda") for v in sizes]
    torch.cuda.synchronize()
    with zq.Server(0, lanes=lanes, max_batch=128) as srv:
        t0 = time.perf_counter()
        due = t0 + np.cumsum(gaps)
        tickets = []
        for r in range(count):
            while time.perf_counter() < due[r]:
                pass
            tickets.append(srv.submit(sizes[r], work[r].data_ptr(), sizes[r], eig[r].data_ptr(), jobz[r]))
        srv.drain()
        recs = {c["ticket"]: c for c in srv.poll()}
    assert sorted(recs) == sorted(tickets)
    assert all(recs[t]["info"] == 0 for t in tickets)
    # batches: (lane, start) -> the class of their requests
    batches = {}
    for r, t in enumerate(tickets):
        c = recs[t]
        assert c["t_submit_us"] <= c["t_start_us"] <= c["t_done_us"]
        key = (c["lane"], c["t_start_us"])
        batches.setdefault(key, set()).add(_size_class(sizes[r]))
        batches[key].add(("end", c["t_done_us"]))
    large = []
    for (lane, ts), v in batches.items():
        classes = {x for x in v if not isinstance(x, tuple)}
        assert len(classes) == 1, "a batch mixes size classes"
        end = max(x[1] for x in v if isinstance(x, tuple))
        if next(iter(classes)) >= 2:
            large.append((ts, end, lane))
    assert large, "no batch of size classes 2-4"
    for ts, end, lane in large:
        overlapping = sum(1 for a, b, _ in large if a <= ts < b)
        assert overlapping <= lanes - 1, "large-class batches held every lane"
    _say(f"batches {len(batches)}, large {len(large)}; checking every result")
    with zq.Context(0) as solo:
        for r in range(count):
            _check_result(pristine[r], work[r], eig[r], jobz[r])
            again = pristine[r].clone()
            e1 = torch.zeros_like(eig[r])
            torch.cuda.synchronize()
            assert solo.zquatev_device(sizes[r], again.data_ptr(), sizes[r], e1.data_ptr(), jobz[r]) == 0
            assert solo.last_info() == 0
            assert torch.equal(e1, eig[r]), f"request {r} eigenvalues alone != in the mix"
            if jobz[r] == "V":
                assert torch.equal(again, work[r]), f"request {r} eigenvectors alone != in the mix"


def test_server_large_classes_leave_a_lane_for_small_requests():
    """A burst of six large requests (classes 2-4, one per batch) and three
    small ones on 3 lanes: at most two large batches run at once, and the
    small requests are served on the third lane before the large burst ends."""
    import torch
    big = [3000, 2200, 1500, 3000, 2200, 1500]
    small = [200, 300, 150]
    sizes = big + small
    _say(f"case: lanes-1 burst sizes={sizes}")
    bufs = [_fock_request(v, 2500 + k) for k, v in enumerate(sizes)]
    eig = [torch.zeros(v // 2, dtype=torch.float64, device="cuda") for v in sizes]
    torch.cuda.synchronize()
    with zq.Server(0, lanes=3, max_batch=1) as srv:
        tickets = [srv.submit(v, b.data_ptr(), v, e.data_ptr(), "V") for v, b, e in zip(sizes, bufs, eig)]
        srv.drain()
        recs = {c["ticket"]: c for c in srv.poll()}
    assert all(recs[t]["info"] == 0 and recs[t]["batch_size"] == 1 for t in tickets)
    iv = [(recs[t]["t_start_us"], recs[t]["t_done_us"], recs[t]["lane"]) for t in tickets]
    for ts, _, _ in iv[:6]:
        assert sum(1 for a, b, _ in iv[:6] if a <= ts < b) <= 2
    last_big_done = max(b for _, b, _ in iv[:6])
    for ts, te, _ in iv[6:]:
        assert te < last_big_done, "a small request waited for the whole large burst"


def test_server_memory_budget_defers_and_rejects():
    """Budget for one large request: a second one waits until the first is
    done (its start stamp after the first's completion), and both results are
    right; a request larger than the whole budget completes at once with
    ZQ_ERR_BUDGET and the server keeps serving."""
    import torch
    n2 = 1600
    need, _, _ = zq.batch_workspace([n2], "V")
    _say(f"case: server budget, one request needs {need} bytes")
    reqs = [_fock_request(n2, 2600 + k) for k in range(2)]
    pristine = [r.clone() for r in reqs]
    eig = [torch.zeros(n2 // 2, dtype=torch.float64, device="cuda") for _ in reqs]
    torch.cuda.synchronize()
    with zq.Server(0, lanes=2, max_batch=4) as srv:
        srv.set_budget(1.5 * need / 2 ** 30)
        t = [srv.submit(n2, r.data_ptr(), n2, e.data_ptr(), "V") for r, e in zip(reqs, eig)]
        srv.drain()
        recs = {c["ticket"]: c for c in srv.poll()}
        a, b = recs[t[0]], recs[t[1]]
        assert a["info"] == 0 and b["info"] == 0
        assert a["batch_size"] == 1 and b["batch_size"] == 1
        first, second = (a, b) if a["t_start_us"] <= b["t_start_us"] else (b, a)
        assert second["t_start_us"] >= first["t_done_us"], "the second request ran beside the first"
        for k in range(2):
            _check_result(pristine[k], reqs[k], eig[k], "V")
        # over the whole budget: rejected at once, later requests still served
        srv.set_budget(1e-3)
        big = _fock_request(n2, 2700)
        eb = torch.zeros(n2 // 2, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        tb = srv.submit(n2, big.data_ptr(), n2, eb.data_ptr(), "V")
        srv.drain()
        rb = {c["ticket"]: c for c in srv.poll()}[tb]
        assert rb["info"] == zq.ERR_BUDGET and rb["batch_size"] == 0
        srv.set_budget(0)
        small = _fock_request(300, 2800)
        ps = small.clone()
        es = torch.zeros(150, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ts = srv.submit(300, small.data_ptr(), 300, es.data_ptr(), "V")
        srv.drain()
        assert {c["ticket"]: c for c in srv.poll()}[ts]["info"] == 0
        _check_result(ps, small, es, "V")


def test_server_nonfinite_request_fails_alone():
    """A request whose input holds a NaN completes with ZQ_ERR_NONFINITE; the
    requests batched with it are solved as usual."""
    import torch
    sizes = [100, 100, 100, 100]
    _say("case: server NaN request in a batch")
    bufs = [_fock_request(v, 2900 + k) for k, v in enumerate(sizes)]
    bufs[2][3, 10] = float("nan")                      # column 3, row 10 of A (lower triangle)
    pristine = [b.clone() for b in bufs]
    eig = [torch.zeros(v // 2, dtype=torch.float64, device="cuda") for v in sizes]
    torch.cuda.synchronize()
    with zq.Server(0, lanes=1, max_batch=8) as srv:
        tickets = [srv.submit(v, b.data_ptr(), v, e.data_ptr(), "V") for v, b, e in zip(sizes, bufs, eig)]
        srv.drain()
        recs = {c["ticket"]: c for c in srv.poll()}
    infos = [recs[t]["info"] for t in tickets]
    assert infos == [0, 0, zq.ERR_NONFINITE, 0]
    for k in (0, 1, 3):
        _check_result(pristine[k], bufs[k], eig[k], "V")
