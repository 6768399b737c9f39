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
