"""Failure reporting, forced from the host (every kernel still exits normally).

The reference returns zhbev's info (zquatev.cc:84, :99): 0, or > 0 when the
QL iteration does not converge.  This build returns the same for its own
tridiagonal solver (leaf QL, secular-equation roots, multisection), and
ZQ_ERR_NONFINITE for an input holding a NaN or an infinity (the reference
does not check).  Non-convergence is forced through the context's iteration
caps; each case is followed by a normal solve on the same context, which
must still be correct."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import zquatev_amd as zq  # noqa: E402
from zquatev_amd import gen  # noqa: E402
from tests import invariants as inv  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(D):
    import torch
    return torch.from_numpy(np.ascontiguousarray(D.T)).cuda()   # row-major D^T == column-major D


def _clean_solve_ok(ctx, n, seed):
    D = gen.uniform_case(n, seed)
    X = np.array(D, order="F", copy=True)
    info, eig = ctx.zquatev(X)
    assert info == 0
    inv.assert_solution(inv.full_h(D, n), X, eig)


@pytest.mark.parametrize("jobz,where", [("V", "A"), ("N", "A"), ("V", "B"), ("N", "B")])
def test_nonfinite_input_device(gpu_ctx, jobz, where):
    import torch
    n = 70
    _say(f"case: non-finite input jobz={jobz} in {where}")
    D = gen.uniform_case(n, 6100)
    if where == "A":
        D[5, 2] = np.nan                       # lower triangle of A
    else:
        D[n + 9, 3] = complex(np.inf, 0.0)     # lower triangle of B
    Dt = _dev(D)
    eig = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gpu_ctx.zquatev_device(2 * n, Dt.data_ptr(), 2 * n, eig.data_ptr(), jobz)
    assert rc == 0                             # launches went through
    assert gpu_ctx.last_info() == zq.ERR_NONFINITE
    assert torch.isnan(eig).all().item(), "eigenvalues of a non-finite input are NaN"
    _clean_solve_ok(gpu_ctx, n, 6101)


def test_nonfinite_input_host_entry(gpu_ctx):
    n = 33
    _say("case: non-finite input, host entry")
    D = gen.uniform_case(n, 6200)
    D[n + 4, 1] = complex(0.0, np.inf)
    X = np.array(D, order="F", copy=True)
    info, eig = gpu_ctx.zquatev(X)
    assert info == zq.ERR_NONFINITE
    assert np.isnan(eig).all()
    _clean_solve_ok(gpu_ctx, n, 6201)
    # the drop-in symbol returns the same code
    info, _ = zq.zquatev(np.array(D, order="F", copy=True))
    assert info == zq.ERR_NONFINITE


def test_nonfinite_member_of_a_batch(gpu_ctx):
    """A NaN in one problem marks that problem only; the others equal their solo solves."""
    import torch
    sizes = [48, 100, 31, 100, 7]
    _say(f"case: batch with a NaN member sizes={sizes}")
    Ds = [gen.uniform_case(n, 6300 + k) for k, n in enumerate(sizes)]
    Ds[1][60, 10] = np.nan
    dev = [_dev(D) for D in Ds]
    eig = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in sizes]
    torch.cuda.synchronize()
    rc, info = gpu_ctx.zquatev_batched_device([2 * n for n in sizes], [t.data_ptr() for t in dev],
                                              [2 * n for n in sizes], [e.data_ptr() for e in eig], "V",
                                              want_info=True)
    assert rc == zq.ERR_NONFINITE
    assert info == [0, zq.ERR_NONFINITE, 0, 0, 0]
    assert torch.isnan(eig[1]).all().item()
    for k, (n, D) in enumerate(zip(sizes, Ds)):
        if k == 1:
            continue
        t = _dev(D)
        e1 = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert gpu_ctx.zquatev_device(2 * n, t.data_ptr(), 2 * n, e1.data_ptr(), "V") == 0
        gpu_ctx.synchronize()
        assert torch.equal(e1, eig[k]) and torch.equal(t, dev[k]), f"member {k}"
    # asynchronous batch: the aggregate through zq_last_info
    dev = [_dev(D) for D in Ds]
    torch.cuda.synchronize()
    rc = gpu_ctx.zquatev_batched_device([2 * n for n in sizes], [t.data_ptr() for t in dev],
                                        [2 * n for n in sizes], [e.data_ptr() for e in eig], "N")
    assert rc == 0 and gpu_ctx.last_info() == zq.ERR_NONFINITE


@pytest.mark.parametrize("caps,jobz", [((1, 0, 0), "V"), ((0, 1, 0), "V"), ((0, 0, 2), "N")])
def test_forced_nonconvergence(caps, jobz):
    """Iteration caps too small to converge: the solve reports info > 0 (as
    zhbev reports a failed QL iteration) and returns normally."""
    import torch
    n = 200
    _say(f"case: forced non-convergence caps={caps} jobz={jobz}")
    D = gen.uniform_case(n, 6400)
    with zq.Context(0) as ctx:
        ctx.set_iteration_caps(*caps)
        Dt = _dev(D)
        eig = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert ctx.zquatev_device(2 * n, Dt.data_ptr(), 2 * n, eig.data_ptr(), jobz) == 0
        info = ctx.last_info()
        assert 0 < info <= n, info
        assert torch.isfinite(eig).all().item()
        X = np.array(D, order="F", copy=True)
        hinfo, _ = ctx.zquatev(X)
        if jobz == "V":
            assert hinfo > 0                   # the host entry returns it directly
        ctx.set_iteration_caps(0, 0, 0)
        _clean_solve_ok(ctx, n, 6401)


def test_solve_torch_orders_on_current_stream(gpu_ctx):
    """solve_torch runs on torch's current stream: the producer of D queued
    just before it is awaited, and eig / D are read without any explicit
    synchronisation."""
    import torch
    n = 300
    _say("case: solve_torch stream order")
    D = gen.uniform_case(n, 6500)
    src = _dev(D)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Dt = torch.empty_like(src)
        Dt.copy_(src)                           # queued on s, not yet synchronised
        rc, eig = zq.solve_torch(gpu_ctx, Dt.T, "V")
        got_eig = eig.cpu()                     # a copy on s: waits for the solve
        got_X = Dt.T.cpu()
    assert rc == 0
    assert gpu_ctx.last_info() == 0
    X = np.asfortranarray(got_X.numpy())
    inv.assert_solution(inv.full_h(D, n), X, got_eig.numpy())


def test_workspace_limit_splits_a_batch(gpu_ctx):
    """With a workspace cap below the batch's need, the batch runs as
    consecutive runs; every result still equals the uncapped batch's bit for
    bit, and a single problem over the cap returns ZQ_ERR_NOMEM."""
    import torch
    sizes = [300, 40, 260, 90, 300, 12, 150, 64]
    _say(f"case: workspace-capped batch sizes={sizes}")
    Ds = [gen.uniform_case(n, 6600 + k) for k, n in enumerate(sizes)]
    n2 = [2 * n for n in sizes]

    def run(ctx):
        dev = [_dev(D) for D in Ds]
        eig = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in sizes]
        torch.cuda.synchronize()
        rc, info = ctx.zquatev_batched_device(n2, [t.data_ptr() for t in dev], n2, [e.data_ptr() for e in eig],
                                              "V", want_info=True)
        assert rc == 0 and info == [0] * len(sizes)
        return dev, eig

    with zq.Context(0) as ref_ctx:
        dev0, eig0 = run(ref_ctx)
    whole, _, _ = zq.batch_workspace(n2, "V")
    limit = whole // 2
    need, runs, _ = zq.batch_workspace(n2, "V", budget=limit)
    assert runs >= 2 and need <= limit
    with zq.Context(0) as ctx:
        ctx.set_workspace_limit(limit)
        dev1, eig1 = run(ctx)
        for k in range(len(sizes)):
            assert torch.equal(eig0[k], eig1[k]) and torch.equal(dev0[k], dev1[k]), f"member {k}"
        ctx.set_workspace_limit(1 << 20)       # far below one 2n = 600 solve
        t = _dev(gen.uniform_case(300, 6700))
        e = torch.zeros(300, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert ctx.zquatev_device(600, t.data_ptr(), 600, e.data_ptr(), "V") == zq.ERR_NOMEM
        ctx.set_workspace_limit(0)
        _clean_solve_ok(ctx, 300, 6701)
