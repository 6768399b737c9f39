"""Batched solves of uneven sizes (zq_zquatev_batched_device / zq_zquatev_batched).

The reference has no batch entry point (one matrix per ts::zquatev call,
zquatev.cc:42-72); a batch must give every problem exactly what its own call
gives.  Every batch member is compared bitwise with its solo device solve on
the same context, and checked against the invariants of the full 2n x 2n
problem (residual, orthonormality, eigenvalues of zheev)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from zquatev_amd import gen  # noqa: E402
from tests import invariants as inv  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(D):
    import torch
    return torch.from_numpy(np.ascontiguousarray(D.T)).cuda()   # row-major D^T == column-major D


def _solo(ctx, D, jobz="V"):
    import torch
    n2 = D.shape[1]
    Dt = _dev(D)
    eig = torch.zeros(max(n2 // 2, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.zquatev_device(n2, Dt.data_ptr(), D.shape[0], eig.data_ptr(), jobz)
    ctx.synchronize()
    assert rc == 0
    return Dt.T.cpu().numpy(), eig[: n2 // 2].cpu().numpy()


def _batch(ctx, Ds, jobz="V", want_info=True):
    import torch
    Dts = [_dev(D) for D in Ds]
    eigs = [torch.zeros(max(D.shape[1] // 2, 1), dtype=torch.float64, device="cuda") for D in Ds]
    torch.cuda.synchronize()
    out = ctx.zquatev_batched_device([D.shape[1] for D in Ds], [t.data_ptr() for t in Dts],
                                     [D.shape[0] for D in Ds], [e.data_ptr() for e in eigs], jobz, want_info)
    ctx.synchronize()
    rc, info = out if want_info else (out, None)
    return rc, info, [t.T.cpu().numpy() for t in Dts], [e[: D.shape[1] // 2].cpu().numpy() for e, D in zip(eigs, Ds)]


# uneven sizes on both sides of the back-transform block switch (n = 448 |
# 449: two groups), the smallest cases (n = 1, 2) and sizes about the 16-row
# D&C leaves and the 64-row reduction tail
SIZES = [300, 1, 65, 257, 2, 64, 130, 33, 256, 3, 200, 7, 449, 17, 448]


def test_batch_equals_solo_bitwise(gpu_ctx):
    _say(f"case: batch of {len(SIZES)} uniform problems, sizes {SIZES}")
    Ds = [gen.uniform_case(n, 5000 + k) for k, n in enumerate(SIZES)]
    rc, info, Xs, eigs = _batch(gpu_ctx, Ds)
    assert rc == 0 and info == [0] * len(SIZES)
    for k, (n, D) in enumerate(zip(SIZES, Ds)):
        Xs1, e1 = _solo(gpu_ctx, D)
        assert np.array_equal(eigs[k], e1), f"problem {k} (n={n}): eigenvalues differ from the solo solve"
        assert np.array_equal(Xs[k], Xs1), f"problem {k} (n={n}): eigenvectors differ from the solo solve"
        H = gen.full_hermitian(D, n)
        inv.assert_solution(H, Xs[k], eigs[k])


def test_batch_fock_workload(gpu_ctx):
    """A small draw of the batched workload: Fock-like matrices, 2n uniform over even [64, 1024]."""
    n2s = [int(v) for v in gen.batched_sizes(10, 1002)]
    _say(f"case: fock batch 2n={n2s}")
    Ds = [gen.fock_case(v // 2, 1002 + k) for k, v in enumerate(n2s)]
    rc, info, Xs, eigs = _batch(gpu_ctx, Ds)
    assert rc == 0 and info == [0] * len(Ds)
    for k, (v, D) in enumerate(zip(n2s, Ds)):
        n = v // 2
        H = gen.full_hermitian(D, n)
        inv.assert_solution(H, Xs[k], eigs[k])
    # one member against its solo solve
    Xs1, e1 = _solo(gpu_ctx, Ds[3])
    assert np.array_equal(eigs[3], e1) and np.array_equal(Xs[3], Xs1)


def test_batch_wide_sizes(gpu_ctx):
    """A member large enough for the wide column kernel (trailing sizes above
    1024) next to small ones: each member bitwise equal to its solo solve."""
    sizes = [1100, 40, 700]
    _say(f"case: batch with wide column kernel, sizes {sizes}")
    Ds = [gen.uniform_case(n, 6600 + n) for n in sizes]
    rc, info, Xs, eigs = _batch(gpu_ctx, Ds)
    assert rc == 0 and info == [0] * len(sizes)
    for k, (n, D) in enumerate(zip(sizes, Ds)):
        Xs1, e1 = _solo(gpu_ctx, D)
        assert np.array_equal(eigs[k], e1), f"problem {k} (n={n}): eigenvalues differ from the solo solve"
        assert np.array_equal(Xs[k], Xs1), f"problem {k} (n={n}): eigenvectors differ from the solo solve"
        if n <= 700:
            inv.assert_solution(gen.full_hermitian(D, n), Xs[k], eigs[k])
        else:
            H = gen.full_hermitian(D, n)
            assert max(inv.eig_error(H, eigs[k])) <= inv.EIG_TOL


def test_batch_values_only(gpu_ctx):
    sizes = [90, 1, 300, 17, 256]
    _say(f"case: values-only batch {sizes}")
    Ds = [gen.uniform_case(n, 6000 + n) for n in sizes]
    rc, info, Xs, eigs = _batch(gpu_ctx, Ds, jobz="N")
    assert rc == 0 and info == [0] * len(sizes)
    for k, (n, D) in enumerate(zip(sizes, Ds)):
        _, e1 = _solo(gpu_ctx, D, jobz="N")
        assert np.array_equal(eigs[k], e1)
        assert np.array_equal(Xs[k][:, :n], D[:, :n]), "jobz='N' must leave the input untouched"
        H = gen.full_hermitian(D, n)
        assert max(inv.eig_error(H, eigs[k])) <= inv.EIG_TOL


def test_batch_host_entry(gpu_ctx):
    sizes = [40, 12, 70]
    _say(f"case: host batch {sizes}")
    Ds = [gen.uniform_case(n, 7100 + n) for n in sizes]
    # a padded leading dimension on one problem
    Dp = np.zeros((2 * sizes[1] + 6, 2 * sizes[1]), dtype=np.complex128, order="F")
    Dp[: 2 * sizes[1], :] = Ds[1]
    work = [np.asfortranarray(Ds[0].copy()), Dp, np.asfortranarray(Ds[2].copy())]
    rc, info, eigs = gpu_ctx.zquatev_batched(work)
    assert rc == 0 and info == [0, 0, 0]
    for k, n in enumerate(sizes):
        X = work[k][: 2 * n, :]
        Xs1, e1 = _solo(gpu_ctx, Ds[k])
        assert np.array_equal(eigs[k], e1)
        assert np.array_equal(X, Xs1)


def test_batch_host_entry_chunked(gpu_ctx):
    """Host batch whose copies fill several pinned chunks and split problems
    between them; padding rows below a problem stay untouched, and an empty
    member (2n = 0) moves nothing."""
    sizes = [900, 7, 0, 260, 1000]
    _say(f"case: host batch chunked {sizes}")
    Ds = [gen.uniform_case(n, 7200 + n) if n else None for n in sizes]
    work = []
    for k, (n, D) in enumerate(zip(sizes, Ds)):
        pad = 3 if k % 2 else 0
        X = np.zeros((max(2 * n + pad, 1), 2 * n), dtype=np.complex128, order="F")
        if n:
            X[: 2 * n, :] = D
        X[2 * n:, :] = 4.0 + 1.0j
        work.append(X)
    rc, info, eigs = gpu_ctx.zquatev_batched(work)
    assert rc == 0 and info == [0] * len(sizes)
    for k, n in enumerate(sizes):
        assert np.all(work[k][2 * n:, :] == 4.0 + 1.0j)
        if n == 0:
            assert eigs[k].size == 0
            continue
        Xs1, e1 = _solo(gpu_ctx, Ds[k])
        assert np.array_equal(eigs[k], e1)
        assert np.array_equal(work[k][: 2 * n, :], Xs1)


def test_batch_twice_then_smaller(gpu_ctx):
    """Consecutive batches on one context reuse its workspace and staging."""
    a = [gen.uniform_case(n, 7300 + n) for n in (150, 20, 80)]
    b = [gen.uniform_case(n, 7400 + n) for n in (30, 10)]
    _say("case: batch 150/20/80 then 30/10")
    rc1, _, X1, e1 = _batch(gpu_ctx, a, want_info=False)
    rc2, _, X2, e2 = _batch(gpu_ctx, b, want_info=False)
    assert rc1 == 0 and rc2 == 0
    for Ds, Xs, es in ((a, X1, e1), (b, X2, e2)):
        for D, X, e in zip(Ds, Xs, es):
            n = D.shape[1] // 2
            inv.assert_solution(gen.full_hermitian(D, n), X, e)


def test_batch_argument_errors(gpu_ctx):
    import torch
    _say("case: batch argument errors")
    D = _dev(gen.uniform_case(4, 1))
    e = torch.zeros(4, dtype=torch.float64, device="cuda")
    ok = (8, D.data_ptr(), 8, e.data_ptr())
    # count 0 is a no-op
    assert gpu_ctx.zquatev_batched_device([], [], [], []) == 0
    # problem 1 has an odd size, problem 2 a short leading dimension: nothing runs
    rc, info = gpu_ctx.zquatev_batched_device([8, 7, 8], [ok[1]] * 3, [8, 8, 6], [ok[3]] * 3, "V", True)
    assert rc == -8 and info == [0, -1, -3]
    assert gpu_ctx.zquatev_batched_device([8], [ok[1]], [8], [ok[3]], "X") == -7


def test_solve_is_deterministic(gpu_ctx):
    """No order-dependent reductions: two solves of one input agree bit for bit."""
    _say("case: determinism n=520")
    D = gen.uniform_case(520, 8520)
    X1, e1 = _solo(gpu_ctx, D)
    X2, e2 = _solo(gpu_ctx, D)
    assert np.array_equal(e1, e2) and np.array_equal(X1, X2)
