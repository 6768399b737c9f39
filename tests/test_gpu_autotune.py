"""Panel-width autotuner (zq_autotune): the reference hard-codes nb = 20
(zquatev.cc:64); here the width is timed per size bucket on the device.  The
pick must come from {16, 32, 64, 128}, apply to every size of its bucket, yield
to an explicit override, and leave solves correct."""
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


def test_autotune_picks_and_applies():
    _say("case: autotune sizes 50, 700")
    with zq.Context(0) as ctx:
        chosen, ms = ctx.autotune([50, 700])
        # n = 50: 128 is wider than the matrix and skipped
        assert chosen[0] in (16, 32, 64) and ms[0, 3] == -1 and (ms[0, :3] > 0).all()
        assert chosen[1] in (16, 32, 64, 128) and (ms[1] > 0).all()
        assert chosen[1] == (16, 32, 64, 128)[int(np.argmin(ms[1]))]
        # the pick covers the power-of-two bucket: [32, 64) and [512, 1024)
        assert ctx.panel_width(50) == chosen[0] and ctx.panel_width(33) == chosen[0]
        assert ctx.panel_width(700) == chosen[1] and ctx.panel_width(1000) == chosen[1]
        # an explicit width overrides the tuned one; 0 restores it
        ctx.set_panel_width(24)
        assert ctx.panel_width(700) == 24
        ctx.set_panel_width(0)
        assert ctx.panel_width(700) == chosen[1]
        n = 700
        D = gen.uniform_case(n, 7007)
        work = np.asfortranarray(D.copy())
        info, eig = ctx.zquatev(work)
        assert info == 0
        inv.assert_solution(gen.full_hermitian(D, n), work, eig)


def test_autotune_argument_errors():
    _say("case: autotune argument errors")
    with zq.Context(0) as ctx:
        with pytest.raises(Exception):
            ctx.autotune([0])
        chosen, ms = ctx.autotune([])
        assert chosen == [] and ms.shape == (0, 4)


def test_autotune_tile_shapes_and_overrides():
    """zq_autotune also picks the matvec strip switch and the GEMM tile-rows
    threshold for the bucket; overrides take precedence; every switch gives a
    correct solve, and a batch equals its solo solves bit for bit under a
    non-default switch (the batch groups problems by it)."""
    import torch
    _say("case: tile-shape autotune 1500")
    with zq.Context(0) as ctx:
        assert ctx.tile_shape(1500) == (768, 16)           # defaults before tuning
        ctx.autotune([1500])
        tsm, t64 = ctx.tile_shape(1500)
        assert tsm in (512, 768, 1024, 1536) and t64 in (4, 16, 64)
        assert ctx.tile_shape(1100) == (tsm, t64)            # the same bucket [1024, 2048)
        for sw in (0, 300, 2048):
            ctx.set_tile_shape(sw, 4)
            assert ctx.tile_shape(1500) == (sw, 4)
            n = 500
            D = gen.uniform_case(n, 7100 + sw)
            work = np.asfortranarray(D.copy())
            info, eig = ctx.zquatev(work)
            assert info == 0
            inv.assert_solution(gen.full_hermitian(D, n), work, eig)
        ctx.set_tile_shape(300, 0)
        sizes = [500, 120, 380]
        Ds = [gen.uniform_case(n, 7200 + k) for k, n in enumerate(sizes)]
        dev = [torch.from_numpy(np.ascontiguousarray(D.T)).cuda() for D in Ds]
        eig = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in sizes]
        torch.cuda.synchronize()
        rc, info = ctx.zquatev_batched_device([2 * n for n in sizes], [t.data_ptr() for t in dev],
                                              [2 * n for n in sizes], [e.data_ptr() for e in eig], "V",
                                              want_info=True)
        assert rc == 0 and info == [0, 0, 0]
        for k, (n, D) in enumerate(zip(sizes, Ds)):
            t = torch.from_numpy(np.ascontiguousarray(D.T)).cuda()
            e = torch.zeros(n, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            assert ctx.zquatev_device(2 * n, t.data_ptr(), 2 * n, e.data_ptr(), "V") == 0
            ctx.synchronize()
            assert torch.equal(e, eig[k]) and torch.equal(t, dev[k]), f"member {k}"
        ctx.set_tile_shape(-1, 0)
        assert ctx.tile_shape(1500) == (tsm, t64)
        with pytest.raises(zq.ZQError):
            ctx.set_tile_shape(4096, 0)
