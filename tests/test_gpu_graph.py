"""HIP graph capture of the device entry points (zq_zquatev_device,
zq_zquatev_batched_device): no counterpart in the reference, which has no
streams.  A captured solve replayed several times, with eager work and
synchronisations between the replays, must give the eager result bit for
bit every time (the solve is deterministic), and the context must refuse to
reallocate the workspace a live graph references."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import zquatev_amd as zq  # noqa: E402
from zquatev_amd import gen  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(D):
    import torch
    return torch.from_numpy(np.ascontiguousarray(D.T)).cuda()   # row-major D^T == column-major D


def test_graph_capture_replays_equal_eager():
    import torch
    _say("case: graph capture of a batch (150, 40, 90) and a single solve (120)")
    with zq.Context(0) as ctx:
        s = torch.cuda.Stream()
        ctx.set_stream(s.cuda_stream)
        sizes = [150, 40, 90]
        Dts = [_dev(gen.uniform_case(n, 8800 + n)) for n in sizes]
        pristine = [t.clone() for t in Dts]
        eigs = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in sizes]
        D1 = _dev(gen.uniform_case(120, 8920))
        p1 = D1.clone()
        e1 = torch.zeros(120, dtype=torch.float64, device="cuda")
        small = _dev(gen.uniform_case(30, 8930))
        ps = small.clone()
        es = torch.zeros(30, dtype=torch.float64, device="cuda")

        def restore():
            for t, p0 in zip(Dts, pristine):
                t.copy_(p0)
            D1.copy_(p1)

        def batch():
            return ctx.zquatev_batched_device([2 * n for n in sizes], [t.data_ptr() for t in Dts],
                                              [2 * n for n in sizes], [e.data_ptr() for e in eigs], "V", False)

        def single():
            return ctx.zquatev_device(240, D1.data_ptr(), 240, e1.data_ptr(), "V")

        # eager results (these calls also size the workspace the capture will use)
        with torch.cuda.stream(s):
            restore()
            assert batch() == 0 and single() == 0
        s.synchronize()
        ref = [t.clone() for t in Dts] + [D1.clone()]
        ref_e = [e.clone() for e in eigs] + [e1.clone()]

        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            restore()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            rc_b = batch()
            rc_s = single()
        assert rc_b == 0 and rc_s == 0
        for rep in range(3):
            with torch.cuda.stream(s):
                restore()
                small.copy_(ps)
                # eager work on the same context between replays (smaller: fits the held workspace)
                assert ctx.zquatev_device(60, small.data_ptr(), 60, es.data_ptr(), "V") == 0
            s.synchronize()
            g.replay()                  # on torch's current stream
            torch.cuda.synchronize()
            for a, b in zip([t for t in Dts] + [D1], ref):
                assert torch.equal(a, b), f"replay {rep}: eigenvectors differ from the eager solve"
            for a, b in zip(eigs + [e1], ref_e):
                assert torch.equal(a, b), f"replay {rep}: eigenvalues differ from the eager solve"

        # the graph references the workspace: a solve that needs a larger one is
        # refused until release
        nb_ = 1500
        big = torch.zeros(2 * nb_, 2 * nb_, dtype=torch.complex128, device="cuda")
        gen.uniform_case_device(nb_, 8940, big)
        eb = torch.zeros(nb_, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert ctx.zquatev_device(2 * nb_, big.data_ptr(), 2 * nb_, eb.data_ptr(), "V") == zq.ERR_GRAPH
        del g
        torch.cuda.synchronize()
        ctx.graph_release()
        assert ctx.zquatev_device(2 * nb_, big.data_ptr(), 2 * nb_, eb.data_ptr(), "V") == 0
        ctx.synchronize()
