"""Randomized parity sweep through the C ABI: random sizes (2n up to 1400),
panel widths, matvec strip switches and GEMM tile-row thresholds, uniform
and Fock-like inputs, each solve held to the bounds of tests/invariants.py
against LAPACK eigenvalues; every tenth case also a batch of five mixed
sizes under the same settings, bitwise equal to its solo solves.  Seeded, a
fixed number of cases."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from zquatev_amd import gen  # noqa: E402
from tests import invariants as inv  # noqa: E402

CASES = 120


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch_equals_solo(ctx, sizes, seed):
    import torch
    Ds = [gen.uniform_case(m, seed + k) for k, m in enumerate(sizes)]
    dev = [torch.from_numpy(np.ascontiguousarray(X.T)).cuda() for X in Ds]
    eg = [torch.zeros(m, dtype=torch.float64, device="cuda") for m in sizes]
    torch.cuda.synchronize()
    rc, infos = ctx.zquatev_batched_device([2 * m for m in sizes], [t.data_ptr() for t in dev],
                                           [2 * m for m in sizes], [e.data_ptr() for e in eg], "V", want_info=True)
    assert rc == 0 and infos == [0] * len(sizes)
    for k, (m, X) in enumerate(zip(sizes, Ds)):
        t = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        e = torch.zeros(m, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert ctx.zquatev_device(2 * m, t.data_ptr(), 2 * m, e.data_ptr(), "V") == 0
        ctx.synchronize()
        assert torch.equal(e, eg[k]) and torch.equal(t, dev[k]), f"batch member {k} (n={m})"


def test_random_settings_sweep(gpu_ctx):
    rng = np.random.default_rng(20261203)
    try:
        for case in range(CASES):
            n = int(rng.integers(1, 80)) if rng.random() < 0.5 else int(rng.integers(80, 700))
            nb = int(rng.choice([0, 1, 3, 16, 32, 64, 100]))
            tsm = int(rng.choice([-1, 0, 100, 300, 768]))
            t64 = int(rng.choice([0, 1, 64]))
            seed = int(rng.integers(1, 1 << 30))
            fock = rng.random() >= 0.6
            gpu_ctx.set_panel_width(nb)
            gpu_ctx.set_tile_shape(tsm, t64)
            _say(f"case {case}: n={n} nb={nb} matvec switch={tsm} tile64={t64} seed={seed} fock={fock}")
            D = gen.fock_case(n, seed) if fock else gen.uniform_case(n, seed)
            work = np.asfortranarray(D.copy())
            info, eig = gpu_ctx.zquatev(work)
            assert info == 0
            inv.assert_solution(inv.full_h(D, n), work, eig)
            if case % 10 == 9:
                sizes = [int(rng.integers(1, 300)) for _ in range(5)]
                _say(f"  batch {sizes}")
                _batch_equals_solo(gpu_ctx, sizes, seed)
    finally:
        gpu_ctx.set_panel_width(0)
        gpu_ctx.set_tile_shape(-1, 0)
