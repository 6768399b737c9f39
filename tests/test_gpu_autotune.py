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
