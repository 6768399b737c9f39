"""Real solves through the drop-in C++ symbol ts::zquatev (zquatev.h:54), as
the reference's harness calls it (test.cc:97-101): through the exported,
Itanium-mangled symbol, and from a C++ program that includes only
include/zquatev.h and links libzquatev_amd (tests/dropin/caller.cc).  Results
are compared with the reference's own outputs (golden fixtures of the
compiled reference) and with the oracle."""
import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import zquatev_amd as zq  # noqa: E402
from zquatev_amd import gen  # noqa: E402
from oracle import zq_oracle  # noqa: E402
from tests import invariants as inv  # noqa: E402
from tests.test_oracle import _testcc_input  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "reference_outputs.npz")
CALLER = os.path.join(ROOT, "tests", "dropin", "zquatev_caller")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ts_zquatev():
    f = zq.lib()._ZN2ts7zquatevEiPSt7complexIdEiPd
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return f


def _check_against(D, X, eig, n, ref_eig, ref_X=None, ref_pairs=None, ref_Xpairs=None):
    H = inv.full_h(D, n)
    inv.assert_solution(H, X, eig)
    nh = np.linalg.norm(H, 2)
    assert np.all(np.diff(eig) >= 0), "eigenvalues not ascending"
    assert np.abs(eig - ref_eig).max() <= 1e-12 * nh
    if ref_X is not None:
        assert inv.projector_distance(X, ref_X, ref_eig, n) <= 1e-10
    if ref_pairs is not None:
        for k, i in enumerate(ref_pairs):
            Y1 = X[:, [i, i + n]]
            Y2 = ref_Xpairs[:, [2 * k, 2 * k + 1]]
            assert np.abs(Y1 @ Y1.conj().T - Y2 @ Y2.conj().T).max() <= 1e-10, f"pair {i}"
    _, Xo, eo = zq_oracle.zquatev(D)
    assert np.abs(eig - eo).max() <= 1e-12 * nh
    assert inv.projector_distance(X, Xo, eo, n) <= 1e-10


def test_symbol_2n64_matches_reference_outputs():
    """2n = 64: eigenvalues and every Kramers-pair projector against the
    compiled reference's full output."""
    g = np.load(GOLD, allow_pickle=False)
    n = 32
    _say("case: ts::zquatev symbol 2n=64")
    D = np.zeros((2 * n, 2 * n), dtype=np.complex128, order="F")
    D[:, :n] = g[f"n{n}_left"]
    X = np.array(D, order="F", copy=True)
    eig = np.zeros(n)
    info = _ts_zquatev()(2 * n, X.ctypes.data, 2 * n, eig.ctypes.data)
    assert info == 0
    _check_against(D, X, eig, n, g[f"n{n}_eig"], ref_X=g[f"n{n}_X"])


def test_symbol_2n512_matches_reference_outputs():
    """2n = 512 (the CPU workload, uniform_case(256, 1000)): eigenvalues and
    the stored Kramers pairs of the compiled reference."""
    g = np.load(GOLD, allow_pickle=False)
    n = 256
    _say("case: ts::zquatev symbol 2n=512")
    D = gen.uniform_case(n, 1000)
    X = np.array(D, order="F", copy=True)
    eig = np.zeros(n)
    info = _ts_zquatev()(2 * n, X.ctypes.data, 2 * n, eig.ctypes.data)
    assert info == 0
    _check_against(D, X, eig, n, g["c0_eig"], ref_pairs=g["c0_pairs"], ref_Xpairs=g["c0_Xpairs"])


def test_symbol_padded_leading_dimension():
    """nld2 > n2: the solve reads and writes the leading n2 x n2 block only;
    rows n2..nld2 of every column are left as they were."""
    n = 40
    _say("case: ts::zquatev symbol 2n=80 ld=97")
    D = gen.uniform_case(n, 5150)
    ld = 97
    P = np.full((ld, 2 * n), 7.0 + 3.0j, dtype=np.complex128, order="F")
    P[: 2 * n, :n] = D[:, :n]
    eig = np.zeros(n)
    info = _ts_zquatev()(2 * n, P.ctypes.data, ld, eig.ctypes.data)
    assert info == 0
    assert np.all(P[2 * n:, :] == 7.0 + 3.0j)
    X = np.asfortranarray(P[: 2 * n, :])
    inv.assert_solution(inv.full_h(D, n), X, eig)


def _run_caller(D, ld2):
    n2 = D.shape[1]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([n2, ld2], dtype=np.int32).tobytes())
            f.write(np.asfortranarray(D).tobytes(order="F"))
        r = subprocess.run([CALLER, fin, fout], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        raw = open(fout, "rb").read()
    info = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    eig = np.frombuffer(raw[4: 4 + 8 * (n2 // 2)], dtype=np.float64).copy()
    X = np.frombuffer(raw[4 + 8 * (n2 // 2):], dtype=np.complex128).reshape((ld2, n2), order="F").copy()
    return info, eig, X


def test_cpp_caller_reference_harness_input():
    """The reference harness's own matrix (test.cc:59-78: srand(32), n = 200)
    solved by a C++ program through include/zquatev.h, against the
    reference's eigenvalues for that input."""
    g = np.load(GOLD, allow_pickle=False)
    n = 200
    _say("case: C++ caller test.cc input 2n=400")
    A, B = _testcc_input(n)
    D = np.zeros((2 * n, 2 * n), dtype=np.complex128, order="F")
    D[:n, :n] = A
    D[n:, :n] = B
    info, eig, X = _run_caller(D, 2 * n)
    assert info == 0
    _check_against(D, X, eig, n, g["testcc_eig"])


def test_cpp_caller_2n512_and_argument_error():
    g = np.load(GOLD, allow_pickle=False)
    n = 256
    _say("case: C++ caller 2n=512")
    D = gen.uniform_case(n, 1000)
    info, eig, X = _run_caller(D, 2 * n)
    assert info == 0
    _check_against(D, X, eig, n, g["c0_eig"], ref_pairs=g["c0_pairs"], ref_Xpairs=g["c0_Xpairs"])
    # an odd order is rejected with -1 before any device work
    info, _, _ = _run_caller(np.zeros((6, 5), dtype=np.complex128, order="F"), 6)
    assert info == -1
