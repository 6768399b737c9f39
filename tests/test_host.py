"""CPU tests of the C ABI boundary: the library loads, exports every symbol
declared in include/*.h, and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import zquatev_amd as zq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in ("zq.h",):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(zq_\w+)\s*\(", src, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = zq.lib()
    declared = _declared_functions()
    assert len(declared) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", zq.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = declared - exported
    assert not missing, missing
    for name in declared:
        getattr(lib, name)
    # the C++ drop-in ts::zquatev (include/zquatev.h) with the reference's mangled name
    assert "_ZN2ts7zquatevEiPSt7complexIdEiPd" in exported


def test_python_binding_covers_abi():
    assert _declared_functions() <= set(zq._SIGNATURES)


def test_argument_validation_without_device():
    D = np.zeros((4, 4), dtype=np.complex128, order="F")
    eig = np.zeros(2)
    f = zq.lib().zq_zquatev
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert f(3, D.ctypes.data, 4, eig.ctypes.data) == -1       # odd n2
    assert f(-2, D.ctypes.data, 4, eig.ctypes.data) == -1
    assert f(4, None, 4, eig.ctypes.data) == -2
    assert f(4, D.ctypes.data, 2, eig.ctypes.data) == -3       # nld2 < n2
    assert f(4, D.ctypes.data, 4, None) == -4
    assert f(0, None, 1, None) == 0                            # empty problem
    ts = zq.lib()._ZN2ts7zquatevEiPSt7complexIdEiPd
    ts.restype = ctypes.c_int
    ts.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert ts(5, D.ctypes.data, 4, eig.ctypes.data) == -1
    info, _ = zq.zquatev(np.zeros((4, 4), dtype=np.complex128, order="F"), n2=3)
    assert info == -1


def test_batched_entry_points_validate_without_device():
    i32 = ctypes.c_int * 1
    f = zq.lib().zq_zquatev_batched_device
    assert f(None, 1, i32(4), None, i32(4), None, ord("V"), None) == zq.ERR_ARG   # null context
    g = zq.lib().zq_zquatev_batched
    assert g(None, 0, None, None, None, None, None) == zq.ERR_ARG


def test_autotune_validates_without_device():
    L = zq.lib()
    assert L.zq_autotune(None, None, 0, None, None) == zq.ERR_ARG
    assert L.zq_panel_width(None, 5) == 0


def test_graph_calls_validate_without_device():
    L = zq.lib()
    assert L.zq_graph_reserve(None, 1 << 20) == zq.ERR_ARG
    assert L.zq_graph_release(None) == zq.ERR_ARG
    assert "graph" in zq.error_string(zq.ERR_GRAPH)


def test_error_strings_and_tuning_table():
    assert zq.error_string(0) == "success"
    assert "odd" in zq.error_string(-1)
    assert "memory" in zq.error_string(zq.ERR_NOMEM)
    assert "converge" in zq.error_string(3)
    for n in (1, 32, 256, 4096, 16384):
        assert 1 <= zq.tuned_panel_width(n) <= 128
    assert zq.version() >= 100


def test_input_type_checks():
    with pytest.raises(TypeError):
        zq.zquatev(np.zeros((4, 4)))
    with pytest.raises(ValueError):
        zq.zquatev(np.zeros((4, 4), dtype=np.complex128, order="C")[:, :3])


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_latency_stats_synthetic_clock():
    # arrivals every 10 ms, each served 3 ms after arrival except the last (13 ms)
    arr = np.arange(100) * 0.010
    done = arr + 0.003
    done[-1] = arr[-1] + 0.013
    st = zq.latency_stats(done, arrivals=arr)
    assert st["completed"] == 100
    assert abs(st["p50_ms"] - 3.0) < 1e-9
    assert abs(st["max_ms"] - 13.0) < 1e-9
    assert 3.0 < st["p99_ms"] <= 13.0
    assert abs(st["solves_per_s"] - 100 / (0.99 + 0.013)) < 1e-9
    recs = [{"t_submit_us": a * 1e6, "t_done_us": d * 1e6} for a, d in zip(arr, done)]
    st2 = zq.latency_stats(recs)
    assert abs(st2["p50_ms"] - st["p50_ms"]) < 1e-6 and abs(st2["solves_per_s"] - st["solves_per_s"]) < 1e-6


def test_server_validates_without_device():
    L = zq.lib()
    h = ctypes.c_void_p()
    assert L.zq_server_create(None, 0, 1, 4, ctypes.c_double(0.0)) == zq.ERR_ARG
    assert L.zq_server_create(ctypes.byref(h), 0, 0, 4, ctypes.c_double(0.0)) == -3
    assert L.zq_server_create(ctypes.byref(h), 0, 2, 0, ctypes.c_double(0.0)) == -4
    assert L.zq_server_create(ctypes.byref(h), 0, 2, 4, ctypes.c_double(-1.0)) == -5
    t = ctypes.c_ulonglong()
    assert L.zq_server_submit(None, 4, None, 4, None, ord("V"), ctypes.byref(t)) == zq.ERR_ARG
    assert L.zq_server_poll(None, None, 0) == zq.ERR_ARG
    assert L.zq_server_destroy(None) == 0


def test_explicit_n2_must_fit_D():
    D = np.zeros((4, 4), dtype=np.complex128, order="F")
    with pytest.raises(ValueError):
        zq.zquatev(D, n2=6)            # more columns than D has: would overflow on the host
    with pytest.raises(ValueError):
        zq.zquatev(np.zeros((4, 6), dtype=np.complex128, order="F"), n2=6)   # ld2 < n2


def test_new_codes_and_setters_validate_without_device():
    L = zq.lib()
    assert "NaN" in zq.error_string(zq.ERR_NONFINITE)
    assert "budget" in zq.error_string(zq.ERR_BUDGET)
    assert L.zq_last_info(None) == zq.ERR_ARG
    assert L.zq_set_workspace_limit(None, 0) == zq.ERR_ARG
    assert L.zq_set_iteration_caps(None, 1, 1, 1) == zq.ERR_ARG
    assert L.zq_server_set_budget(None, ctypes.c_double(1.0)) == zq.ERR_ARG


def test_batch_workspace_split_plan():
    """A heavy-tailed batch (one 2n = 8192 problem and many of 2n = 2100 in one
    panel-width group) needs a slab sized for the largest problem per member;
    split to a budget, it becomes consecutive runs of the size-sorted problems,
    each within the budget, covering every problem once."""
    sizes = [2100] * 60 + [8192] + [2100] * 60
    whole, runs, _ = zq.batch_workspace(sizes, "V")
    assert runs == 1
    budget = 40 << 30
    assert whole > budget
    need, runs, run = zq.batch_workspace(sizes, "V", budget=budget)
    assert runs >= 2 and need <= budget
    assert run[60] == 0                                   # the largest problem leads the first run
    assert sorted(set(run)) == list(range(runs))
    for r in range(runs):
        members = [s for s, q in zip(sizes, run) if q == r]
        got, k, _ = zq.batch_workspace(members, "V")
        assert k == 1 and got <= budget
    # consecutive in descending size order: no run holds a larger problem than an earlier run
    firsts = [max(s for s, q in zip(sizes, run) if q == r) for r in range(runs)]
    assert firsts == sorted(firsts, reverse=True)
    # values only needs less than vectors; a problem larger than the budget forms its own run
    assert zq.batch_workspace(sizes, "N")[0] < whole
    need1, _, run1 = zq.batch_workspace([8192, 8192, 64], "V", budget=1 << 20)
    assert len(set(run1)) == 3 and need1 > (1 << 20)
    assert zq.batch_workspace([0, 64], "V")[2][0] == -1   # empty problems belong to no run
    with pytest.raises(zq.ZQError):
        zq.batch_workspace([63], "V")


def test_dropin_caller_is_built():
    exe = os.path.join(ROOT, "tests", "dropin", "zquatev_caller")
    assert os.path.exists(exe), "built by zquatev_amd/csrc/Makefile (__graft_entry__.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
