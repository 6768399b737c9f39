"""The bench.py contract on small configurations, each run as a subprocess
on the GPU: one JSON line with the keys the line promises, the roofline of
the dominant kernel, the size-independent checks, and the batched and GEMM
workloads' own lines."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, timeout=600):
    print(f"case: bench.py {' '.join(args)}", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.timeout(600)
def test_single_solve_line():
    d = _run(["--n2", "512", "--steps", "2", "--warmup", "1", "--no-serving", "--no-batched", "--no-host-entry",
              "--no-cpu-baseline"])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["dtype"] == "f64" and d["config"]["n2"] == 512 and "2n=512" in d["config"]["workload"]
    r = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in r, k
    assert 0.0 < r["frac"] < 1.0 and r["achieved"] > 0
    assert d["check"]["ok"], d["check"]
    # the eigenvalues-only summary agrees with the vectors solve
    assert d["values_only"]["max_abs_diff_vs_vectors"] <= 1e-11


@pytest.mark.timeout(600)
def test_batched_line():
    d = _run(["--workload", "batched", "--batch-count", "24", "--steps", "2", "--warmup", "1"])
    assert d["unit"] == "solves/s" and d["value"] > 0 and d["steps"] == 2
    b = d["batched"]
    assert b["problems"] == 24 and b["speedup_vs_loop"] > 0
    assert b["check"]["ok"], b["check"]


@pytest.mark.timeout(300)
def test_qgemm_line():
    d = _run(["--workload", "qgemm", "--qgemm-shape", "300,260,96", "--steps", "3", "--warmup", "1"])
    assert d["unit"] == "TFLOP/s" and d["value"] > 0
    assert 0.0 < d["frac_of_fp64_mfma_peak"] < 1.0
    assert d["config"]["flops_per_call"] == 32.0 * 300 * 260 * 96
