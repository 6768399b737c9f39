"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

zquatev_amd/csrc/Makefile's `sanitize` target builds a copy of the library
with the host side instrumented (device code unchanged) and the driver
tests/sanitize/host_checks.cc, which exercises what runs without a GPU:
argument validation of the entry points, error strings, the batch workspace
planner with its split of oversized batches on 300 random batches, and the
failure paths of context and server creation.  Any sanitizer report fails
the run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zquatev_amd", "csrc")
EXE = os.path.join(CSRC, "_build_san", "host_checks")


def test_host_code_under_asan_and_ubsan():
    # incremental: rebuilds whatever changed since the last sanitized build
    subprocess.run(["make", "-C", CSRC, "-j8", "sanitize"], check=True, capture_output=True, timeout=1200)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "halt_on_error=1:detect_leaks=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["HIP_VISIBLE_DEVICES"] = ""
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "0 failure(s)" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
