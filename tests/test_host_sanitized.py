"""The runtime's host code under AddressSanitizer, UndefinedBehaviorSanitizer and
LeakSanitizer, on the CPU: tests/host/build.py compiles map_amd/csrc's host sources with a
double of the device layer (tests/host/dev_double.cpp: host memory for HBM, launches and
copies as bookkeeping, checks of its own on buffer lifetimes against recorded launches) and a
driver (tests/host/host_driver.cpp) that goes through node lifetimes, batched and recorded
evaluation, the resident cache, spilling under an HBM cap and the error past it, serving with
continuous batching, and DAG collection. The run must end cleanly with no sanitizer report.
"""
import os
import subprocess

from tests.host import build as host_build


def test_host_runtime_under_sanitizers():
    exe = host_build.build()
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1:strict_string_checks=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, env=env,
                       cwd="/tmp")
    out = r.stdout
    assert "ERROR: AddressSanitizer" not in out and "ERROR: LeakSanitizer" not in out, out[-4000:]
    assert "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert "host driver: ok" in out
    stats = out[out.index("host driver: ok") + len("host driver: ok"):]
    for key in ('"spills": 0,', '"unspills": 0,', '"dropped_for_space": 0,', '"graphs": 0,'):
        assert key not in stats, "the driver did not exercise %s" % key
