"""Host sanitizer builds (CPU): the oracle's restatements and the host side
of the C-ABI (pp_geom_build in both modes with fresh cells,
pp_geom_build_surfaces, every pp_geom getter) run once each under
AddressSanitizer + UndefinedBehaviorSanitizer, and the set-up code's host
worker pool under ThreadSanitizer, without a report (tests/sanitize/)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not available")
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(SAN, "_build")


@pytest.mark.parametrize("prog", ["oracle_check", "geom_check"])
def test_runs_clean_under_sanitizers(built, prog):
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.stdout.strip().endswith("ok")


def test_host_pool_runs_clean_under_thread_sanitizer(built):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, "pool_check")], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "WARNING: ThreadSanitizer" not in out, out[-4000:]
    assert r.stdout.strip().endswith("ok")
