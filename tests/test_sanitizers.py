"""The host-side C and C++ under AddressSanitizer and UBSan (CPU only): the
oracle's C source runs the oracle test module through its Python wrapper with
the instrumented library preloaded, and the library's table builder runs its
invariant checks.  Built with gcc/g++ from tests/sanitize/Makefile."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "sanitize")
BUILD = os.path.join(SAN, "_build")


def _build():
    subprocess.run(["make", "-s", "-C", SAN], check=True, capture_output=True, text=True)


def _runtime(name):
    out = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def built():
    if _runtime("libasan.so") is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    _build()
    return BUILD


def test_table_builder_under_asan_ubsan(built):
    r = subprocess.run([os.path.join(built, "tables_check")], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and "tables_check ok" in r.stdout


def test_oracle_suite_under_asan_ubsan(built):
    """test_oracle.py against the instrumented oracle (the heavier Monte Carlo
    cases are deselected to keep the sanitized run to about a minute)."""
    env = dict(os.environ,
               PDSPY_ORACLE_LIB=os.path.join(built, "liboracle_asan.so"),
               LD_PRELOAD=" ".join(p for p in (_runtime("libasan.so"), _runtime("libubsan.so"))
                                   if p),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="2")
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
           os.path.join(HERE, "test_oracle.py"),
           "-k", "not epoch_lag_bias and not mrw_temperatures and not thread_count"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900, env=env, cwd=ROOT)
    print(r.stdout[-3000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
