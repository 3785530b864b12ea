"""The oracle's C++ under AddressSanitizer and UndefinedBehaviorSanitizer:
the oracle test files run again in a child process against
oracle/liboracle_san.so (built by the oracle Makefile, host code only) with
the ASan runtime preloaded; any report aborts the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
FILES = ["test_oracle_pins.py", "test_oracle_math.py", "test_oracle_reference_tests.py", "test_oracle_io.py",
         "test_oracle_adaptive.py"]


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    p = out.stdout.strip()
    return p if out.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no AddressSanitizer runtime for g++ here")
    subprocess.run(["make", "-s", "-C", ORACLE, "liboracle_san.so"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "VR_ORACLE_LIB": os.path.join(ORACLE, "liboracle_san.so"),
    })
    # the child loads the sanitizer build, with the ASan runtime in the process
    probe = ("import ctypes, oracle; oracle.lib(); ctypes.CDLL(None).__asan_report_load4; "
             "print(oracle.LIB_PATH)")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ORACLE, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("liboracle_san.so"), r.stdout + r.stderr
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + [
        os.path.join(ROOT, "tests", f) for f in FILES]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail
