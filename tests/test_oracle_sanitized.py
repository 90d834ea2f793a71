"""The oracle's own tests again, against the oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
`sanitized`): an out-of-bounds access, a use after free, a signed overflow
or a misaligned load in the C restatement aborts the run.  The tests run in
a child process with the sanitizer runtime preloaded (Python itself is not
instrumented, so leak checking is off)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    if not shutil.which("gcc"):
        return None
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if p.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_tests_pass_under_address_and_ub_sanitizers():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitized"], check=True)
    env = dict(os.environ)
    env["KGO_LIB"] = os.path.join(ORACLE, "_san", "libkgo.so")
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    tests = [os.path.join(ROOT, "tests", t) for t in
             ("test_oracle.py", "test_oracle_filter.py", "test_oracle_frame.py")]
    # -s: a sanitizer's report goes to the child's stderr, not to a captured
    # file that an abort would lose
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-p", "no:cacheprovider",
                        *tests], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:] + r.stderr[-4000:])
    assert "passed" in r.stdout
