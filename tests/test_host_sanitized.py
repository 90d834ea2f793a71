"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(kgpu/build.py, asan=True: each -fsanitize= applies to the host pass only,
the gfx950 code is built as always): the host-side tests of the C ABI (symbol
exports, error text, the Arrow C Data Interface import) run against it in a
child process with the compiler's sanitizer runtime preloaded."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "krishiv-engine_amd"))


def test_host_abi_tests_pass_with_the_host_sanitized_library():
    from kgpu import build
    rt = build.asan_runtime()
    if not os.path.isabs(rt) or not os.path.exists(rt):
        pytest.skip("the compiler's AddressSanitizer runtime is not installed")
    lib = build.ASAN_LIB
    srcs = [os.path.join(build.CSRC, f) for f in os.listdir(build.CSRC)]
    if not os.path.exists(lib) or os.path.getmtime(lib) < max(os.path.getmtime(s) for s in srcs):
        build.build(verbose=False, asan=True)
    env = dict(os.environ)
    env["KGPU_LIB"] = lib
    env["LD_PRELOAD"] = rt
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    tests = [os.path.join(ROOT, "tests", t) for t in ("test_abi.py", "test_arrow_boundary.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-p", "no:cacheprovider",
                        *tests], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:] + r.stderr[-4000:])
    assert "passed" in r.stdout
