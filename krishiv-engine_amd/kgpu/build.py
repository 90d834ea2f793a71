"""Compiles libkgpu.so (the HIP sources under ../csrc) for gfx950, in-tree.

    python -m kgpu.build            (from krishiv-engine_amd/)

hipcc cross-compiles for gfx950 without a GPU.  Objects are rebuilt every
call (the build check compiles every time); each source compiles in its own
process.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(HERE, "libkgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-munsafe-fp-atomics",
    # no packed-fp32 instructions (some forms have misbehaved on these boxes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


# The host code under AddressSanitizer and UndefinedBehaviorSanitizer (each
# -fsanitize= right after -Xarch_host: the device code is built as always).
ASAN_DIR = os.path.join(HERE, "_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libkgpu.so")
ASAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
              "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer",
              "-Xarch_host", "-shared-libasan"]


def _compile(src: str, asan: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + (".asan.o" if asan else ".o"))
    cmd = [HIPCC, *FLAGS, *(ASAN_FLAGS if asan else []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    noise = [l for l in r.stderr.splitlines()
             if l.strip() and "packed-fp32-ops" not in l]
    if noise:
        sys.stderr.write("\n".join(noise[-40:]) + "\n")
    return obj


def build(verbose: bool = True, asan: bool = False) -> str:
    """Builds libkgpu.so, or with asan the host-sanitized copy in _asan/
    (loaded instead when KGPU_LIB names it, with the sanitizer runtime
    preloaded: tests/test_host_sanitized.py)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, asan), srcs))
    out = ASAN_LIB if asan else LIB
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
           *(["-fsanitize=address,undefined", "-shared-libasan"] if asan else []),
           "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(srcs)} sources")
    return out


def asan_runtime() -> str:
    """The compiler's shared AddressSanitizer runtime (to preload)."""
    r = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    return r.stdout.strip()


if __name__ == "__main__":
    build(asan="--asan" in sys.argv)
