"""Shared helpers of the GPU parity tests: one device context per process and
comparison of GPU results with the oracle's."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "krishiv-engine_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import kgpu  # noqa: E402

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = kgpu.Context(0)
    return _CTX


def say(msg):
    print(msg, flush=True)


def rows_by_key(keys, kvalid, cols):
    """{key or None: tuple(values)} from parallel result columns."""
    out = {}
    for i in range(len(keys)):
        k = int(keys[i]) if kvalid[i] else None
        out[k] = tuple(c[i] for c in cols)
    return out


def same_float(a, b, rel=1e-12):
    if a is None or b is None:
        return a is None and b is None
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b or math.isclose(a, b, rel_tol=rel, abs_tol=1e-300)
