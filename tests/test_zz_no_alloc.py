"""The allocation-free Bloom probe (kg_bloom_probe_noalloc) through the C ABI:
the same survivors and count as kg_bloom_probe, on two calls in a row (the
second smaller, so it runs over the scratch the first left behind), for both
filter layouts; no allocation through the context's allocator during the
calls; sizes past the limit and host columns refused."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import ctx, say
import kgpu
from kgpu import Col, DeviceBuffer
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def probe_noalloc(bloom, col: Col, nrows: int, out: DeviceBuffer) -> int:
    cnt = C.c_int64(-1)
    kc = col.c()
    A.check(kgpu.lib().kg_bloom_probe_noalloc(bloom.h, C.byref(kc), nrows, C.c_void_p(out.ptr),
                                              C.byref(cnt)))
    return cnt.value


@pytest.mark.parametrize("layout", ["split-block", "spark"])
def test_noalloc_probe_equals_probe_on_two_calls(layout):
    c = ctx()
    rng = np.random.default_rng(5)
    build = rng.integers(0, 1 << 40, 300_000)
    if layout == "spark":
        bits = kgpu.RuntimeBloom.spark_bits_for(len(build), 0.03)
        b = kgpu.RuntimeBloom(c, spark_bits=bits,
                              spark_hashes=kgpu.RuntimeBloom.spark_hashes_for(len(build), bits))
    else:
        b = kgpu.RuntimeBloom(c, kgpu.RuntimeBloom.bytes_for(len(build), 0.03))
    b.insert(Col(A.T_INT64, build))
    # 40M rows (three 16M-row chunks, the last partial) then 5M rows, both
    # with NULLs and a nonzero column offset
    for n in (40_000_000, 5_000_000):
        say(f"noalloc: {layout}, {n} rows")
        off = 3
        keys = rng.integers(0, 1 << 40, n + off)
        keys[off::5] = build[rng.integers(0, len(build), len(keys[off::5]))]
        valid = rng.random(n + off) > 0.01
        kd = DeviceBuffer.from_numpy(c, keys)
        vd = DeviceBuffer.from_numpy(c, np.packbits(valid, bitorder="little"))
        col = Col(A.T_INT64, kd, vd, length=n, offset=off)
        want_rows, want_n = b.probe(col, nrows=n)
        out = DeviceBuffer(c, n * 4)
        before = c.mem_stats()
        got_n = probe_noalloc(b, col, n, out)
        after = c.mem_stats()
        assert (after["hip_mallocs"], after["cache_hits"], after["in_use"]) == \
            (before["hip_mallocs"], before["cache_hits"], before["in_use"])
        assert got_n == want_n and want_n > n // 6
        assert np.array_equal(out.to_numpy(np.int32, got_n), want_rows.to_numpy(np.int32, want_n))


def test_noalloc_probe_refusals():
    c = ctx()
    b = kgpu.RuntimeBloom(c, 4096)
    b.insert(Col(A.T_INT64, np.arange(100, dtype=np.int64)))
    kd = DeviceBuffer(c, 64)
    out = DeviceBuffer(c, 64)
    with pytest.raises(kgpu.KgError, match="KRISHIV_UNSUPPORTED_FEATURE"):
        probe_noalloc(b, Col(A.T_INT64, kd, length=1 << 31), 1 << 31, out)
    with pytest.raises(kgpu.KgError, match="KRISHIV_UNSUPPORTED_FEATURE"):
        probe_noalloc(b, Col(A.T_INT64, np.arange(8, dtype=np.int64)), 8, out)
    assert probe_noalloc(b, Col(A.T_INT64, kd, length=0), 0, out) == 0
