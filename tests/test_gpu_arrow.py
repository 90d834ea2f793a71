"""The Arrow C Data Interface at the C ABI on the device: operators run on
pyarrow arrays viewed in place (offsets and NULLs included), and device
batches leave as Arrow struct arrays -- copied to host memory (pyarrow
imports them) or zero-copy in HBM as an ArrowDeviceArray whose columns an
operator can take back."""
import ctypes as C

import numpy as np
import pyarrow as pa
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def test_operators_on_arrow_arrays():
    say("arrow: select_range and a hash aggregate on pyarrow arrays")
    rng = np.random.default_rng(1)
    n = 200_000
    key = pa.array(rng.integers(0, 5000, n), mask=rng.random(n) < 0.02)
    val = pa.array(rng.integers(-1000, 1000, n).astype(np.int32), mask=rng.random(n) < 0.05)
    k, v = key.slice(37), val.slice(37)  # offsets into the buffers
    out, cnt = kgpu.select_range(ctx(), kgpu.arrow_col(k), 100, 2000)
    got = out.to_numpy(np.int32, cnt)
    kn = k.to_numpy(zero_copy_only=False)
    ok = ~np.asarray(k.is_null())
    want = np.flatnonzero(ok & (np.nan_to_num(kn, nan=-1) >= 100) & (np.nan_to_num(kn, nan=-1) <= 2000))
    assert np.array_equal(got, want)
    g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), [(A.T_INT32, 0, 0)],
                           [(A.AGG_COUNT_STAR, 0), (A.AGG_SUM, 0)])
    g.sink(kgpu.arrow_col(k), [kgpu.arrow_col(v)])
    g.finalize()
    res = g.result().to_arrow(["k", "n", "s"])
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT32, 0, 0)], [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0)])
    o.update(kgo.Col(kgo.T_INT64, np.nan_to_num(kn, nan=0).astype(np.int64), ok),
             [kgo.Col(kgo.T_INT32, np.nan_to_num(v.to_numpy(zero_copy_only=False), nan=0)
                      .astype(np.int32), ~np.asarray(v.is_null()))])
    rk, rkv, rres = o.finish()
    want = {(int(a) if b else None): (int(c), None if not d else int(e))
            for a, b, c, d, e in zip(rk, rkv, rres[0][0], rres[1][1], rres[1][0])}
    gotd = {a: (b, c) for a, b, c in zip(res.column(0).to_pylist(), res.column(1).to_pylist(),
                                         res.column(2).to_pylist())}
    assert gotd == want


def test_batches_leave_as_arrow():
    say("arrow: host and device export of a batch")
    rng = np.random.default_rng(2)
    n = 100_000
    key = rng.integers(0, 1 << 40, n)
    ok = rng.random(n) > 0.05
    f = rng.standard_normal(n)
    flag = rng.random(n) > 0.5
    cols = [Col(A.T_INT64, key, ok), Col(A.T_FLOAT64, f), Col(A.T_BOOL, flag)]
    batch, offs = kgpu.repartition(ctx(), cols, [0], 16)
    pids = kgo.route([kgo.Col(kgo.T_INT64, key, ok)], 16)
    order = np.argsort(pids, kind="stable")
    rb = batch.to_arrow(["k", "f", "flag"])
    rb.validate(full=True)
    assert rb.schema.names == ["k", "f", "flag"]
    assert rb.column(0).equals(pa.array(key[order], mask=~ok[order]))
    assert rb.column(1).equals(pa.array(f[order]))
    assert rb.column(2).equals(pa.array(flag[order]))
    # zero-copy in HBM: the columns as device arrays an operator takes back
    before = kgpu.abi.lib().kg_ctx_hbm_in_use(ctx().h)
    arr, sch = batch.export_device(["k", "f", "flag"])
    assert arr.device_type == A.ARROW_DEVICE_ROCM and arr.array.n_children == 3
    assert sch.format == b"+s" and sch.children[0].contents.format == b"l"
    child = A.ArrowDeviceArray()
    child.array = arr.array.children[0].contents
    child.device_id, child.device_type = arr.device_id, A.ARROW_DEVICE_ROCM
    kc = kgpu.arrow_col(child, schema=sch.children[0].contents)
    assert kc.device and kc.length == n
    h = kgpu.hash_columns(ctx(), [kc])
    assert np.array_equal(h, kgo.hash_columns([kgo.Col(kgo.T_INT64, key[order], ok[order])]))
    arr.array.release(C.pointer(arr.array))
    sch.release(C.pointer(sch))
    kgpu.abi.lib().kg_ctx_synchronize(ctx().h)
    assert kgpu.abi.lib().kg_ctx_hbm_in_use(ctx().h) < before
