"""NeedMemory at the boundary (crates/krishiv-exec/src/operator.rs:144-161,
314-329): a sink or finalize refused for HBM leaves the operator's state
unchanged, so the same input retried once memory is free gives the answer of
a run that was never refused.  Memory pressure is made by filling the
context's HBM budget with 1 MiB blocks, then releasing them."""
import numpy as np
import pytest

from gpu_util import say
import kgo
import kgpu
from kgpu import Col, DeviceBuffer
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def fill_budget(c, leave_mib):
    """Holds 1 MiB blocks until the budget refuses one, then frees `leave_mib`
    of them; returns the held blocks."""
    held = []
    while True:
        try:
            held.append(DeviceBuffer(c, 1 << 20))
        except kgpu.KgError as e:
            assert e.code == A.ERR_MEMORY
            break
    for _ in range(min(leave_mib, len(held))):
        held.pop().free()
    return held


def release(held):
    for b in held:
        b.free()
    held.clear()


def test_aggregate_sink_and_finalize_refused_then_retried():
    say("needmemory: aggregate")
    c = kgpu.Context(0, 2 << 30)
    n, nd = 3_000_000, 200_000
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(21, 0, 2 * n, cdf, nd)
    types = [(A.T_DECIMAL64, 12, 2), (A.T_FLOAT64, 0, 0), (A.T_INT32, 0, 0)]
    aggs = [(A.AGG_COUNT_STAR, 0)] + [(fn, k) for k in range(3)
                                      for fn in (A.AGG_SUM, A.AGG_MIN, A.AGG_MAX, A.AGG_AVG)]
    dev = {nm: DeviceBuffer.from_numpy(c, v) for nm, v in
           (("key", key), ("dec", dec), ("f", f), ("i32", i32))}

    def batch(b):
        lo = b * n
        return (Col(A.T_INT64, dev["key"], length=n, offset=lo),
                [Col(A.T_DECIMAL64, dev["dec"], precision=12, scale=2, length=n, offset=lo),
                 Col(A.T_FLOAT64, dev["f"], length=n, offset=lo),
                 Col(A.T_INT32, dev["i32"], length=n, offset=lo)])

    g = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), types, aggs, expected_groups=1 << 12)
    held = fill_budget(c, 8)
    k0, in0 = batch(0)
    with pytest.raises(kgpu.KgError, match="UNABLE_TO_ACQUIRE_MEMORY"):
        g.sink(k0, in0, n)
    release(held)
    g.sink(k0, in0, n)  # the same batch again
    k1, in1 = batch(1)
    g.sink(k1, in1, n)
    held = fill_budget(c, 8)
    with pytest.raises(kgpu.KgError, match="UNABLE_TO_ACQUIRE_MEMORY"):
        g.finalize()  # the spill's merge is refused before it starts
    release(held)
    g.finalize()
    r = g.result()
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0),
                                      (kgo.T_INT32, 0, 0)], aggs, exact_f64=True)
    o.update(kgo.Col(kgo.T_INT64, key), [kgo.Col(kgo.T_DECIMAL64, dec, precision=12, scale=2),
                                         kgo.Col(kgo.T_FLOAT64, f), kgo.Col(kgo.T_INT32, i32)])
    rk, _, rres = o.finish(raw=True)
    gk, _ = r.column_raw(0)
    assert len(gk) == len(rk)
    go, ro = np.argsort(gk), np.argsort(rk)
    assert np.array_equal(gk[go], rk[ro])
    for j in range(len(aggs)):
        gv, _ = r.column_raw(1 + j)
        rv = np.asarray(rres[j][0])
        if j in (5, 8, 12):  # float sums and averages
            np.testing.assert_allclose(gv[go], rv[ro], rtol=1e-12)
        else:
            assert np.array_equal(gv[go], rv[ro]), j


def test_gathered_host_batches_refused_then_retried():
    """Small host batches are gathered in HBM; the batch that does not fit
    makes the gathered rows aggregate first.  Refused for memory there, the
    gathered rows stay and the batch is not taken: retried, it is counted
    once."""
    say("needmemory: gathered host batches")
    c = kgpu.Context(0, 2 << 30)
    n, nd, m = 4_400_000, 200_000, 100_000
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(22, 0, n, cdf, nd)
    types = [(A.T_DECIMAL64, 12, 2), (A.T_FLOAT64, 0, 0), (A.T_INT32, 0, 0)]
    aggs = [(A.AGG_COUNT_STAR, 0), (A.AGG_SUM, 0), (A.AGG_MAX, 0), (A.AGG_SUM, 2),
            (A.AGG_MIN, 2)]
    g = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), types, aggs, expected_groups=1 << 12)
    kc = Col(A.T_INT64, key)
    ins = [Col(A.T_DECIMAL64, dec, precision=12, scale=2), Col(A.T_FLOAT64, f),
           Col(A.T_INT32, i32)]
    refused = 0
    for lo in range(0, n, m):
        kc.offset = lo
        for col in ins:
            col.offset = lo
        if lo == 41 * m:  # 4.1M rows gathered: 100K more do not fit its 2^22 rows
            held = fill_budget(c, 8)
            with pytest.raises(kgpu.KgError, match="UNABLE_TO_ACQUIRE_MEMORY"):
                g.sink(kc, ins, min(m, n - lo))
            refused += 1
            release(held)
        g.sink(kc, ins, min(m, n - lo))
    assert refused == 1
    g.finalize()
    r = g.result()
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0),
                                      (kgo.T_INT32, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT64, key), [kgo.Col(kgo.T_DECIMAL64, dec, precision=12, scale=2),
                                         kgo.Col(kgo.T_FLOAT64, f), kgo.Col(kgo.T_INT32, i32)])
    rk, _, rres = o.finish(raw=True)
    gk, _ = r.column_raw(0)
    assert len(gk) == len(rk)
    go, ro = np.argsort(gk), np.argsort(rk)
    assert np.array_equal(gk[go], rk[ro])
    for j in range(len(aggs)):
        gv, _ = r.column_raw(1 + j)
        assert np.array_equal(gv[go], np.asarray(rres[j][0])[ro]), j


def test_join_build_sink_and_finalize_refused_then_retried():
    say("needmemory: join build")
    c = kgpu.Context(0, 512 << 20)
    rng = np.random.default_rng(4)
    nb = 2_000_000
    bk = rng.integers(0, 1_500_000, nb)
    bv = np.arange(nb, dtype=np.int64)
    j = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0], expected_build_rows=nb)
    half = nb // 2
    first = [Col(A.T_INT64, bk[:half]), Col(A.T_INT64, bv[:half])]
    held = fill_budget(c, 4)
    with pytest.raises(kgpu.KgError, match="UNABLE_TO_ACQUIRE_MEMORY"):
        j.build_sink(first, half)
    release(held)
    j.build_sink(first, half)  # retried
    j.build_sink([Col(A.T_INT64, bk[half:]), Col(A.T_INT64, bv[half:])], nb - half)
    held = fill_budget(c, 4)
    with pytest.raises(kgpu.KgError, match="UNABLE_TO_ACQUIRE_MEMORY"):
        j.build_finalize()
    release(held)
    j.build_finalize()
    assert j.build_rows() == nb
    pk = rng.integers(0, 1_600_000, 1_000_000)
    out = j.probe([Col(A.T_INT64, pk)])
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk)], [kgo.Col(kgo.T_INT64, bk)])
    assert out.num_rows == len(op)
    got = np.sort(out.column(0)[0] * 4_000_000 + out.column(2)[0])
    want = np.sort(pk[op] * 4_000_000 + bv[ob])
    assert np.array_equal(got, want)
