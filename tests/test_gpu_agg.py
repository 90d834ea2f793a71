"""GPU parity of the hash aggregation (kg_agg_*) against the CPU oracle.

Integer and decimal results bit-exact; float64 SUM/AVG within 1e-12
relative (their value depends on summation order)."""
import json
import math
import os

import numpy as np
import pytest

from gpu_util import ctx, say, same_float
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

T_MAP = {kgo.T_INT64: A.T_INT64, kgo.T_INT32: A.T_INT32, kgo.T_INT8: A.T_INT8,
         kgo.T_DECIMAL64: A.T_DECIMAL64, kgo.T_DECIMAL128: A.T_DECIMAL128,
         kgo.T_FLOAT64: A.T_FLOAT64, kgo.T_INT16: A.T_INT16}


def run_both(key, kvalid, cols, types, aggs, ansi=False, batches=1, device=False,
             expected_groups=1 << 12, merge=None):
    """Runs the oracle and the GPU aggregate on the same rows; returns
    (oracle result, gpu batch)."""
    n = len(key) if key is not None else len(cols[0][0])
    ktype = (kgo.T_INT64 if key is not None and key.dtype == np.int64 else
             kgo.T_INT8 if key is not None and key.dtype == np.int8 else
             kgo.T_INT32) if key is not None else None
    o = kgo.Agg((ktype, 0, 0) if key is not None else None, types, aggs, ansi=ansi)
    o.update(kgo.Col(ktype, key, kvalid) if key is not None else None,
             [kgo.Col(t[0], v, ok, precision=t[1], scale=t[2]) for (v, ok), t in zip(cols, types)])
    ref = o.finish()
    g = kgpu.HashAggregate(ctx(), (T_MAP[ktype], 0, 0) if key is not None else None,
                           [(T_MAP[t[0]], t[1], t[2]) for t in types], aggs, ansi=ansi,
                           expected_groups=expected_groups)
    if merge is not None:
        g.set_merge(*merge)
    bounds = np.linspace(0, n, batches + 1).astype(np.int64)
    keep = []
    for b in range(batches):
        lo, hi = int(bounds[b]), int(bounds[b + 1])
        if device:
            kc = None
            if key is not None:
                kb = kgpu.DeviceBuffer.from_numpy(ctx(), key[lo:hi])
                vb = (kgpu.DeviceBuffer.from_numpy(ctx(), np.packbits(kvalid[lo:hi], bitorder="little"))
                      if kvalid is not None else None)
                kc = Col(T_MAP[ktype], kb, vb, length=hi - lo)
                keep += [kb, vb]
            ic = []
            for (v, ok), t in zip(cols, types):
                vb = kgpu.DeviceBuffer.from_numpy(ctx(), v[lo:hi])
                bb = (kgpu.DeviceBuffer.from_numpy(ctx(), np.packbits(ok[lo:hi], bitorder="little"))
                      if ok is not None else None)
                ic.append(Col(T_MAP[t[0]], vb, bb, precision=t[1], scale=t[2], length=hi - lo))
                keep += [vb, bb]
        else:
            kc = (Col(T_MAP[ktype], key[lo:hi], None if kvalid is None else kvalid[lo:hi])
                  if key is not None else None)
            ic = [Col(T_MAP[t[0]], v[lo:hi], None if ok is None else ok[lo:hi], precision=t[1],
                      scale=t[2]) for (v, ok), t in zip(cols, types)]
        g.sink(kc, ic, hi - lo)
    g.finalize()
    return ref, g.result(), g


def compare(ref, batch, aggs, has_key=True):
    rk, rkv, rres = ref
    base = 1 if has_key else 0
    assert batch.num_rows == len(rk), (batch.num_rows, len(rk))
    if has_key:
        gk, gkv = batch.column(0)
        gmap = {(int(k) if v else None): i for i, (k, v) in enumerate(zip(gk, gkv))}
    else:
        gmap = {0: 0}
    cols = [batch.column(base + j) for j in range(len(aggs))]
    types = [batch.column_type(base + j) for j in range(len(aggs))]
    for ri in range(len(rk)):
        k = (int(rk[ri]) if rkv[ri] else None) if has_key else 0
        gi = gmap[k]
        for j in range(len(aggs)):
            rv, rok, (rt, rp, rs) = rres[j]
            gv, gok = cols[j]
            assert types[j][0] == T_MAP.get(rt, rt) or (rt, types[j][0]) in (
                (kgo.T_FLOAT64, A.T_FLOAT64),), (j, rt, types[j])
            assert bool(rok[ri]) == bool(gok[gi]), (k, j, rok[ri], gok[gi])
            if not rok[ri]:
                continue
            a, b = rv[ri], gv[gi]
            if rt == kgo.T_FLOAT64:
                assert same_float(float(a), float(b)), (k, j, a, b)
            else:
                assert int(a) == int(b), (k, j, a, b)


ALL6 = [kgo.AGG_COUNT, kgo.AGG_SUM, kgo.AGG_MIN, kgo.AGG_MAX, kgo.AGG_AVG]


def config2_aggs(ncols=3):
    aggs = [(kgo.AGG_COUNT_STAR, 0)]
    for c in range(ncols):
        aggs += [(kgo.AGG_SUM, c), (kgo.AGG_MIN, c), (kgo.AGG_MAX, c), (kgo.AGG_AVG, c)]
    return aggs


def test_config2_shape_zipf_int64_key():
    say("agg: config-2 shape, 300K rows, Zipf(1.1) over 100K keys")
    n, nd = 300_000, 100_000
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(3, 0, n, cdf, nd)
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0), (kgo.T_INT32, 0, 0)]
    aggs = config2_aggs()
    ref, b, _ = run_both(key, None, [(dec, None), (f, None), (i32, None)], types, aggs,
                         expected_groups=nd)
    compare(ref, b, aggs)


def test_config2_device_resident_multi_batch():
    say("agg: device-resident inputs, 5 batches")
    n, nd = 200_000, 50_000
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(5, 0, n, cdf, nd)
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0), (kgo.T_INT32, 0, 0)]
    aggs = config2_aggs()
    ref, b, _ = run_both(key, None, [(dec, None), (f, None), (i32, None)], types, aggs,
                         batches=5, device=True, expected_groups=1024)  # forces growth
    compare(ref, b, aggs)


def test_int8_key_direct_path():
    say("agg: int8 key with 4 values (direct path)")
    n = 100_000
    cdf = kgo.zipf_cdf(1000, 1.1)
    _, dec, f, i32, k8 = kgo.gen_agg(7, 0, n, cdf, 1000)
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0), (kgo.T_INT32, 0, 0)]
    aggs = config2_aggs()
    ref, b, _ = run_both(k8, None, [(dec, None), (f, None), (i32, None)], types, aggs)
    compare(ref, b, aggs)
    # negative one-byte keys and a NULL key
    k = np.array([-128, -1, 0, 127, -1, 5] * 100, dtype=np.int8)
    kv = np.array([True, True, True, True, False, True] * 100)
    v = np.arange(600, dtype=np.int64) - 300
    ref, b, _ = run_both(k, kv, [(v, None)], [(kgo.T_INT64, 0, 0)],
                         [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_MIN, 0)])
    compare(ref, b, [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_MIN, 0)])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_nulls_and_signs(seed):
    say(f"agg: random nulls/signs seed={seed}")
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1000, 60000))
    ng = int(rng.integers(1, 5000))
    key = rng.integers(-ng, ng, n).astype(np.int64)
    key[rng.random(n) < 0.01] = np.iinfo(np.int64).min  # the table's sentinel value
    kv = rng.random(n) > 0.03
    i64 = rng.integers(-(1 << 40), 1 << 40, n)
    i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
    dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, n)
    f = rng.random(n) * 1e3  # one sign: float sums compare relatively
    f[rng.random(n) < 0.01] = np.nan
    f[rng.random(n) < 0.01] = -np.inf
    oks = [rng.random(n) > 0.1 for _ in range(4)]
    types = [(kgo.T_INT64, 0, 0), (kgo.T_INT32, 0, 0), (kgo.T_DECIMAL64, 12, 2),
             (kgo.T_FLOAT64, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0)] + [(fn, c) for c in range(4) for fn in ALL6]
    ref, b, _ = run_both(key, kv, list(zip([i64, i32, dec, f], oks)), types, aggs,
                         batches=3, expected_groups=ng)
    compare(ref, b, aggs)


def test_global_aggregate_and_empty_input():
    say("agg: global aggregate")
    v = np.arange(1, 10001, dtype=np.int64)
    aggs = [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_AVG, 0), (kgo.AGG_MAX, 0)]
    ref, b, _ = run_both(None, None, [(v, None)], [(kgo.T_INT64, 0, 0)], aggs)
    compare(ref, b, aggs, has_key=False)
    # a global aggregate on empty input is one row: count 0, the rest NULL
    g = kgpu.HashAggregate(ctx(), None, [(A.T_INT64, 0, 0)], aggs)
    g.finalize()
    r = g.result()
    assert r.num_rows == 1
    assert r.column(0)[0][0] == 0 and not r.column(1)[1][0]
    # a grouped aggregate on empty input has no rows
    g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), [(A.T_INT64, 0, 0)], aggs)
    g.finalize()
    assert g.result().num_rows == 0


def test_int_sum_wraps_or_raises():
    say("agg: int64 sum overflow")
    big = np.array([1 << 62] * 3, dtype=np.int64)
    aggs = [(kgo.AGG_SUM, 0)]
    ref, b, _ = run_both(None, None, [(big, None)], [(kgo.T_INT64, 0, 0)], aggs)
    compare(ref, b, aggs, has_key=False)
    with pytest.raises(kgpu.KgError, match="ARITHMETIC_OVERFLOW"):
        g = kgpu.HashAggregate(ctx(), None, [(A.T_INT64, 0, 0)], aggs, ansi=True)
        g.sink(None, [Col(A.T_INT64, big)])
        g.finalize()
        g.result()


def test_spark_decimal_goldens_single_partition():
    """Spark 4.2.0 decimal sum/avg goldens whose input is one partition
    (row order replayed exactly on the device when the proof fails)."""
    with open(os.path.join(GOLD, "spark_decimal_sum_avg.json")) as f:
        cases = [c for c in json.load(f) if c["partitions"] == 1]
    say(f"agg: {len(cases)} Spark decimal goldens")
    for c in cases:
        p, s = c["type"]
        func = A.AGG_SUM if c["fn"] == "sum" else A.AGG_AVG
        vals = [int(x) for x in c["values"]]
        groups = c["groups"]
        g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), [(A.T_DECIMAL128, p, s)], [(func, 0)],
                               ansi=c["ansi"])
        keys = np.array([i % groups for i in range(len(vals))], dtype=np.int64)
        try:
            g.sink(Col(A.T_INT64, keys), [Col(A.T_DECIMAL128, vals, precision=p, scale=s)])
            g.finalize()
            r = g.result()
            err = None
        except kgpu.KgError as e:
            err = str(e)
        if "error" in c:
            assert err is not None and c["error"] in err, (c, err)
            continue
        assert err is None, (c, err)
        t, op, os_ = r.column_type(1)
        assert c["type_out"] == f"decimal({op},{os_})"
        gk, _ = r.column(0)
        gv, gok = r.column(1)
        got = {int(k): (v if ok else None) for k, v, ok in zip(gk, gv, gok)}
        for gi, want in enumerate(c["value"]):
            v = got[gi]
            if want is None:
                assert v is None, (c, gi, v)
            else:
                neg = v < 0
                d = str(abs(v)).rjust(os_ + 1, "0")
                text = ("-" if neg else "") + (d[:-os_] + "." + d[-os_:] if os_ else d)
                assert text == want, (c, gi, text)


def test_tpch_q1_matches_spark():
    """TPC-H Q1 aggregates (decimal128 sums/averages) on the device against
    Spark's answer."""
    import pyarrow.ipc as ipc
    say("agg: TPC-H Q1")
    t = ipc.open_file(os.path.join(GOLD, "tpch_q1_input.arrow")).read_all()
    code = {}
    key = np.array([code.setdefault((a, b), len(code)) for a, b in
                    zip(t.column("l_returnflag").to_pylist(), t.column("l_linestatus").to_pylist())],
                   dtype=np.int64)
    cols, types = [], []
    for name in ("l_quantity", "l_extendedprice", "disc_price", "charge", "l_discount"):
        col = t.column(name).combine_chunks()
        cols.append(Col.from_arrow(col))
        types.append((A.T_DECIMAL128, col.type.precision, col.type.scale))
    aggs = [(A.AGG_SUM, 0), (A.AGG_SUM, 1), (A.AGG_SUM, 2), (A.AGG_SUM, 3), (A.AGG_AVG, 0),
            (A.AGG_AVG, 1), (A.AGG_AVG, 4), (A.AGG_COUNT_STAR, 0)]
    g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), types, aggs, ansi=True)
    g.sink(Col(A.T_INT64, key), cols)
    g.finalize()
    r = g.result()
    inv = {v: k for k, v in code.items()}
    gk, _ = r.column(0)
    rows = []
    for gi in range(r.num_rows):
        f = list(inv[int(gk[gi])])
        for j in range(len(aggs)):
            tt, p, s = r.column_type(1 + j)
            v = r.column(1 + j)[0][gi]
            if tt in (A.T_DECIMAL64, A.T_DECIMAL128):
                neg = v < 0
                d = str(abs(int(v))).rjust(s + 1, "0")
                f.append(("-" if neg else "") + (d[:-s] + "." + d[-s:] if s else d))
            else:
                f.append(str(int(v)))
        rows.append(",".join(f))
    with open(os.path.join(GOLD, "tpch_q1_expected.csv")) as fh:
        want = sorted(l.strip() for l in fh if l.strip())
    assert sorted(rows) == want


def _two_phase(keys_parts, vals_parts, types, aggs, ansi):
    """Partial aggregate per partition (device), then the final aggregate
    merging the partial states in partition order."""
    final = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), types, aggs, ansi=ansi, merge=True)
    keep = []
    for keys, vals in zip(keys_parts, vals_parts):
        p = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), types, aggs, ansi=ansi, partial=True)
        p.sink(Col(A.T_INT64, keys), vals)
        p.finalize()
        st = p.result()
        keep += [p, st]
        final.sink(st.as_col(0), [st.as_col(i) for i in range(1, st.num_columns)], st.num_rows)
    final.finalize()
    return final.result()


def test_spark_decimal_goldens_two_phase():
    """Every Spark decimal golden through PartialAggregate + the final merge
    (contiguous partitions merged in order, as Spark's range partitioning)."""
    with open(os.path.join(GOLD, "spark_decimal_sum_avg.json")) as f:
        cases = json.load(f)
    say(f"agg: {len(cases)} Spark decimal goldens, two-phase")
    for c in cases:
        p, s = c["type"]
        func = A.AGG_SUM if c["fn"] == "sum" else A.AGG_AVG
        vals = [int(x) for x in c["values"]]
        groups, parts = c["groups"], c["partitions"]
        n = len(vals)
        kp, vp = [], []
        for k in range(parts):
            lo, hi = k * n // parts, (k + 1) * n // parts
            kp.append(np.array([i % groups for i in range(lo, hi)], dtype=np.int64))
            vp.append([Col(A.T_DECIMAL128, vals[lo:hi], precision=p, scale=s)])
        try:
            r = _two_phase(kp, vp, [(A.T_DECIMAL128, p, s)], [(func, 0)], c["ansi"])
            err = None
        except kgpu.KgError as e:
            err = str(e)
        if "error" in c:
            assert err is not None and c["error"] in err, (c, err)
            continue
        assert err is None, (c, err)
        t, op, os_ = r.column_type(1)
        assert c["type_out"] == f"decimal({op},{os_})"
        gk, _ = r.column(0)
        gv, gok = r.column(1)
        got = {int(k): (v if ok else None) for k, v, ok in zip(gk, gv, gok)}
        for gi, want in enumerate(c["value"]):
            v = got[gi]
            if want is None:
                assert v is None, (c, gi, v)
            else:
                neg = v < 0
                d = str(abs(v)).rjust(os_ + 1, "0")
                assert ("-" if neg else "") + (d[:-os_] + "." + d[-os_:] if os_ else d) == want


def test_two_phase_equals_single_phase():
    say("agg: two-phase vs oracle")
    rng = np.random.default_rng(21)
    n = 90000
    key = rng.integers(0, 3000, n)
    i64 = rng.integers(-(1 << 40), 1 << 40, n)
    dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, n)
    f = rng.random(n)
    oks = [rng.random(n) > 0.1 for _ in range(3)]
    types = [(kgo.T_INT64, 0, 0), (kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0)] + [(fn, c) for c in range(3) for fn in ALL6]
    o = kgo.Agg((kgo.T_INT64, 0, 0), types, aggs)
    o.update(kgo.Col(kgo.T_INT64, key), [kgo.Col(t[0], v, ok, precision=t[1], scale=t[2])
                                         for v, ok, t in zip((i64, dec, f), oks, types)])
    ref = o.finish()
    cuts = [0, 20000, 55000, n]
    kp = [key[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    vp = [[Col(T_MAP[t[0]], v[a:b], ok[a:b], t[1], t[2]) for v, ok, t in zip((i64, dec, f), oks, types)]
          for a, b in zip(cuts[:-1], cuts[1:])]
    r = _two_phase(kp, vp, [(T_MAP[t[0]], t[1], t[2]) for t in types], aggs, False)
    compare(ref, r, aggs)


@pytest.mark.parametrize("merge", [(A.AGG_MERGE_BY_KEY, False), (A.AGG_MERGE_WINDOWS, False),
                                   (A.AGG_MERGE_AUTO, True)],
                         ids=["by-key-grows", "windows-grow", "auto-sketch"])
def test_merge_resumes_after_table_growth(merge):
    """200K distinct keys into a table sized for 16 groups.  By key, with the
    sketch's sizing off: the spill merge runs out of table room, stops, the
    table grows, the leftover records and pending rows merge with atomics
    and the stopped partitions resume.  In windows, sizing off: windows fill,
    are left as they were, and run again after each growth of the table,
    the windows done skipped.  Auto: the sketch of the spilled keys sizes
    the table before the merge."""
    say(f"agg: merge with table growth {merge}")
    rng = np.random.default_rng(11)
    n = 600_000
    key = (rng.integers(0, 200_000, n) * 7919 - (1 << 40)).astype(np.int64)
    v = rng.integers(-(1 << 50), 1 << 50, n)
    f = rng.random(n)
    ok = rng.random(n) > 0.05
    types = [(kgo.T_INT64, 0, 0), (kgo.T_FLOAT64, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_MIN, 0), (kgo.AGG_MAX, 1),
            (kgo.AGG_AVG, 1), (kgo.AGG_COUNT, 0)]
    ref, b, _ = run_both(key, None, [(v, ok), (f, None)], types, aggs, batches=4,
                         expected_groups=16, merge=merge)
    compare(ref, b, aggs)


def test_host_batches_are_gathered():
    """Host batches smaller than a staging slot (512K rows here) are gathered
    in HBM before they are aggregated: ragged batches at odd row and bit
    offsets into the caller's arrays, NULL keys and values, the table's
    sentinel key, batches of a slot or more (the gathered rows are
    aggregated first), a device batch, and more rows than the 4M-row HBM
    buffer holds; bit-exact against the oracle."""
    say("agg: gathered host batches")
    rng = np.random.default_rng(12)
    n = 5_000_000
    key = rng.integers(-3000, 3000, n).astype(np.int64)
    key[rng.random(n) < 0.001] = np.iinfo(np.int64).min
    kv = rng.random(n) > 0.02
    v = rng.integers(-(1 << 40), 1 << 40, n)
    vv = rng.random(n) > 0.05
    types = [(kgo.T_INT64, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_MIN, 0), (kgo.AGG_MAX, 0),
            (kgo.AGG_COUNT, 0)]
    o = kgo.Agg((kgo.T_INT64, 0, 0), types, aggs)
    o.update(kgo.Col(kgo.T_INT64, key, kv), [kgo.Col(kgo.T_INT64, v, vv)])
    ref = o.finish()
    g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), [(A.T_INT64, 0, 0)], aggs,
                           expected_groups=1 << 10)
    sizes = []
    left = n
    while left > 0:
        r = rng.random()
        m = int(rng.integers(600_000, 700_000)) if r < 0.02 else int(rng.integers(1, 70_000))
        m = min(m, left)
        sizes.append(m)
        left -= m
    dev_at = len(sizes) // 2
    lo = 0
    keep = []
    # host batches view the whole arrays at element (and bit) offset lo
    kc = Col(A.T_INT64, key, kv, length=n)
    vc = Col(A.T_INT64, v, vv, length=n)
    for b, m in enumerate(sizes):
        if b == dev_at:
            kb = kgpu.DeviceBuffer.from_numpy(ctx(), key[lo:lo + m])
            kvb = kgpu.DeviceBuffer.from_numpy(ctx(), np.packbits(kv[lo:lo + m], bitorder="little"))
            vb = kgpu.DeviceBuffer.from_numpy(ctx(), v[lo:lo + m])
            vvb = kgpu.DeviceBuffer.from_numpy(ctx(), np.packbits(vv[lo:lo + m], bitorder="little"))
            keep += [kb, kvb, vb, vvb]
            g.sink(Col(A.T_INT64, kb, kvb, length=m), [Col(A.T_INT64, vb, vvb, length=m)], m)
        else:
            kc.offset = vc.offset = lo
            g.sink(kc, [vc], m)
        lo += m
    g.finalize()
    compare(ref, g.result(), aggs)


def test_host_partial_states_are_gathered():
    """The final phase of a two-phase aggregate fed host copies of the
    partial states (decimal128 sums, BOOL nonempty flags, counts) in small
    batches, which it gathers; against the oracle over all rows."""
    say("agg: gathered host partial states")
    rng = np.random.default_rng(13)
    n = 120_000
    key = rng.integers(0, 5000, n)
    dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, n)
    i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
    oks = [rng.random(n) > 0.1, rng.random(n) > 0.3]
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_INT32, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0)] + [(fn, c) for c in range(2) for fn in ALL6]
    o = kgo.Agg((kgo.T_INT64, 0, 0), types, aggs)
    o.update(kgo.Col(kgo.T_INT64, key), [kgo.Col(t[0], vals, ok, precision=t[1], scale=t[2])
                                         for vals, ok, t in zip((dec, i32), oks, types)])
    ref = o.finish()
    at = [(T_MAP[t[0]], t[1], t[2]) for t in types]
    final = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), at, aggs, merge=True)
    cuts = [0, 40_000, 95_000, n]
    for a_, b_ in zip(cuts[:-1], cuts[1:]):
        p = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), at, aggs, partial=True)
        p.sink(Col(A.T_INT64, key[a_:b_]),
               [Col(T_MAP[t[0]], vals[a_:b_], ok[a_:b_], t[1], t[2])
                for vals, ok, t in zip((dec, i32), oks, types)])
        p.finalize()
        st = p.result()
        host = []
        for i in range(st.num_columns):
            ct, cp, cs = st.column_type(i)
            vals, ok = st.column_raw(i)
            if ct == A.T_DECIMAL128:
                vals = vals.view(np.uint64)
            host.append((ct, cp, cs, vals, ok))
        m = st.num_rows
        for lo in range(0, m, 777):
            hi = min(m, lo + 777)
            cols = [Col(ct, vals[lo:hi], ok[lo:hi], cp, cs) for ct, cp, cs, vals, ok in host]
            final.sink(cols[0], cols[1:], hi - lo)
    final.finalize()
    compare(ref, final.result(), aggs)
