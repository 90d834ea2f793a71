"""GPU parity of the hash join (build sink + probe operator) and of the join
fused with GROUP BY the build key, against the oracle: join outputs compared
as sorted row multisets with Spark NULL semantics, aggregates bit-exact."""
import numpy as np
import pytest

from gpu_util import ctx, same_float, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu

JT = {A.JOIN_INNER: kgo.JOIN_INNER, A.JOIN_LEFT_OUTER: kgo.JOIN_LEFT_OUTER,
      A.JOIN_LEFT_SEMI: kgo.JOIN_LEFT_SEMI, A.JOIN_LEFT_ANTI: kgo.JOIN_LEFT_ANTI}


def gpu_join(jt, probe_cols, build_cols, pk, bk, null_safe=None, build_batches=1, direct=True):
    j = kgpu.HashJoin(ctx(), jt, pk, bk, null_safe, expected_build_rows=1024, direct_index=direct)
    n = build_cols[0].length
    cuts = np.linspace(0, n, build_batches + 1).astype(int)
    for a, b in zip(cuts[:-1], cuts[1:]):
        j.build_sink([Col(c.type, c.values_np[a:b], None if c.valid_np is None else c.valid_np[a:b],
                          c.precision, c.scale) for c in build_cols], b - a)
    j.build_finalize()
    out = j.probe(probe_cols)
    return j, out


def mk(ctype, values, valid=None):
    c = Col(ctype, values, valid)
    c.values_np = np.asarray(values)
    c.valid_np = None if valid is None else np.asarray(valid)
    return c


def rows_of(batch, ncols):
    cols = [batch.column(i) for i in range(ncols)]
    out = []
    for r in range(batch.num_rows):
        out.append(tuple((int(v[r]) if ok[r] else None) for v, ok in cols))
    return sorted(out, key=lambda t: tuple((x is None, x or 0) for x in t))


def expected_rows(jt, pairs, probe, build):
    op, ob = pairs
    out = []
    for i, j in zip(op, ob):
        row = [(int(v[i]) if ok is None or ok[i] else None) for v, ok in probe]
        if jt in (A.JOIN_INNER, A.JOIN_LEFT_OUTER):
            row += [None] * len(build) if j < 0 else \
                [(int(v[j]) if ok is None or ok[j] else None) for v, ok in build]
        out.append(tuple(row))
    return sorted(out, key=lambda t: tuple((x is None, x or 0) for x in t))


def probed_direct(run):
    """Runs `run()` with the library's timers on; returns whether the probe
    went through the direct index (its own timer family) or the table."""
    kgpu.prof_reset()
    kgpu.prof_enable(True)
    try:
        run()
    finally:
        kgpu.prof_enable(False)
    d = kgpu.prof_get("join_probe_direct")[1] + kgpu.prof_get("joinagg_direct")[1] + \
        kgpu.prof_get("joinagg_probe_direct")[1]
    h = kgpu.prof_get("join_probe")[1] + kgpu.prof_get("joinagg_probe")[1] + \
        kgpu.prof_get("jp_merge")[1]
    assert d + h > 0
    return d > 0 and h == 0


@pytest.mark.parametrize("direct", [True, False])
@pytest.mark.parametrize("jt", [A.JOIN_INNER, A.JOIN_LEFT_OUTER, A.JOIN_LEFT_SEMI,
                                A.JOIN_LEFT_ANTI])
def test_join_types_with_duplicates_and_nulls(jt, direct):
    """3,000 build rows on keys spanning 2,000 values: the direct index
    (join/direct.rs) when allowed, the hashed table otherwise; the same rows
    either way."""
    say(f"join: type {jt}, direct index {direct}")
    rng = np.random.default_rng(10 + jt)
    nb, npr = 3000, 8000
    bk = rng.integers(0, 2000, nb)
    bv = rng.random(nb) > 0.05
    bpay = rng.integers(-(1 << 40), 1 << 40, nb)
    bpv = rng.random(nb) > 0.2
    pk = rng.integers(0, 2500, npr)
    pv = rng.random(npr) > 0.05
    ppay = np.arange(npr, dtype=np.int64)
    build = [mk(A.T_INT64, bk, bv), mk(A.T_INT64, bpay, bpv)]
    probe = [mk(A.T_INT64, pk, pv), mk(A.T_INT64, ppay)]
    res = {}
    assert probed_direct(lambda: res.update(out=gpu_join(jt, probe, build, [0], [0], build_batches=3,
                                                         direct=direct)[1])) == direct
    out = res["out"]
    pairs = kgo.join(JT[jt], [kgo.Col(kgo.T_INT64, pk, pv)], [kgo.Col(kgo.T_INT64, bk, bv)])
    want = expected_rows(jt, pairs, [(pk, pv), (ppay, None)], [(bk, bv), (bpay, bpv)])
    ncols = 4 if jt in (A.JOIN_INNER, A.JOIN_LEFT_OUTER) else 2
    assert out.num_columns == ncols
    assert rows_of(out, ncols) == want


@pytest.mark.parametrize("case", ["sentinel", "int32", "too_sparse", "too_many_rows"])
def test_direct_index_qualification_and_edges(case):
    """Direct-index edges (join/direct.rs:96-160): keys at the bottom of the
    int64 range (the table's EMPTY value among them), an int32 key with
    probe keys outside the span, and builds that keep the hashed table (a
    span above 4 slots a row; more than 4,096 rows)."""
    say(f"join: direct index, {case}")
    rng = np.random.default_rng(len(case))
    t, kt = (A.T_INT32, kgo.T_INT32) if case == "int32" else (A.T_INT64, kgo.T_INT64)
    nb = 5000 if case == "too_many_rows" else 1200
    if case == "sentinel":
        bk = np.iinfo(np.int64).min + rng.integers(0, 900, nb)
        pk = np.iinfo(np.int64).min + rng.integers(0, 1000, 30000)
        pk[::7] = rng.integers(-(1 << 62), 1 << 62, len(pk[::7]))
    elif case == "int32":
        bk = rng.integers(-2000, 2000, nb).astype(np.int32)
        pk = rng.integers(-(1 << 31), (1 << 31) - 1, 30000).astype(np.int32)
        pk[::2] = rng.integers(-2100, 2100, 15000)
    elif case == "too_sparse":
        bk = rng.integers(0, 1_000_000, nb)
        pk = np.concatenate([bk[rng.integers(0, nb, 20000)], rng.integers(0, 1_000_000, 10000)])
    else:
        bk = rng.integers(0, 6000, nb)
        pk = rng.integers(-100, 6100, 30000)
    bv = rng.random(nb) > 0.03
    pv = rng.random(len(pk)) > 0.03
    bpay = np.arange(nb, dtype=np.int64)
    ppay = np.arange(len(pk), dtype=np.int64)
    build = [mk(t, bk, bv), mk(A.T_INT64, bpay)]
    probe = [mk(t, pk, pv), mk(A.T_INT64, ppay)]
    res = {}
    used = probed_direct(lambda: res.update(out=gpu_join(A.JOIN_LEFT_OUTER, probe, build, [0], [0])[1]))
    assert used == (case in ("sentinel", "int32"))
    pairs = kgo.join(kgo.JOIN_LEFT_OUTER, [kgo.Col(kt, pk, pv)], [kgo.Col(kt, bk, bv)])
    want = expected_rows(A.JOIN_LEFT_OUTER, pairs, [(pk, pv), (ppay, None)], [(bk, bv), (bpay, None)])
    assert rows_of(res["out"], 4) == want


@pytest.mark.parametrize("nprobe", [200_000, 3_000_000])
def test_direct_index_fused_aggregate(nprobe):
    """The fused GROUP BY build key over a direct index: records privatized
    per workgroup in LDS (probes of 64K rows and more), duplicate build keys
    (the row count multiplies), NULL probe keys and inputs, probe keys
    outside the span."""
    say(f"join: direct-index fused aggregate, {nprobe} probe rows")
    rng = np.random.default_rng(nprobe)
    nb = 2000
    bk = rng.integers(-700, 1300, nb).astype(np.int32)
    pk = rng.integers(-1000, 1600, nprobe).astype(np.int32)
    pkv = rng.random(nprobe) > 0.02
    amt = rng.integers(-(1 << 40), 1 << 40, nprobe)
    amtv = rng.random(nprobe) > 0.03
    f = rng.random(nprobe)
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0), (kgo.AGG_MAX, 1), (kgo.AGG_AVG, 1)]
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT32, pk, pkv)], [kgo.Col(kgo.T_INT32, bk)])
    o = kgo.Agg((kgo.T_INT32, 0, 0), [(kgo.T_INT64, 0, 0), (kgo.T_FLOAT64, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT32, bk[ob]),
             [kgo.Col(kgo.T_INT64, amt[op], amtv[op]), kgo.Col(kgo.T_FLOAT64, f[op])])
    rk, rkv, rres = o.finish()
    want = _groups(rk, rkv, [(np.asarray(v), np.asarray(ok)) for v, ok, _ in rres])
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=nb)
    j.build_sink([Col(A.T_INT32, bk)])
    j.build_finalize()
    j.aggregate([(A.T_INT64, 0, 0), (A.T_FLOAT64, 0, 0)], aggs)
    # (a host batch below 512K rows is gathered and probed when the result is
    # taken: both calls are timed)
    res = {}
    assert probed_direct(lambda: (j.probe_aggregate(
        [Col(A.T_INT32, pk, pkv), Col(A.T_INT64, amt, amtv), Col(A.T_FLOAT64, f)], [1, 2]),
        res.update(r=j.aggregate_result())))
    r = res["r"]
    gk, gkv = r.column(0)
    got = _groups(gk, gkv, [r.column(1 + c) for c in range(len(aggs))])
    assert set(got) == set(want)
    for k, wv in want.items():
        for c, (a, b) in enumerate(zip(wv, got[k])):
            assert (a is None) == (b is None), (k, c)
            if a is None:
                continue
            if isinstance(a, (float, np.floating)):
                assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)) + 1e-300, (k, c, a, b)
            else:
                assert int(a) == int(b), (k, c, a, b)


def test_null_safe_key_and_float_normalization():
    probe = [mk(A.T_FLOAT64, np.array([0.0, np.nan, 1.0, 2.0]), np.array([1, 1, 1, 0], bool))]
    build = [mk(A.T_FLOAT64, np.array([-0.0, -np.nan, 3.0, 0.0]), np.array([1, 1, 1, 0], bool)),
             mk(A.T_INT64, np.array([10, 11, 12, 13]))]
    for ns, want in ((None, [(0, 10), (1, 11)]), ([1], [(0, 10), (1, 11), (3, 13)])):
        j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], ns)
        j.build_sink(build)
        j.build_finalize()
        out = j.probe(probe + [mk(A.T_INT64, np.arange(4))])
        # columns: probe f64, probe row id, build f64, build payload
        rid, _ = out.column(1)
        pay, _ = out.column(3)
        assert sorted(zip(rid.tolist(), pay.tolist())) == want


def test_two_int32_keys():
    rng = np.random.default_rng(5)
    nb, npr = 2000, 5000
    b1, b2 = rng.integers(0, 30, nb).astype(np.int32), rng.integers(-20, 20, nb).astype(np.int32)
    p1, p2 = rng.integers(0, 35, npr).astype(np.int32), rng.integers(-25, 25, npr).astype(np.int32)
    build = [mk(A.T_INT32, b1), mk(A.T_INT32, b2), mk(A.T_INT64, np.arange(nb))]
    probe = [mk(A.T_INT32, p1), mk(A.T_INT32, p2), mk(A.T_INT64, np.arange(npr))]
    j, out = gpu_join(A.JOIN_INNER, probe, build, [0, 1], [0, 1])
    pairs = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT32, p1), kgo.Col(kgo.T_INT32, p2)],
                     [kgo.Col(kgo.T_INT32, b1), kgo.Col(kgo.T_INT32, b2)])
    got = sorted(zip(out.column(2)[0].tolist(), out.column(5)[0].tolist()))
    assert got == sorted(zip(pairs[0].tolist(), pairs[1].tolist()))


def test_config1_shape_join_then_aggregate_and_fused():
    """SELECT key, SUM(amt), COUNT(*) FROM build JOIN probe ON key = fk
    GROUP BY key: 100K build (a permutation), 800K probe (fk over 1.2x the
    range, 1 % NULL): join -> device aggregate, and the fused form, both
    bit-exact against the oracle."""
    say("join: config-1 shape")
    nb, npr = 100_000, 800_000
    key, val = kgo.gen_join_build(1, 0, nb, nb)
    fk, fkv, amt = kgo.gen_join_probe(2, 0, npr, int(nb * 1.2), 10)
    # oracle: join pairs, then the aggregate over the pairs
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, fk, fkv)], [kgo.Col(kgo.T_INT64, key)])
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)])
    o.update(kgo.Col(kgo.T_INT64, key[ob]), [kgo.Col(kgo.T_INT64, amt[op])])
    rk, _, rres = o.finish()
    want = sorted(zip(rk.tolist(), rres[0][0].tolist(), rres[1][0].tolist()))
    # device: join, then aggregate the join output in HBM (no copy)
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=nb)
    j.build_sink([Col(A.T_INT64, key), Col(A.T_INT64, val)])
    j.build_finalize()
    out = j.probe([Col(A.T_INT64, fk, fkv), Col(A.T_INT64, amt)])
    g = kgpu.HashAggregate(ctx(), (A.T_INT64, 0, 0), [(A.T_INT64, 0, 0)],
                           [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)], expected_groups=nb)
    g.sink(out.as_col(2), [out.as_col(1)], out.num_rows)
    g.finalize()
    r = g.result()
    got = sorted(zip(r.column(0)[0].tolist(), r.column(1)[0].tolist(), r.column(2)[0].tolist()))
    assert got == want
    # fused: the aggregate state in the build table's slots
    j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
    j.probe_aggregate([Col(A.T_INT64, fk, fkv), Col(A.T_INT64, amt)], [1])
    r = j.aggregate_result()
    got = sorted(zip(r.column(0)[0].tolist(), r.column(1)[0].tolist(), r.column(2)[0].tolist()))
    assert got == want


def test_fused_aggregate_with_duplicate_build_keys():
    rng = np.random.default_rng(9)
    bk = rng.integers(0, 500, 3000)
    pk = rng.integers(0, 600, 20000)
    amt = rng.integers(-1000, 1000, 20000)
    f = rng.random(20000)
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk)], [kgo.Col(kgo.T_INT64, bk)])
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0), (kgo.AGG_MIN, 0), (kgo.AGG_MAX, 1),
            (kgo.AGG_AVG, 0)]
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0), (kgo.T_FLOAT64, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT64, bk[ob]), [kgo.Col(kgo.T_INT64, amt[op]),
                                            kgo.Col(kgo.T_FLOAT64, f[op])])
    rk, _, rres = o.finish()
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0])
    j.build_sink([Col(A.T_INT64, bk)])
    j.build_finalize()
    j.aggregate([(A.T_INT64, 0, 0), (A.T_FLOAT64, 0, 0)], aggs)
    j.probe_aggregate([Col(A.T_INT64, pk), Col(A.T_INT64, amt), Col(A.T_FLOAT64, f)], [1, 2])
    r = j.aggregate_result()
    gk = r.column(0)[0]
    gi = {int(k): i for i, k in enumerate(gk)}
    assert len(gi) == len(rk)
    for ri, k in enumerate(rk):
        i = gi[int(k)]
        assert int(r.column(1)[0][i]) == int(rres[0][0][ri])
        assert int(r.column(2)[0][i]) == int(rres[1][0][ri])
        assert int(r.column(3)[0][i]) == int(rres[2][0][ri])
        assert float(r.column(4)[0][i]) == float(rres[3][0][ri])
        assert abs(float(r.column(5)[0][i]) - float(rres[4][0][ri])) <= 1e-12 * abs(float(rres[4][0][ri]))


# probe rows: the per-row probe (global atomics), the direct index (dense
# build keys, LDS records) and the radix-partitioned path
@pytest.mark.parametrize("path,nb_keys,npr", [("probe", 500, 20_000), ("direct", 500, 200_000),
                                              ("partitioned", 300_000, 1_200_000)])
def test_fused_aggregate_decimal_and_ansi_sums(path, nb_keys, npr):
    """Decimal SUM/AVG and the ANSI integer SUM (checked sums) in the fused
    aggregate, with duplicate build keys (each probe row counts once per
    pair) and NULL inputs: every batch is proven unable to overflow, so the
    sums are exact; against the oracle's join + ANSI aggregate."""
    say(f"join: fused decimal/ANSI sums, {path} path")
    rng = np.random.default_rng(nb_keys + npr)
    nb = nb_keys * 3
    bk = rng.integers(0, nb_keys, nb).astype(np.int64)
    if path == "probe":
        bk = bk * 7919 + 12345  # sparse keys: no direct index
    pk = bk[rng.integers(0, nb, npr)]
    pk[rng.random(npr) < 0.2] += 1
    dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, npr)
    decv = rng.random(npr) > 0.05
    # (|value| x rows x pairs a row stays below 2^63: the ANSI sum is proven)
    i64 = rng.integers(-(1 << 30), 1 << 30, npr)
    i64v = rng.random(npr) > 0.05
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk)], [kgo.Col(kgo.T_INT64, bk)])
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_AVG, 0), (kgo.AGG_SUM, 1), (kgo.AGG_COUNT_STAR, 0),
            (kgo.AGG_MAX, 0)]
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_INT64, 0, 0)]
    o = kgo.Agg((kgo.T_INT64, 0, 0), types, aggs, ansi=True)
    o.update(kgo.Col(kgo.T_INT64, bk[ob]),
             [kgo.Col(kgo.T_DECIMAL64, dec[op], decv[op], precision=12, scale=2),
              kgo.Col(kgo.T_INT64, i64[op], i64v[op])])
    rk, _, rres = o.finish()
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=nb)
    j.build_sink([Col(A.T_INT64, bk)])
    j.build_finalize()
    j.aggregate([(A.T_DECIMAL64, 12, 2), (A.T_INT64, 0, 0)], aggs, ansi=True)
    half = npr // 2 if path != "partitioned" else npr
    for a, b in ((0, half), (half, npr)):
        if b > a:
            j.probe_aggregate([Col(A.T_INT64, pk[a:b]),
                               Col(A.T_DECIMAL64, dec[a:b], decv[a:b], 12, 2),
                               Col(A.T_INT64, i64[a:b], i64v[a:b])], [1, 2])
    r = j.aggregate_result()
    assert r.num_rows == len(rk)
    gk = r.column(0)[0]
    gi = {int(k): i for i, k in enumerate(gk)}
    cols = [r.column(1 + c) for c in range(len(aggs))]
    for ri, k in enumerate(rk):
        i = gi[int(k)]
        for c in range(len(aggs)):
            rv, rok, _ = rres[c]
            gv, gok = cols[c]
            assert bool(rok[ri]) == bool(gok[i]), (k, c)
            if rok[ri]:
                assert int(rv[ri]) == int(gv[i]), (k, c, rv[ri], gv[i])


def test_fused_ansi_sum_that_could_overflow_is_refused():
    """An ANSI SUM whose batch could pass 2^63 (values near 2^61, four
    build rows a key) is refused before anything is applied; a batch that
    is proven safe is then taken as usual."""
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0])
    j.build_sink([Col(A.T_INT64, np.repeat(np.arange(10, dtype=np.int64), 4))])
    j.build_finalize()
    j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)], ansi=True)
    c = ctx()
    pk = kgpu.DeviceBuffer.from_numpy(c, np.arange(10, dtype=np.int64))
    big = kgpu.DeviceBuffer.from_numpy(c, np.full(10, 1 << 61, dtype=np.int64))
    small = kgpu.DeviceBuffer.from_numpy(c, np.arange(10, dtype=np.int64) * 3)
    with pytest.raises(kgpu.KgError, match="could overflow"):
        j.probe_aggregate([Col(A.T_INT64, pk, length=10), Col(A.T_INT64, big, length=10)], [1])
    j.probe_aggregate([Col(A.T_INT64, pk, length=10), Col(A.T_INT64, small, length=10)], [1])
    r = j.aggregate_result()
    got = dict(zip(r.column(0)[0].tolist(), zip(r.column(1)[0].tolist(), r.column(2)[0].tolist())))
    assert got == {k: (4 * 3 * k, 4) for k in range(10)}
    # a host batch is gathered: its refusal comes when the gathered rows are
    # taken (here by the result), and leaves them gathered
    j2 = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0])
    j2.build_sink([Col(A.T_INT64, np.repeat(np.arange(10, dtype=np.int64), 4))])
    j2.build_finalize()
    j2.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0)], ansi=True)
    j2.probe_aggregate([Col(A.T_INT64, np.arange(10)), Col(A.T_INT64, np.full(10, 1 << 61))], [1])
    with pytest.raises(kgpu.KgError, match="could overflow"):
        j2.aggregate_result()


def test_gathered_host_batches_build_and_probe():
    """Small host batches are gathered in HBM before the build and the fused
    aggregate take them: ragged build and probe batches at odd row and bit
    offsets, NULL keys and inputs, duplicate build keys, a device probe batch
    and a probe batch too large to gather between them; against the oracle,
    and a plain probe of the gathered build."""
    say("join: gathered host batches, build and fused probe")
    rng = np.random.default_rng(31)
    nb, npr = 300_000, 2_000_000
    bk = rng.integers(0, 100_000, nb).astype(np.int64)
    bkv = rng.random(nb) > 0.01
    pk = rng.integers(0, 120_000, npr).astype(np.int64)
    pkv = rng.random(npr) > 0.02
    amt = rng.integers(-(1 << 40), 1 << 40, npr)
    amtv = rng.random(npr) > 0.05
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk, pkv)],
                      [kgo.Col(kgo.T_INT64, bk, bkv)])
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0), (kgo.AGG_MIN, 0), (kgo.AGG_COUNT, 0)]
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT64, bk[ob]), [kgo.Col(kgo.T_INT64, amt[op], amtv[op])])
    rk, _, rres = o.finish()
    c = ctx()
    j = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0])
    bc = Col(A.T_INT64, bk, bkv, length=nb)
    lo = 0
    while lo < nb:
        m = min(int(rng.integers(1, 5000)), nb - lo)
        bc.offset = lo
        j.build_sink([bc], m)
        lo += m
    assert j.build_rows() == nb
    j.build_finalize()
    # a plain probe of the gathered build
    sp = pk[:5000]
    out = j.probe([Col(A.T_INT64, sp, pkv[:5000])])
    pp, _ = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, sp, pkv[:5000])],
                     [kgo.Col(kgo.T_INT64, bk, bkv)])
    assert out.num_rows == len(pp)
    j.aggregate([(A.T_INT64, 0, 0)], aggs)
    kc = Col(A.T_INT64, pk, pkv, length=npr)
    ac = Col(A.T_INT64, amt, amtv, length=npr)
    keep = []
    lo, b = 0, 0
    while lo < npr:
        m = min(int(rng.integers(1, 70_000)) if b != 7 else 600_000, npr - lo)
        if b == 3:
            d = [kgpu.DeviceBuffer.from_numpy(c, x) for x in
                 (pk[lo:lo + m], np.packbits(pkv[lo:lo + m], bitorder="little"),
                  amt[lo:lo + m], np.packbits(amtv[lo:lo + m], bitorder="little"))]
            keep += d
            j.probe_aggregate([Col(A.T_INT64, d[0], d[1], length=m),
                               Col(A.T_INT64, d[2], d[3], length=m)], [1], m)
        else:
            kc.offset = ac.offset = lo
            j.probe_aggregate([kc, ac], [1], m)
        lo += m
        b += 1
    r = j.aggregate_result()
    assert r.num_rows == len(rk)
    gk = r.column(0)[0]
    gi = {int(k): i for i, k in enumerate(gk)}
    cols = [r.column(1 + c_) for c_ in range(len(aggs))]
    for ri, k in enumerate(rk):
        i = gi[int(k)]
        for c_ in range(len(aggs)):
            rv, rok, _ = rres[c_]
            gv, gok = cols[c_]
            assert bool(rok[ri]) == bool(gok[i]), (k, c_)
            if rok[ri]:
                assert int(rv[ri]) == int(gv[i]), (k, c_, rv[ri], gv[i])


def test_empty_probe_and_no_match():
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0])
    j.build_sink([Col(A.T_INT64, np.array([1, 2, 3]))])
    j.build_finalize()
    assert j.probe([Col(A.T_INT64, np.array([], dtype=np.int64))]).num_rows == 0
    assert j.probe([Col(A.T_INT64, np.array([7, 8]))]).num_rows == 0


# build sizes: 128 and 2048 partitions (LDS-staged scatter), 4096 (direct
# scatter), 32768 (more than the partition kernels count: global atomics)
@pytest.mark.parametrize("nb_keys,nb_rows", [(5000, 20000), (300_000, 300_000),
                                             (600_000, 600_000), (4_000_000, 4_300_000)])
def test_partitioned_fused_aggregate(nb_keys, nb_rows):
    """Probe batches of a million rows and more take the radix-partitioned
    fused aggregate (LDS windows of the build table): duplicate build keys
    (the row count multiplies), NULL probe keys and NULL inputs, the key equal
    to the table's sentinel, a small table (one window) and larger ones."""
    say(f"join: partitioned fused aggregate, {nb_rows} build rows over {nb_keys} keys")
    rng = np.random.default_rng(nb_keys)
    bk = rng.integers(0, nb_keys, nb_rows).astype(np.int64)
    # at most ~300 build rows hold the sentinel key: its group gets ~1300 probe
    # rows times that many pairs, and the oracle's float sum over them in pair
    # order must stay within the tolerance of a sum in any other order
    bk[::max(997, nb_rows // 300)] = np.iinfo(np.int64).min
    npr = 1_300_000
    pk = rng.integers(0, int(nb_keys * 1.2), npr).astype(np.int64)
    pk[::1009] = np.iinfo(np.int64).min
    pkv = rng.random(npr) > 0.02
    amt = rng.integers(-(1 << 40), 1 << 40, npr)
    amtv = rng.random(npr) > 0.03
    f = rng.random(npr)
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk, pkv)], [kgo.Col(kgo.T_INT64, bk)])
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0), (kgo.AGG_MIN, 0), (kgo.AGG_MAX, 1),
            (kgo.AGG_AVG, 1), (kgo.AGG_COUNT, 0)]
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0), (kgo.T_FLOAT64, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT64, bk[ob]),
             [kgo.Col(kgo.T_INT64, amt[op], amtv[op]), kgo.Col(kgo.T_FLOAT64, f[op])])
    rk, _, rres = o.finish()
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=nb_rows)
    j.build_sink([Col(A.T_INT64, bk)])
    j.build_finalize()
    j.aggregate([(A.T_INT64, 0, 0), (A.T_FLOAT64, 0, 0)], [(a, c) for a, c in aggs])
    cut = 1_100_000  # one batch on the partitioned path, one on the direct path
    for a, b in ((0, cut), (cut, npr)):
        j.probe_aggregate([Col(A.T_INT64, pk[a:b], pkv[a:b]), Col(A.T_INT64, amt[a:b], amtv[a:b]),
                           Col(A.T_FLOAT64, f[a:b])], [1, 2])
    r = j.aggregate_result()
    assert r.num_rows == len(rk)
    gk = r.column(0)[0]
    order = np.argsort(gk, kind="stable")
    ro = np.argsort(np.asarray(rk), kind="stable")
    assert np.array_equal(np.asarray(gk)[order], np.asarray(rk)[ro])
    for c in range(len(aggs)):
        rv, rok, (rt, _, _) = rres[c]
        gv, gok = r.column(1 + c)
        rv, rok, gv, gok = np.asarray(rv)[ro], np.asarray(rok)[ro], gv[order], gok[order]
        assert np.array_equal(rok.astype(bool), gok.astype(bool)), c
        if rt == kgo.T_FLOAT64:
            a, b = rv[rok].astype(np.float64), gv[rok].astype(np.float64)
            assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(a), np.abs(b)) + 1e-300), c
        else:
            assert np.array_equal(rv[rok].astype(np.int64), gv[rok].astype(np.int64)), c


def _groups(keys, kvalid, cols):
    """{key or None: tuple of (value or None)} of an aggregate result."""
    out = {}
    for g in range(len(keys)):
        k = int(keys[g]) if kvalid[g] else None
        out[k] = tuple((v[g] if ok[g] else None) for v, ok in cols)
    return out


def test_partitioned_fused_aggregate_type_dispatch():
    """The partitioned fused aggregate on rows whose loads need the type
    dispatch: an int32 `<=>` key with NULLs on both sides (the NULL key's
    fixed slot, applied with atomics), int16 and float32 inputs with NULLs."""
    say("join: partitioned fused aggregate, int32 null-safe key, int16/float32 inputs")
    rng = np.random.default_rng(77)
    nb, npr = 40_000, 1_200_000
    bk = rng.integers(-30000, 30000, nb).astype(np.int32)
    bkv = rng.random(nb) > 0.001
    pk = rng.integers(-33000, 33000, npr).astype(np.int32)
    pkv = rng.random(npr) > 0.01
    a16 = rng.integers(-30000, 30000, npr).astype(np.int16)
    a16v = rng.random(npr) > 0.05
    f32 = rng.standard_normal(npr).astype(np.float32)
    f32v = rng.random(npr) > 0.05
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT32, pk, pkv)],
                      [kgo.Col(kgo.T_INT32, bk, bkv)], null_safe=[1])
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0), (kgo.AGG_MIN, 1), (kgo.AGG_MAX, 1),
            (kgo.AGG_SUM, 1), (kgo.AGG_AVG, 0), (kgo.AGG_COUNT, 1)]
    o = kgo.Agg((kgo.T_INT32, 0, 0), [(kgo.T_INT16, 0, 0), (kgo.T_FLOAT32, 0, 0)], aggs)
    o.update(kgo.Col(kgo.T_INT32, bk[ob], bkv[ob]),
             [kgo.Col(kgo.T_INT16, a16[op], a16v[op]), kgo.Col(kgo.T_FLOAT32, f32[op], f32v[op])])
    rk, rkv, rres = o.finish()
    want = _groups(rk, rkv, [(np.asarray(v), np.asarray(ok)) for v, ok, _ in rres])
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], [1], expected_build_rows=nb)
    j.build_sink([Col(A.T_INT32, bk, bkv)])
    j.build_finalize()
    j.aggregate([(A.T_INT16, 0, 0), (A.T_FLOAT32, 0, 0)], aggs)
    j.probe_aggregate([Col(A.T_INT32, pk, pkv), Col(A.T_INT16, a16, a16v),
                       Col(A.T_FLOAT32, f32, f32v)], [1, 2])
    r = j.aggregate_result()
    gk, gkv = r.column(0)
    got = _groups(gk, gkv, [r.column(1 + c) for c in range(len(aggs))])
    assert None in want and set(got) == set(want)
    for k, wv in want.items():
        for c, (a, b) in enumerate(zip(wv, got[k])):
            assert (a is None) == (b is None), (k, c)
            if a is None:
                continue
            if isinstance(a, (float, np.floating)):
                assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)) + 1e-300, (k, c, a, b)
            else:
                assert int(a) == int(b), (k, c, a, b)


# ---------------------------------------------------------------- every join type
ALL_TYPES = [A.JOIN_INNER, A.JOIN_LEFT_OUTER, A.JOIN_LEFT_SEMI, A.JOIN_LEFT_ANTI,
             A.JOIN_NULL_AWARE_ANTI, A.JOIN_EXISTENCE, A.JOIN_RIGHT_SEMI, A.JOIN_RIGHT_ANTI,
             A.JOIN_RIGHT_OUTER, A.JOIN_FULL_OUTER]
# the reference's VLess residual filter (tests/hash_join/main.rs:72-91) over
# column 2 of each side: (p + b) % 7 != 0 AND p < b
VLESS = [(A.X_PROBE, 2, 0), (A.X_BUILD, 2, 0), (A.X_ADD, 0, 0), (A.X_CONST, 0, 7),
         (A.X_MOD, 0, 0), (A.X_CONST, 0, 0), (A.X_NE, 0, 0), (A.X_PROBE, 2, 0),
         (A.X_BUILD, 2, 0), (A.X_LT, 0, 0), (A.X_AND, 0, 0)]
NULLKEY = None


def _side(rows, kdt):
    """Rows (k1, k2, v) -> columns k1, k2 (nullable, of dtype kdt) and v (int64)."""
    n = len(rows)
    k1 = np.array([0 if r[0] is None else r[0] for r in rows], dtype=kdt)
    k2 = np.array([0 if r[1] is None else r[1] for r in rows], dtype=kdt)
    v = np.array([r[2] for r in rows], dtype=np.int64)
    ok1 = np.array([r[0] is not None for r in rows], dtype=bool)
    ok2 = np.array([r[1] is not None for r in rows], dtype=bool)
    return (k1, ok1), (k2, ok2), (v, np.ones(n, dtype=bool))


def expected_any(jt, pairs, probe, build):
    """Output rows of kgo.join_ex's pairs, as the device returns them."""
    op, ob = pairs
    out = []
    for i, j in zip(op.tolist(), ob.tolist()):
        pr = [None] * len(probe) if i < 0 else [int(v[i]) if ok[i] else None for v, ok in probe]
        br = [] if jt == A.JOIN_EXISTENCE else \
            [None] * len(build) if j < 0 else [int(v[j]) if ok[j] else None for v, ok in build]
        if jt in (A.JOIN_LEFT_SEMI, A.JOIN_LEFT_ANTI, A.JOIN_NULL_AWARE_ANTI):
            out.append(tuple(pr))
        elif jt == A.JOIN_EXISTENCE:
            out.append(tuple(pr) + (bool(j),))
        elif jt in (A.JOIN_RIGHT_SEMI, A.JOIN_RIGHT_ANTI):
            out.append(tuple(br))
        else:
            out.append(tuple(pr) + tuple(br))
    return sorted(out, key=lambda t: tuple((x is None, x or 0) for x in t))


def batch_rows(batches):
    out = []
    for b in batches:
        if b.num_rows == 0:
            continue
        cols = [b.column(i) for i in range(b.num_columns)]
        for r in range(b.num_rows):
            out.append(tuple((bool(v[r]) if v.dtype == bool else int(v[r])) if ok[r] else None
                             for v, ok in cols))
    return sorted(out, key=lambda t: tuple((x is None, x or 0) for x in t))


def run_any(jt, keys, safe, filt, probe_rows, build_rows, max_rows, build_batches=2,
            probe_batches=2, direct=False):
    # one key: int64 (the untyped 8-byte fast path); two keys: int32 + int32,
    # packed into 8 bytes
    kdt, kt, okt = (np.int64, A.T_INT64, kgo.T_INT64) if keys == 1 else \
        (np.int32, A.T_INT32, kgo.T_INT32)
    types = [kt, kt, A.T_INT64]
    P, B = _side(probe_rows, kdt), _side(build_rows, kdt)
    j = kgpu.HashJoin(ctx(), jt, list(range(keys)), list(range(keys)), safe[:keys],
                      expected_build_rows=64, direct_index=direct)
    nb = len(build_rows)
    if nb == 0:  # an empty build still gives its schema
        j.build_sink([Col(t, v, ok) for t, (v, ok) in zip(types, B)], 0)
    cuts = np.linspace(0, nb, build_batches + 1).astype(int)
    for a, b in zip(cuts[:-1], cuts[1:]):
        if b > a:
            j.build_sink([Col(t, v[a:b], ok[a:b]) for t, (v, ok) in zip(types, B)], b - a)
    j.build_finalize()
    if filt:
        j.set_filter(VLESS)
    loc = j.local()
    got = []
    npr = len(probe_rows)
    cuts = np.linspace(0, npr, probe_batches + 1).astype(int)
    for a, b in zip(cuts[:-1], cuts[1:]):
        if b > a:
            for out in loc.probe_all([Col(t, v[a:b], ok[a:b]) for t, (v, ok) in zip(types, P)],
                                     max_rows, b - a):
                assert out.num_rows <= max_rows
                got.append(out)
    got += j.tail([(t, 0, 0) for t in types], max_rows)
    kc = lambda side: [kgo.Col(t, v, ok) for t, (v, ok) in zip([okt, okt, kgo.T_INT64], side)]
    pairs = kgo.join_ex(jt, kc(P)[:keys], kc(B)[:keys], safe[:keys],
                        filt=VLESS if filt else None, probe_cols=kc(P), build_cols=kc(B))
    return batch_rows(got), expected_any(jt, pairs, P, B)


A_, B_, X_ = 1, 2, 3
EDGE_PROBE = [(1, A_, 1), (1, A_, 9), (2, None, 3), (None, A_, 4), (None, None, 5), (7, X_, 6)]
EDGE_BUILD = [(1, A_, 8), (1, A_, 2), (1, B_, 3), (2, None, 12), (None, None, 10), (None, A_, 11)]
ALL_NULL = [(None, None, 1), (None, None, 2)]


@pytest.mark.parametrize("jt", ALL_TYPES)
def test_every_join_type_on_the_reference_edge_cases(jt):
    """The reference's edge cases (tests/hash_join/main.rs:232-290): NULL keys,
    duplicates on both sides, empty and all-NULL sides, one and two keys,
    `=` and `<=>`, with and without the residual filter, output bounded at
    5 rows a batch (the reference's tiny batch size)."""
    say(f"join types: edge cases, type {jt}")
    inputs = [(EDGE_PROBE, EDGE_BUILD), (EDGE_PROBE, []), ([], EDGE_BUILD),
              (EDGE_PROBE, ALL_NULL), (ALL_NULL, EDGE_BUILD)]
    for keys in (1, 2):
        for s in (False, True):
            if keys == 2 and s:
                continue  # two null-safe keys are refused on the device
            for filt in (False, True):
                if jt == A.JOIN_NULL_AWARE_ANTI and (keys > 1 or s or filt):
                    continue
                for p, b in inputs:
                    got, want = run_any(jt, keys, [s, s], filt, p, b, 5)
                    assert got == want, (jt, keys, s, filt, len(p), len(b))


def test_not_in_spark_null_semantics_on_device():
    """tests/hash_join/main.rs:300-321 on the device."""
    probe = [(1, None, 0), (2, None, 0), (None, None, 0)]
    jt = A.JOIN_NULL_AWARE_ANTI
    assert run_any(jt, 1, [False], False, probe, [], 7)[0] == [(1, None, 0), (2, None, 0),
                                                                (None, None, 0)]
    assert run_any(jt, 1, [False], False, probe, [(1, None, 0), (None, None, 0)], 7)[0] == []
    assert run_any(jt, 1, [False], False, probe, [(1, None, 0)], 7)[0] == [(2, None, 0)]


@pytest.mark.parametrize("jt", ALL_TYPES)
def test_every_join_type_random_larger(jt):
    """Random rows with NULLs and duplicates (keys in 0..2000, values in
    0..20), one or two keys, `=`/`<=>`, with and without the filter, output
    bounded at 777 rows a batch; the direct index on for one-key builds."""
    rng = np.random.default_rng(100 + jt)

    def rows(n, nk):
        return [(None if rng.random() < 0.05 else int(rng.integers(0, nk)),
                 None if rng.random() < 0.05 else int(rng.integers(0, 3)), int(rng.integers(0, 20)))
                for _ in range(n)]

    for case in range(4):
        naaj = jt == A.JOIN_NULL_AWARE_ANTI
        keys = 1 if naaj or case % 2 == 0 else 2
        safe = [False, False] if naaj or keys == 2 else [bool(case & 2)] * 2
        filt = (not naaj) and case >= 2
        say(f"join types: random, type {jt}, case {case}")
        probe, build = rows(6000, 2000), rows(3000, 2000)
        if naaj and case == 3:
            build = [r for r in build if r[0] is not None]  # a NOT IN build without NULLs
        got, want = run_any(jt, keys, safe, filt, probe, build, 777, build_batches=3,
                            probe_batches=3, direct=(keys == 1 and case == 1))
        assert got == want, (jt, case)


def test_locals_probe_concurrently_from_threads():
    """Two drivers' locals over one finalized build, probing different
    batches from two threads at once (ctypes releases the GIL during the
    calls): each gets the rows it gets alone."""
    import threading
    rng = np.random.default_rng(9)
    bk = rng.integers(0, 50_000, 200_000)
    pay = np.arange(len(bk), dtype=np.int64)
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=len(bk))
    j.build_sink([Col(A.T_INT64, bk), Col(A.T_INT64, pay)])
    j.build_finalize()
    batches = [rng.integers(0, 60_000, 400_000) for _ in range(6)]
    alone = []
    for pk in batches:
        out = j.local().probe_all([Col(A.T_INT64, pk)], 1 << 30)[0]
        alone.append(np.sort(out.column(2)[0] * 100_000 + out.column(0)[0]))
    got = [None] * len(batches)
    err = []

    def drive(which):
        try:
            loc = j.local()
            for k in which:
                outs = loc.probe_all([Col(A.T_INT64, batches[k])], 50_000)
                v = np.concatenate([o.column(2)[0] * 100_000 + o.column(0)[0] for o in outs])
                got[k] = np.sort(v)
        except Exception as e:  # noqa: BLE001
            err.append(e)

    ts = [threading.Thread(target=drive, args=(w,)) for w in ([0, 2, 4], [1, 3, 5])]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not err, err
    for a, b in zip(alone, got):
        assert np.array_equal(a, b)
