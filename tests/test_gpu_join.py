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
    assert probed_direct(lambda: j.probe_aggregate(
        [Col(A.T_INT32, pk, pkv), Col(A.T_INT64, amt, amtv), Col(A.T_FLOAT64, f)], [1, 2]))
    r = j.aggregate_result()
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
