"""GPU parity at the benchmarked sizes: each workload the benchmark times,
run through the C ABI at its full configured shape and checked against the
oracle on the same seeded inputs (generated in HBM, then copied to the host
for the oracle).

  configuration 1: 1M-row permuted build, 8M-row probe (1 % NULL fk), the
                   fused join aggregate and the unfused join + aggregate
  configuration 2: 200M rows, int64 key Zipf(1.1) over 10M and the int8
                   4-value key, COUNT(*) + SUM/MIN/MAX/AVG of decimal(12,2),
                   float64, int32; integer and decimal results bit-exact,
                   float64 SUM/AVG within 1e-12 of a compensated sum
  configuration 3: the 15M-row build's filter bytes, the survivors of 60M
                   probe rows and their join SUM/COUNT
  configuration 4: 1/100 of the grace join (15M build x 60M probe rows,
                   Zipf(0.8) over 12M keys) against the per-key factorization
"""
import math
import os

import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col, DeviceBuffer
from kgpu import abi as A

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
REL = 1e-12  # float64 SUM/AVG: their value depends on summation order


def host(buf: DeviceBuffer, dtype, n):
    return buf.to_numpy(dtype, n)


def compare_groups(gpu_batch, ref, ncols, float_cols=(), what=""):
    """GPU result batch [key, agg...] against the oracle's finish(raw=True):
    same key set, integer/decimal columns and validity bit-exact, float
    columns within REL (or both NaN/NULL)."""
    rk, rkv, rres = ref
    gk, gkv = gpu_batch.column_raw(0)
    assert len(gk) == len(rk), (what, len(gk), len(rk))
    go = np.lexsort((gk, ~gkv))
    ro = np.lexsort((rk, ~rkv))
    assert np.array_equal(gk[go], rk[ro]) and np.array_equal(gkv[go], rkv[ro]), what
    for j in range(ncols):
        gv, gok = gpu_batch.column_raw(1 + j)
        rv, rok, _ = rres[j]
        gv, gok = np.asarray(gv)[go], gok[go]
        rv, rok = np.asarray(rv)[ro], rok[ro]
        assert np.array_equal(gok, rok), (what, j)
        if j in float_cols:
            a, b = gv[gok], rv[rok]
            both_nan = np.isnan(a) & np.isnan(b)
            err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            assert np.all(both_nan | (a == b) | (err <= REL)), (what, j, float(np.max(err)))
        else:
            gv = np.where(gok[:, None] if gv.ndim == 2 else gok, gv, 0)
            rv = np.where(rok[:, None] if rv.ndim == 2 else rok, rv, 0)
            assert np.array_equal(gv, rv), (what, j)


# ---------------------------------------------------------------- configuration 1
@pytest.mark.parametrize("fused", [True, False])
def test_config1_join_sum_count_exact(fused):
    """SELECT key, SUM(amt), COUNT(*) FROM build JOIN probe ON key = fk GROUP
    BY key at exactly 1M x 8M rows (fk uniform in [0, 1.2M), 1 % NULL)."""
    say(f"config 1: 1M x 8M, fused={fused}")
    c = ctx()
    L = kgpu.lib()
    nb, npr = 1_000_000, 8_000_000
    bk, bv = DeviceBuffer(c, nb * 8), DeviceBuffer(c, nb * 8)
    A.check(L.kg_gen_join_build(c.h, 1, 0, nb, nb, bk.ptr, bv.ptr))
    fk, fkv, amt = DeviceBuffer(c, npr * 8), DeviceBuffer(c, npr // 8 + 8), DeviceBuffer(c, npr * 8)
    A.check(L.kg_gen_join_probe(c.h, 2, 0, npr, 1_200_000, 10, fk.ptr, fkv.ptr, amt.ptr))
    probe = [Col(A.T_INT64, fk, fkv, length=npr), Col(A.T_INT64, amt, length=npr)]
    j = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0], expected_build_rows=nb)
    j.build_sink([Col(A.T_INT64, bk, length=nb), Col(A.T_INT64, bv, length=nb)], nb)
    j.build_finalize()
    if fused:
        j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
        j.probe_aggregate(probe, [1], npr)
        out = j.aggregate_result()
    else:
        pairs = j.probe(probe, npr)
        g = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), [(A.T_INT64, 0, 0)],
                               [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)], expected_groups=nb)
        g.sink(pairs.as_col(2), [pairs.as_col(1)], pairs.num_rows)
        g.finalize()
        out = g.result()
    # the oracle: the same rows on the host, joined and aggregated
    hbk = host(bk, np.int64, nb)
    hfk, hamt = host(fk, np.int64, npr), host(amt, np.int64, npr)
    hok = np.unpackbits(host(fkv, np.uint8, npr // 8), bitorder="little").astype(bool)
    want_k, want_ok, want_amt = kgo.gen_join_probe(2, 0, 4096, 1_200_000, 10)
    assert np.array_equal(hfk[:4096][want_ok], want_k[want_ok])  # the generators agree
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, hfk, hok)],
                      [kgo.Col(kgo.T_INT64, hbk)])
    o = kgo.AggPar((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                   [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)], THREADS)
    o.update(kgo.Col(kgo.T_INT64, hbk[ob]), [kgo.Col(kgo.T_INT64, hamt[op])])
    ref = o.finish(raw=True)
    assert 990_000 < len(ref[0]) < 1_000_000
    compare_groups(out, ref, 2, what="config 1")


# ---------------------------------------------------------------- configuration 2
AGGS2 = [(A.AGG_COUNT_STAR, 0)] + [(f, c) for c in range(3)
                                   for f in (A.AGG_SUM, A.AGG_MIN, A.AGG_MAX, A.AGG_AVG)]
TYPES2 = [(A.T_DECIMAL64, 12, 2), (A.T_FLOAT64, 0, 0), (A.T_INT32, 0, 0)]
FLOAT_COLS2 = (5, 8, 12)  # SUM/AVG of the float64 column, AVG of int32 (after the key)


def gen_config2(n, nd):
    c = ctx()
    L = kgpu.lib()
    cdf = np.zeros(nd)
    A.check(L.kg_zipf_cdf(nd, 1.1, cdf.ctypes.data))
    cdf_d = DeviceBuffer.from_numpy(c, cdf)
    bufs = {"key": DeviceBuffer(c, n * 8), "k8": DeviceBuffer(c, n), "dec": DeviceBuffer(c, n * 8),
            "f": DeviceBuffer(c, n * 8), "i32": DeviceBuffer(c, n * 4)}
    A.check(L.kg_gen_agg(c.h, 3, 0, n, cdf_d.ptr, nd, bufs["key"].ptr, bufs["dec"].ptr,
                         bufs["f"].ptr, bufs["i32"].ptr, bufs["k8"].ptr))
    c.synchronize()
    return bufs


def run_config2(bufs, n, key_type, batches, expected):
    c = ctx()
    kbuf = bufs["key"] if key_type == A.T_INT64 else bufs["k8"]
    w = 8 if key_type == A.T_INT64 else 1
    g = kgpu.HashAggregate(c, (key_type, 0, 0), TYPES2, AGGS2, expected_groups=expected)
    step = n // batches
    for a in range(0, n, step):
        m = min(step, n - a)
        g.sink(Col(key_type, kbuf, length=m, offset=a),
               [Col(A.T_DECIMAL64, bufs["dec"], precision=12, scale=2, length=m, offset=a),
                Col(A.T_FLOAT64, bufs["f"], length=m, offset=a),
                Col(A.T_INT32, bufs["i32"], length=m, offset=a)], m)
    g.finalize()
    return g.result()


def oracle_config2(bufs, n, key_type):
    kt = kgo.T_INT64 if key_type == A.T_INT64 else kgo.T_INT8
    key = host(bufs["key"] if key_type == A.T_INT64 else bufs["k8"],
               np.int64 if key_type == A.T_INT64 else np.int8, n)
    dec, f, i32 = host(bufs["dec"], np.int64, n), host(bufs["f"], np.float64, n), \
        host(bufs["i32"], np.int32, n)
    o = kgo.AggPar((kt, 0, 0), [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0),
                                (kgo.T_INT32, 0, 0)], AGGS2, THREADS, exact_f64=True)
    o.update(kgo.Col(kt, key), [kgo.Col(kgo.T_DECIMAL64, dec, precision=12, scale=2),
                                kgo.Col(kgo.T_FLOAT64, f), kgo.Col(kgo.T_INT32, i32)])
    return o.finish(raw=True)


def test_config2_200M_rows_int64_zipf_and_int8_keys():
    """The benchmark's aggregation at its full 200M rows: the int64 Zipf(1.1)
    key over 10M values (about 6.7M groups, a top key of 11.6 % of the rows:
    the LDS pass with its eviction windows, the spill drain and its partition
    passes, and table growth from a small expected group count) and the int8
    key with 4 values (the direct path)."""
    n, nd = 200_000_000, 10_000_000
    say("config 2: generating 200M rows")
    bufs = gen_config2(n, nd)
    for key_type, batches, expected in ((A.T_INT64, 4, 1 << 20), (A.T_INT8, 1, 4)):
        say(f"config 2: key type {key_type}, {batches} sink batches, expected {expected}")
        out = run_config2(bufs, n, key_type, batches, expected)
        ref = oracle_config2(bufs, n, key_type)
        if key_type == A.T_INT64:
            assert 6_600_000 < len(ref[0]) < 6_800_000
        else:
            assert len(ref[0]) == 4
        compare_groups(out, ref, len(AGGS2), FLOAT_COLS2, f"config 2 key {key_type}")
        del out


def test_pass1_eviction_with_null_and_sentinel_keys():
    """20M rows over 3,000 Zipf(1.1) keys with NULL keys and INT64_MIN (the
    table's empty marker) mixed in, one batch: every pass-1 workgroup runs
    past its 64K-row eviction windows (cold records move out, a key can take
    a second LDS slot after a hole), checked bit-exact."""
    n, nd = 20_000_000, 3_000
    say("agg: eviction windows, 20M rows over 3,000 keys")
    c = ctx()
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(11, 0, n, cdf, nd)
    key = key - 1500
    key[::1013] = np.iinfo(np.int64).min
    kv = np.ones(n, dtype=bool)
    kv[::997] = False
    cols = [Col(A.T_DECIMAL64, dec, precision=12, scale=2), Col(A.T_FLOAT64, f),
            Col(A.T_INT32, i32)]
    g = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), TYPES2, AGGS2, expected_groups=nd)
    g.sink(Col(A.T_INT64, key, kv), cols, n)
    g.finalize()
    out = g.result()
    o = kgo.AggPar((kgo.T_INT64, 0, 0), [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0),
                                         (kgo.T_INT32, 0, 0)], AGGS2, THREADS, exact_f64=True)
    o.update(kgo.Col(kgo.T_INT64, key, kv), [kgo.Col(kgo.T_DECIMAL64, dec, precision=12, scale=2),
                                             kgo.Col(kgo.T_FLOAT64, f), kgo.Col(kgo.T_INT32, i32)])
    ref = o.finish(raw=True)
    assert len(ref[0]) == nd + 2
    compare_groups(out, ref, len(AGGS2), FLOAT_COLS2, "eviction")


# ---------------------------------------------------------------- configuration 3
def sbbf_passes(words, hashes):
    nblk = len(words) // 8
    blocks = ((hashes >> np.uint64(32)) * np.uint64(nblk)) >> np.uint64(32)
    salt = np.array([0x47b6137b, 0x44974d91, 0x8824ad5b, 0xa2b7289d, 0x705495c7, 0x2df1424b,
                     0x9efc4947, 0x5c6bfb31], dtype=np.uint64)
    key32 = hashes & np.uint64(0xFFFFFFFF)
    bits = ((key32[:, None] * salt[None, :]) & np.uint64(0xFFFFFFFF)) >> np.uint64(27)
    wv = words.reshape(-1, 8)[blocks.astype(np.int64)].astype(np.uint64)
    return np.all((wv >> bits) & np.uint64(1), axis=1)


def test_config3_filter_bytes_survivors_and_join_at_60M():
    """The 15M-row build (keys a permutation, the predicate keeping 5 %), its
    split-block filter sized for 750K items at fpp 0.03 (723,968 bytes)
    byte-identical to the oracle's; 60M probe rows screened (survivors equal
    the oracle's check, in order) and joined, SUM(amt) and COUNT(*) exact."""
    say("config 3: 15M build, 60M probe")
    c = ctx()
    L = kgpu.lib()
    nb, npr = 15_000_000, 60_000_000
    bk, bv = DeviceBuffer(c, nb * 8), DeviceBuffer(c, nb * 8)
    A.check(L.kg_gen_join_build(c.h, 4, 0, nb, nb, bk.ptr, bv.ptr))
    fk, amt = DeviceBuffer(c, npr * 8), DeviceBuffer(c, npr * 8)
    fkv = DeviceBuffer(c, npr // 8 + 8)
    A.check(L.kg_gen_join_probe(c.h, 5, 0, npr, nb, 0, fk.ptr, fkv.ptr, amt.ptr))
    rows, kept = kgpu.select_range(c, Col(A.T_INT64, bv, length=nb), 0, 49_999)
    keys = kgpu.gather(c, [Col(A.T_INT64, bk, length=nb)], rows, kept)
    nbytes = kgpu.RuntimeBloom.bytes_for(750_000, 0.03)
    assert nbytes == 723_968
    b = kgpu.RuntimeBloom(c, nbytes)
    b.insert(keys.as_col(0), nrows=kept)
    hbk, hbv = host(bk, np.int64, nb), host(bv, np.int64, nb)
    kept_keys = hbk[hbv < 50_000]
    assert kept == len(kept_keys) and 700_000 < kept < 800_000
    assert np.array_equal(keys.column_raw(0)[0], kept_keys)
    ok, words = kgo.sbbf_build(kgo.hash_columns([kgo.Col(kgo.T_INT64, kept_keys)]), nbytes=nbytes)
    assert ok and b.to_bytes() == words.tobytes()
    j = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0], expected_build_rows=kept)
    j.build_sink([keys.as_col(0)], kept)
    j.build_finalize()
    probe = [Col(A.T_INT64, fk, length=npr), Col(A.T_INT64, amt, length=npr)]
    surv, ns = b.probe(probe[0])
    got = host(surv, np.int32, ns)
    out = j.probe_selected(probe, surv, ns, npr)
    g = kgpu.HashAggregate(c, None, [(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
    g.sink(None, [out.as_col(1)], out.num_rows)
    g.finalize()
    r = g.result()
    hfk, hamt = host(fk, np.int64, npr), host(amt, np.int64, npr)
    want = []
    for a in range(0, npr, 5_000_000):
        h = kgo.hash_columns([kgo.Col(kgo.T_INT64, hfk[a:a + 5_000_000])])
        want.append(np.nonzero(sbbf_passes(words, h))[0] + a)
    want = np.concatenate(want)
    assert np.array_equal(got, want.astype(np.int32))
    is_kept = np.zeros(nb, dtype=bool)
    is_kept[kept_keys] = True
    hit = is_kept[hfk]
    assert np.isin(np.nonzero(hit)[0], got).all()  # no false negatives
    assert int(r.column_raw(0)[0][0]) == int(hamt[hit].sum())
    assert int(r.column_raw(1)[0][0]) == int(hit.sum())


# ---------------------------------------------------------------- configuration 4
def test_config4_grace_at_one_hundredth():
    """Configuration 4 at 1/100 (15M build x 60M probe rows, int64 keys
    Zipf(0.8) over 12M values, payloads in [1, 1e4], 64 partitions, an 80 MB
    table budget): per-partition COUNT and SUM bit-exact against the
    per-key factorization, which the oracle tests pin to the pair walk."""
    say("config 4: 1/100 scale, 15M x 60M")
    c = ctx()
    L = kgpu.lib()
    nb, npr, nd = 15_000_000, 60_000_000, 12_000_000
    cdf = np.zeros(nd)
    A.check(L.kg_zipf_cdf(nd, 0.8, cdf.ctypes.data))
    cdf_d = DeviceBuffer.from_numpy(c, cdf)
    del cdf
    bk, bp = DeviceBuffer(c, nb * 8), DeviceBuffer(c, nb * 8)
    pk, pp = DeviceBuffer(c, npr * 8), DeviceBuffer(c, npr * 8)
    # one key distribution on both sides (permutation seed 6)
    A.check(L.kg_gen_zipf_keys(c.h, 6, 6, 0, nb, cdf_d.ptr, nd, bk.ptr))
    A.check(L.kg_gen_zipf_keys(c.h, 7, 6, 0, npr, cdf_d.ptr, nd, pk.ptr))
    scratch = DeviceBuffer(c, npr * 8)
    A.check(L.kg_gen_join_probe(c.h, 8, 0, nb, 1, 0, scratch.ptr, None, bp.ptr))
    A.check(L.kg_gen_join_probe(c.h, 9, 0, npr, 1, 0, scratch.ptr, None, pp.ptr))
    del scratch
    g = kgpu.GraceJoinAggregate(c, 64, table_budget=80 << 20)
    batch = 1 << 24
    for a in range(0, nb, batch):
        m = min(batch, nb - a)
        g.build_sink(Col(A.T_INT64, bk, length=m, offset=a), Col(A.T_INT64, bp, length=m, offset=a), m)
    for a in range(0, npr, batch):
        m = min(batch, npr - a)
        g.probe_sink(Col(A.T_INT64, pk, length=m, offset=a), Col(A.T_INT64, pp, length=m, offset=a), m)
    gc, gs, gv, _ = g.finish()
    g.close()
    hb = [kgo.Col(kgo.T_INT64, host(x, np.int64, nb)) for x in (bk, bp)]
    hp = [kgo.Col(kgo.T_INT64, host(x, np.int64, npr)) for x in (pk, pp)]
    wc, ws, wv = kgo.grace_factorized(hb[0], hb[1], hp[0], hp[1], 64, 0)
    assert wc.sum() > 10 ** 11  # about 1.3e11 pairs at this scale
    assert np.array_equal(hb[0].values[:1000], kgo.gen_zipf_keys(6, 6, 0, 1000, kgo.zipf_cdf(nd, 0.8), nd))
    assert np.array_equal(gc, wc) and np.array_equal(gs, ws) and np.array_equal(gv, wv)
