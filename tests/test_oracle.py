"""Pins the CPU oracle (oracle/kgo.c) against the reference's golden vectors,
independent implementations (python xxhash, the Parquet SBBF specification,
exact Python integer / Fraction arithmetic, math.fsum) and brute force.
CPU only."""
import json
import math
import os
import random
import struct
from fractions import Fraction

import numpy as np
import pytest

import kgo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "vectors.json")) as _f:
    V = json.load(_f)
R = V["routing"]
U64 = (1 << 64) - 1


def setup_module():
    kgo.build()


# ---------------------------------------------------------------- XXH64
def test_xxh64_reference_vectors():
    for s, h in R["xxh64"].items():
        assert kgo.xxh64(s.encode()) == h


def test_xxh64_general_matches_python_xxhash():
    for c in V["xxh64_bytes"]:
        b = bytes.fromhex(c["hex"])
        assert kgo.xxh64(b, 0) == c["seed0"]
        assert kgo.xxh64(b, 42) == c["seed42"]


def test_xxh64_fixed_widths_match_xxhash():
    for c in V["xxh64_fixed"]:
        v = int(c["v"])
        assert kgo.xxh64_fixed(v & 0xFF, 1) == c["w1"]
        assert kgo.xxh64_fixed(v & 0xFFFF, 2) == c["w2"]
        assert kgo.xxh64_fixed(v & 0xFFFFFFFF, 4) == c["w4"]
        assert kgo.xxh64_fixed(v & U64, 8) == c["w8"]
        assert kgo.xxh64_fixed(v, 16) == c["w16"]


# ---------------------------------------------------------------- H_kg
def test_hkg_int64_goldens_and_null():
    vals = R["int64"]["values"]
    col = kgo.Col(kgo.T_INT64, vals + [0], valid=[True] * len(vals) + [False])
    h = kgo.hash_columns([col])
    assert list(h[:4]) == R["int64"]["hashes"]
    assert int(h[4]) == R["null_hash"]


def test_hkg_int32_float64_decimal_goldens():
    h = kgo.hash_columns([kgo.Col(kgo.T_INT32, R["int32"]["values"])])
    assert list(h) == R["int32"]["hashes"]
    h = kgo.hash_columns([kgo.Col(kgo.T_DATE32, [7])])
    assert int(h[0]) == R["int32"]["hashes"][0], "a date hashes as its days"
    f = [float(x) for x in R["float64"]["values"]]
    f[3] = -f[3]  # -NaN
    h = kgo.hash_columns([kgo.Col(kgo.T_FLOAT64, f)])
    assert list(h) == R["float64"]["hashes"]
    d = R["decimal128"]
    h = kgo.hash_columns([kgo.Col(kgo.T_DECIMAL128, d["values"], precision=20, scale=2)])
    assert list(h) == d["hashes"]
    # Decimal64 carrier hashes as the same 16-byte unscaled value
    h64 = kgo.hash_columns([kgo.Col(kgo.T_DECIMAL64, d["values"], precision=18, scale=2)])
    assert list(h64) == d["hashes"]
    # int8 -1 and bools: one byte
    assert int(kgo.hash_columns([kgo.Col(kgo.T_INT8, [-1])])[0]) == kgo.xxh64(b"\xff")
    hb = kgo.hash_columns([kgo.Col(kgo.T_BOOL, [False, True])])
    assert list(hb) == [kgo.xxh64(b"\x00"), kgo.xxh64(b"\x01")]


def test_hkg_fold_goldens():
    assert kgo.lib().kgo_fold_column(0, 0) == 0
    assert kgo.lib().kgo_fold_column(0x243F6A8885A308D3, 0) == R["fold_init_0"]
    assert kgo.xxh64(R["string"]["value"].encode()) == R["string"]["hash"]
    t = R["two_columns"]
    ha = kgo.hash_columns([kgo.Col(kgo.T_INT64, t["int64"])])
    for i, s in enumerate(t["string"]):
        hb = kgo.xxh64(s.encode())
        h = kgo.lib().kgo_fold_column(kgo.lib().kgo_fold_column(0x243F6A8885A308D3, int(ha[i])), hb)
        assert h == t["hashes"][i]


def test_hkg_multi_column_is_ordered_and_distinct():
    a = kgo.Col(kgo.T_INT64, np.arange(1000))
    b = kgo.Col(kgo.T_INT64, np.arange(1000)[::-1].copy())
    ab, ba = kgo.hash_columns([a, b]), kgo.hash_columns([b, a])
    assert np.sum(ab == ba) <= 1
    aa = kgo.hash_columns([a, a])
    assert len(np.unique(aa)) == 1000


# ---------------------------------------------------------------- partition_of
def test_partition_of_invariants():
    rng = random.Random(7)
    L = kgo.lib()
    for _ in range(3000):
        h = rng.getrandbits(64)
        n = rng.randint(1, 4096)
        p = L.kgo_partition_of(h, n)
        assert p < n
        s, e = kgo.key_group_range(p, n)
        assert s <= (h >> 52) < e
        k = rng.randint(0, 20)
        assert L.kgo_partition_of(h, 1 << k) == L.kgo_radix(h, k)
        r0 = rng.randint(0, 8)
        if r0 <= k:
            assert L.kgo_radix(h, r0) == L.kgo_partition_of(h, 1 << k) >> (k - r0)
        big = L.kgo_partition_of(h, 1 << 20)
        assert big < (1 << 20)
        # above 4096 partitions: (h * n) >> 64
        n2 = rng.randint(4097, 1 << 30)
        assert L.kgo_partition_of(h, n2) == (h * n2) >> 64
    assert L.kgo_partition_of(U64, 0) == 0


def test_key_group_ranges_tile():
    for n in list(range(1, 65)) + [100, 1000, 4095, 4096]:
        at = 0
        for p in range(n):
            s, e = kgo.key_group_range(p, n)
            assert s == at and e > s
            at = e
        assert at == 4096
        assert kgo.key_group_range(n, n) == (0, 0)


def test_route_and_split():
    rng = np.random.default_rng(3)
    keys = rng.integers(-1000, 1000, 5000)
    valid = rng.random(5000) > 0.1
    col = kgo.Col(kgo.T_INT64, keys, valid)
    pids = kgo.route([col], 37)
    h = kgo.hash_columns([col])
    assert all(int(p) == kgo.lib().kgo_partition_of(int(x), 37) for p, x in zip(pids, h))
    offs, rows = kgo.split(pids, 37)
    for p in range(37):
        r = rows[offs[p]:offs[p + 1]]
        assert np.all(np.diff(r) > 0)
        assert np.all(pids[r] == p)
    assert offs[-1] == 5000


# ---------------------------------------------------------------- row hash h / g
def test_row_hash_known_answer():
    f = kgo.Col(kgo.T_FLOAT64, [0.0])
    i = kgo.Col(kgo.T_INT32, [0], valid=[False])
    assert int(kgo.row_hash([f, i])[0]) == V["row_hash"]["h"]
    # a single int64 key: h(v) = fold(fold(v, MIX), FINISH), h(0) = 0
    L = kgo.lib()
    for v in (0, 1, -5, 1 << 50):
        want = L.kgo_fold(L.kgo_fold(v & U64, 0x243F6A8885A308D3), 0xB7E151628AED2A6B)
        assert int(kgo.row_hash([kgo.Col(kgo.T_INT64, [v])])[0]) == want
    # -0.0 and 0.0 hash alike
    g = kgo.row_hash([kgo.Col(kgo.T_FLOAT64, [-0.0, 0.0])])
    assert g[0] == g[1]


def test_join_partition_levels():
    L = kgo.lib()
    assert L.kgo_join_partition(5 << 61, 1, 3, 0) == 5
    assert L.kgo_join_partition(5 << 59, 1, 5, 0) == 5
    assert any(L.kgo_join_partition(x << 20, 1, 3, 1) != L.kgo_join_partition(x << 20, 2, 3, 1)
               for x in range(64))
    for d in range(1, 8):
        assert L.kgo_join_partition(U64, 1, 3, d) < 8


# ---------------------------------------------------------------- SBBF
def test_sbbf_matches_parquet_spec_vector():
    hashes = np.array([kgo.xxh64(struct.pack("<q", i * 7919 - 3)) for i in range(500)],
                      dtype=np.uint64)
    kept, words = kgo.sbbf_build(hashes, nbytes=1024, max_fpp=1.0)
    assert words.tobytes() == bytes.fromhex(V["sbbf_1024"]["bytes_hex"])
    assert kgo.sbbf_check(words, kgo.xxh64(struct.pack("<q", 7919 * 3 - 3)))


def test_sbbf_config3_size_and_digest():
    import hashlib
    L = kgo.lib()
    assert L.kgo_sbbf_bytes_for(750000, 0.03) == 723968
    assert L.kgo_sbbf_bytes_for(750000, 0.01) == 907648
    keys = np.arange(750000, dtype=np.int64) * 20 + 3
    h = kgo.hash_columns([kgo.Col(kgo.T_INT64, keys)])
    kept, words = kgo.sbbf_build(h, nbytes=723968)
    assert kept
    assert hashlib.sha256(words.tobytes()).hexdigest() == V["sbbf_config3"]["sha256"]


def test_sbbf_sizing_and_false_positives():
    # the reference's own bounds (bloom.rs:349-378)
    L = kgo.lib()
    n = 100000
    b = L.kgo_sbbf_bytes_for(n, 0.01)
    assert 120000 <= b <= 140000
    h = kgo.hash_columns([kgo.Col(kgo.T_UINT64, np.arange(n, dtype=np.uint64))])
    kept, words = kgo.sbbf_build(h, nbytes=b)
    assert all(kgo.sbbf_check(words, int(x)) for x in h[:2000])
    probe = kgo.hash_columns([kgo.Col(kgo.T_UINT64, np.arange(n, 2 * n, dtype=np.uint64))])
    fp = sum(kgo.sbbf_check(words, int(x)) for x in probe)
    assert fp < 1600
    batch = kgo.sbbf_check_batch(words, probe)
    assert batch.sum() == fp and all(batch[i] == kgo.sbbf_check(words, int(probe[i]))
                                     for i in range(0, n, 97))
    assert kgo.sbbf_check_batch(words, h).all()
    nb = len(words) // 8
    est = L.kgo_sbbf_expected_fpp(nb, n)
    assert 0.005 < est < 0.016
    assert L.kgo_sbbf_expected_fpp(nb, n * 20) > 0.05
    assert L.kgo_sbbf_expected_fpp(nb, 0) == 0.0
    assert L.kgo_sbbf_expected_fpp(1, 5000) > 0.99


# ---------------------------------------------------------------- Spark-layout Bloom
def test_spark_murmur3_and_xxhash64_match_spark_goldens():
    """Spark 4.2.0's hash() (Murmur3 x86_32, seed 42) and xxhash64() (seed
    42) of int and bigint values: the hashes the Spark-layout filter uses."""
    with open(os.path.join(GOLD, "spark_hash_int_bigint.json")) as f:
        g = json.load(f)
    L = kgo.lib()
    assert len(g["cases"]) >= 12
    for c in g["cases"]:
        v = c["value"]
        if c["type"] == "int":
            assert L.kgo_murmur3_int(v, 42) == c["hash"], c
            assert kgo.spark_xxhash64([v], 4)[0] == c["xxhash64"], c
        else:
            assert L.kgo_murmur3_long(v, 42) == c["hash"], c
            assert kgo.spark_xxhash64([v], 8)[0] == c["xxhash64"], c


def test_spark_bloom_has_no_false_negatives_and_its_expected_fpp():
    """Layout parity is unpinned (the reference has no Spark filter); what
    any correct filter must do: every inserted item passes, and the rate of
    passing non-members is near (1 - e^(-kn/m))^k."""
    L = kgo.lib()
    n = 20000
    m = L.kgo_spark_bloom_bits_for(n, 0.03)
    k = L.kgo_spark_bloom_hashes_for(n, m)
    assert abs(m - (-n * math.log(0.03) / math.log(2) ** 2)) < 1 and k == 5
    items = kgo.spark_xxhash64(np.arange(n), 8)
    words = kgo.spark_bloom_build(items, m, k)
    assert len(words) == (m + 63) // 64
    assert kgo.spark_bloom_check_batch(words, k, items).all()
    other = kgo.spark_xxhash64(np.arange(n, n + 200000), 8)
    fp = kgo.spark_bloom_check_batch(words, k, other).mean()
    want = (1 - math.exp(-k * n / (len(words) * 64))) ** k
    assert 0.7 * want < fp < 1.3 * want


# ---------------------------------------------------------------- generators
def test_perm_is_a_bijection():
    L = kgo.lib()
    for n in (1, 2, 3, 5, 17, 1000, 4097):
        vals = sorted(L.kgo_perm(i, n, 99) for i in range(n))
        assert vals == list(range(n))
    key, _ = kgo.gen_join_build(1, 0, 1 << 16, 1 << 16, with_val=False)
    assert np.array_equal(np.sort(key), np.arange(1 << 16))


def test_zipf_shape():
    cdf = kgo.zipf_cdf(10_000_000, 1.1)
    assert np.all(np.diff(cdf[:1000]) > 0) and cdf[-1] == 1.0
    # top key share 1/H(10M, 1.1) ~ 11.6 %
    assert 0.110 < cdf[0] < 0.122
    L = kgo.lib()
    assert L.kgo_zipf_rank(cdf.ctypes.data, len(cdf), 0.0) == 0
    assert L.kgo_zipf_rank(cdf.ctypes.data, len(cdf), float(cdf[5])) == 6


# ---------------------------------------------------------------- aggregation
def _dec_text(v, s):
    neg = v < 0
    d = str(abs(v))
    if s == 0:
        return ("-" if neg else "") + d
    d = d.rjust(s + 1, "0")
    return ("-" if neg else "") + d[:-s] + "." + d[-s:]


def test_spark_decimal_sum_avg_goldens():
    """Spark 4.2.0's grouped decimal sum/avg (row i in group i % groups,
    contiguous partitions merged in order), as the reference's
    accumulator/proof_tests.rs:205-293 feeds them."""
    with open(os.path.join(GOLD, "spark_decimal_sum_avg.json")) as f:
        cases = json.load(f)
    assert len(cases) == 172
    for c in cases:
        p, s = c["type"]
        func = kgo.AGG_SUM if c["fn"] == "sum" else kgo.AGG_AVG
        vals = [int(x) for x in c["values"]]
        groups, parts = c["groups"], c["partitions"]
        n = len(vals)
        spec = dict(key_type=(kgo.T_INT64, 0, 0), input_types=[(kgo.T_DECIMAL128, p, s)],
                    aggs=[(func, 0)], ansi=c["ansi"])

        def part(lo, hi):
            a = kgo.Agg(**spec)
            key = kgo.Col(kgo.T_INT64, [i % groups for i in range(lo, hi)])
            a.update(key, [kgo.Col(kgo.T_DECIMAL128, vals[lo:hi], precision=p, scale=s)])
            return a

        try:
            if parts == 1:
                a = part(0, n)
            else:
                a = kgo.Agg(**spec)
                for k in range(parts):
                    a.merge(part(k * n // parts, (k + 1) * n // parts))
            keys, kv, res = a.finish()
            got = None
        except ArithmeticError as e:
            got = str(e)
        if "error" in c:
            assert got is not None and c["error"] in got, (c, got)
            continue
        assert got is None, (c, got)
        vals_out, valid, (t, op, os_) = res[0]
        assert c["type_out"] == f"decimal({op},{os_})", c
        by_key = {int(k): (v, ok) for k, v, ok in zip(keys, vals_out, valid)}
        for g, want in enumerate(c["value"]):
            v, ok = by_key[g]
            assert (want is None) == (not ok), (c, g)
            if ok:
                assert _dec_text(v, os_) == want, (c, g, v)


def test_tpch_q1_decimal_aggregates_match_spark():
    """TPC-H Q1's aggregate over Spark 4.2.0's SF0.01 input, against Spark's
    answer (crates/krishiv-exec/tests/aggregate_tpch.rs). Keys
    (returnflag, linestatus) are dictionary-coded to one int64."""
    import pyarrow.ipc as ipc
    t = ipc.open_file(os.path.join(GOLD, "tpch_q1_input.arrow")).read_all()
    rf = t.column("l_returnflag").to_pylist()
    ls = t.column("l_linestatus").to_pylist()
    code = {}
    key = np.array([code.setdefault((a, b), len(code)) for a, b in zip(rf, ls)], dtype=np.int64)
    cols, types = [], []
    for name in ("l_quantity", "l_extendedprice", "disc_price", "charge", "l_discount"):
        col = t.column(name).combine_chunks()
        p, s = col.type.precision, col.type.scale
        ints = [int(x.scaleb(s)) for x in col.to_pylist()]
        cols.append(kgo.Col(kgo.T_DECIMAL128, ints, precision=p, scale=s))
        types.append((kgo.T_DECIMAL128, p, s))
    aggs = [(kgo.AGG_SUM, 0), (kgo.AGG_SUM, 1), (kgo.AGG_SUM, 2), (kgo.AGG_SUM, 3),
            (kgo.AGG_AVG, 0), (kgo.AGG_AVG, 1), (kgo.AGG_AVG, 4), (kgo.AGG_COUNT_STAR, 0)]
    a = kgo.Agg((kgo.T_INT64, 0, 0), types, aggs, ansi=True)
    a.update(kgo.Col(kgo.T_INT64, key), cols)
    keys, kv, res = a.finish()
    inv = {v: k for k, v in code.items()}
    rows = []
    for gi, k in enumerate(keys):
        f = [inv[int(k)][0], inv[int(k)][1]]
        for vals, valid, (tt, p, s) in res:
            v = vals[gi]
            f.append(_dec_text(v, s) if tt in (kgo.T_DECIMAL64, kgo.T_DECIMAL128) else str(v))
        rows.append(",".join(f))
    with open(os.path.join(GOLD, "tpch_q1_expected.csv")) as fh:
        want = sorted(l.strip() for l in fh if l.strip())
    assert sorted(rows) == want


def _py_agg(keys, kvalid, cols, aggs, ansi=False):
    """Brute-force Spark aggregation over Python values (None = NULL)."""
    groups = {}
    order = []
    for i in range(len(keys)):
        k = keys[i] if kvalid[i] else None
        if k not in groups:
            groups[k] = []
            order.append(k)
        groups[k].append(i)
    out = {}
    for k in order:
        rows = groups[k]
        res = []
        for func, inp in aggs:
            if func == kgo.AGG_COUNT_STAR:
                res.append(len(rows))
                continue
            vals = [cols[inp][i] for i in rows if cols[inp][i] is not None]
            if func == kgo.AGG_COUNT:
                res.append(len(vals))
            elif not vals:
                res.append(None)
            elif func == kgo.AGG_SUM:
                res.append(vals)
            elif func == kgo.AGG_MIN:
                res.append(min(vals))
            elif func == kgo.AGG_MAX:
                res.append(max(vals))
            else:
                res.append(vals)
        out[k] = res
    return out


def test_random_aggregates_match_exact_python():
    rng = np.random.default_rng(11)
    for trial in range(6):
        n = int(rng.integers(1, 3000))
        ng = int(rng.integers(1, 200))
        keys = rng.integers(0, ng, n)
        kvalid = rng.random(n) > 0.05
        i64 = rng.integers(-(1 << 40), 1 << 40, n)
        i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
        dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, n)
        f = rng.standard_normal(n) * 1e3
        vals = [rng.random(n) > 0.1 for _ in range(4)]
        cols = [kgo.Col(kgo.T_INT64, i64, vals[0]), kgo.Col(kgo.T_INT32, i32, vals[1]),
                kgo.Col(kgo.T_DECIMAL64, dec, vals[2], precision=12, scale=2),
                kgo.Col(kgo.T_FLOAT64, f, vals[3])]
        py = [[int(x) if ok else None for x, ok in zip(a, v)]
              for a, v in zip((i64, i32, dec), vals[:3])] + \
             [[float(x) if ok else None for x, ok in zip(f, vals[3])]]
        aggs = [(kgo.AGG_COUNT_STAR, 0)]
        for c in range(4):
            aggs += [(kgo.AGG_COUNT, c), (kgo.AGG_SUM, c), (kgo.AGG_MIN, c), (kgo.AGG_MAX, c),
                     (kgo.AGG_AVG, c)]
        a = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0), (kgo.T_INT32, 0, 0),
                                          (kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0)],
                    aggs)
        a.update(kgo.Col(kgo.T_INT64, keys, kvalid), cols)
        gk, gv, res = a.finish()
        want = _py_agg([int(k) for k in keys], kvalid, py, aggs)
        assert len(gk) == len(want)
        for gi in range(len(gk)):
            k = int(gk[gi]) if gv[gi] else None
            w = want[k]
            for j, (func, inp) in enumerate(aggs):
                vals_o, valid_o, (t, p, s) = res[j]
                got = vals_o[gi] if valid_o[gi] else None
                exp = w[j]
                if func in (kgo.AGG_COUNT_STAR, kgo.AGG_COUNT, kgo.AGG_MIN, kgo.AGG_MAX):
                    assert got == exp, (trial, j)
                elif exp is None:
                    assert got is None
                elif func == kgo.AGG_SUM:
                    if inp == 3:
                        assert math.isclose(got, math.fsum(exp), rel_tol=1e-12, abs_tol=1e-9)
                    else:
                        assert got == sum(exp), (trial, j)
                else:  # AVG
                    if inp == 2:  # decimal(12,2) -> decimal(16,6), HALF_UP
                        q = Fraction(sum(exp) * 10 ** 4, len(exp))
                        r = int(abs(q) + Fraction(1, 2))
                        assert got == (r if q >= 0 else -r), (trial, j)
                    else:
                        assert math.isclose(got, math.fsum(exp) / len(exp), rel_tol=1e-12,
                                            abs_tol=1e-12)


def test_int_sum_wraps_or_raises():
    big = [(1 << 62), (1 << 62), (1 << 62)]
    for ansi in (False, True):
        a = kgo.Agg(None, [(kgo.T_INT64, 0, 0)], [(kgo.AGG_SUM, 0)], ansi=ansi)
        if ansi:
            with pytest.raises(ArithmeticError, match="ARITHMETIC_OVERFLOW"):
                a.update(None, [kgo.Col(kgo.T_INT64, big)])
        else:
            a.update(None, [kgo.Col(kgo.T_INT64, big)])
            _, _, res = a.finish()
            want = (3 << 62) - (1 << 64)
            assert res[0][0][0] == want


def test_float_min_max_nan_is_largest():
    a = kgo.Agg(None, [(kgo.T_FLOAT64, 0, 0)], [(kgo.AGG_MIN, 0), (kgo.AGG_MAX, 0)])
    a.update(None, [kgo.Col(kgo.T_FLOAT64, [1.0, float("nan"), -2.0])])
    _, _, res = a.finish()
    assert res[0][0][0] == -2.0 and math.isnan(res[1][0][0])


def test_float32_min_max_per_group_match_numpy():
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 50, 5000).astype(np.int64)
    v = rng.standard_normal(5000).astype(np.float32)
    a = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_FLOAT32, 0, 0)], [(kgo.AGG_MIN, 0), (kgo.AGG_MAX, 0)])
    a.update(kgo.Col(kgo.T_INT64, keys), [kgo.Col(kgo.T_FLOAT32, v)])
    rk, _, res = a.finish()
    for g, k in enumerate(rk):
        sel = v[keys == k]
        assert res[0][0][g] == float(sel.min()) and res[1][0][g] == float(sel.max())


def test_global_aggregate_on_empty_input_is_one_row_of_nulls():
    a = kgo.Agg(None, [(kgo.T_INT64, 0, 0)], [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0)])
    a.update(None, [kgo.Col(kgo.T_INT64, [1])])  # a global aggregate has its group once seen
    b = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)], [(kgo.AGG_COUNT_STAR, 0)])
    _, _, res = b.finish()
    assert len(res[0][0]) == 0  # grouped on empty input: no rows


# ---------------------------------------------------------------- join
def _brute(jt, pk, pv, bk, bv):
    out = []
    for i in range(len(pk)):
        m = [j for j in range(len(bk)) if pv[i] and bv[j] and pk[i] == bk[j]]
        if jt == kgo.JOIN_INNER:
            out += [(i, j) for j in m]
        elif jt == kgo.JOIN_LEFT_OUTER:
            out += [(i, j) for j in m] if m else [(i, -1)]
        elif jt == kgo.JOIN_LEFT_SEMI:
            out += [(i, -1)] if m else []
        else:
            out += [] if m else [(i, -1)]
    return sorted(out)


@pytest.mark.parametrize("jt", [kgo.JOIN_INNER, kgo.JOIN_LEFT_OUTER, kgo.JOIN_LEFT_SEMI,
                                kgo.JOIN_LEFT_ANTI])
def test_join_hash_equals_nested_loop_and_brute_force(jt):
    rng = np.random.default_rng(jt + 5)
    for _ in range(5):
        nb, npr = int(rng.integers(0, 300)), int(rng.integers(0, 400))
        bk = rng.integers(0, 60, nb)
        pk = rng.integers(0, 80, npr)
        bv = rng.random(nb) > 0.1
        pv = rng.random(npr) > 0.1
        P = [kgo.Col(kgo.T_INT64, pk, pv)]
        B = [kgo.Col(kgo.T_INT64, bk, bv)]
        h = sorted(zip(*kgo.join(jt, P, B)))
        n = sorted(zip(*kgo.join(jt, P, B, nested=True)))
        want = _brute(jt, list(pk), list(pv), list(bk), list(bv))
        assert h == n == want


def test_join_null_safe_and_float_normalization():
    P = [kgo.Col(kgo.T_FLOAT64, [0.0, float("nan"), 1.0, 2.0], [True, True, True, False])]
    B = [kgo.Col(kgo.T_FLOAT64, [-0.0, -float("nan"), 3.0, 0.0], [True, True, True, False])]
    pairs = sorted(zip(*kgo.join(kgo.JOIN_INNER, P, B)))
    assert pairs == [(0, 0), (1, 1)]
    pairs = sorted(zip(*kgo.join(kgo.JOIN_INNER, P, B, null_safe=[1])))
    assert pairs == [(0, 0), (1, 1), (3, 3)]
