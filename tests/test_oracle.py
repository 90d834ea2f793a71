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
    42) of int and bigint values."""
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


# ---------------------------------------------------------------- join types
# The reference's nested-loop oracle (crates/krishiv-exec/tests/hash_join/
# main.rs:130-195), restated in Python over rows (k1, k2, v): string keys are
# mapped to small integers.  Existence follows tests/hash_join/existence.rs:
# every probe row once, plus whether some build row matches it.
ALL_TYPES = [kgo.JOIN_INNER, kgo.JOIN_LEFT_OUTER, kgo.JOIN_LEFT_SEMI, kgo.JOIN_LEFT_ANTI,
             kgo.JOIN_NULL_AWARE_ANTI, kgo.JOIN_EXISTENCE, kgo.JOIN_RIGHT_SEMI,
             kgo.JOIN_RIGHT_ANTI, kgo.JOIN_RIGHT_OUTER, kgo.JOIN_FULL_OUTER]
# VLess (main.rs:72-91): (p + b) % 7 != 0 AND p < b, NULL when (p+b) % 7 == 0
VLESS = [(kgo.X_PROBE, 2, 0), (kgo.X_BUILD, 2, 0), (kgo.X_ADD, 0, 0), (kgo.X_CONST, 0, 7),
         (kgo.X_MOD, 0, 0), (kgo.X_CONST, 0, 0), (kgo.X_NE, 0, 0), (kgo.X_PROBE, 2, 0),
         (kgo.X_BUILD, 2, 0), (kgo.X_LT, 0, 0), (kgo.X_AND, 0, 0)]


def ref_oracle(jt, keys, safe, filt, probe, build):
    def eq(s, x, y):
        return x == y if s else (x is not None and x == y)

    def key_eq(p, b):
        return all(eq(safe[k], p[k], b[k]) for k in range(keys))

    def vless(p, b):
        return (p + b) % 7 != 0 and p < b

    def matches(p, b):
        return key_eq(p, b) and (not filt or vless(p[2], b[2]))

    out = []
    for bi, b in enumerate(build):
        m = [pi for pi, p in enumerate(probe) if matches(p, b)]
        if jt == kgo.JOIN_RIGHT_SEMI and m:
            out.append((-1, bi))
        elif jt == kgo.JOIN_RIGHT_ANTI and not m:
            out.append((-1, bi))
        elif jt == kgo.JOIN_FULL_OUTER and not m:
            out.append((-1, bi))
        elif jt == kgo.JOIN_RIGHT_OUTER:
            out += [(pi, bi) for pi in m] or [(-1, bi)]
    for pi, p in enumerate(probe):
        m = [bi for bi, b in enumerate(build) if matches(p, b)]
        if jt in (kgo.JOIN_INNER, kgo.JOIN_LEFT_OUTER, kgo.JOIN_FULL_OUTER):
            out += [(pi, bi) for bi in m]
            if not m and jt != kgo.JOIN_INNER:
                out.append((pi, -1))
        elif jt == kgo.JOIN_LEFT_SEMI and m:
            out.append((pi, -1))
        elif jt == kgo.JOIN_LEFT_ANTI and not m:
            out.append((pi, -1))
        elif jt == kgo.JOIN_NULL_AWARE_ANTI:
            keep = not build or (all(b[0] is not None for b in build) and p[0] is not None
                                 and not m)
            if keep:
                out.append((pi, -1))
        elif jt == kgo.JOIN_EXISTENCE:
            out.append((pi, 1 if m else 0))
    return sorted(out)


def _cols3(rows):
    k1 = np.array([0 if r[0] is None else r[0] for r in rows], dtype=np.int64)
    k2 = np.array([0 if r[1] is None else r[1] for r in rows], dtype=np.int64)
    v = np.array([r[2] for r in rows], dtype=np.int64)
    return [kgo.Col(kgo.T_INT64, k1, [r[0] is not None for r in rows]),
            kgo.Col(kgo.T_INT64, k2, [r[1] is not None for r in rows]),
            kgo.Col(kgo.T_INT64, v)]


def _run(jt, keys, safe, filt, probe, build, nested):
    P, B = _cols3(probe), _cols3(build)
    op, ob = kgo.join_ex(jt, P[:keys], B[:keys], safe[:keys], nested=nested,
                         filt=VLESS if filt else None, probe_cols=P, build_cols=B)
    return sorted(zip(op.tolist(), ob.tolist()))


A_, B_, X_ = 1, 2, 3  # the reference's string keys "a", "b", "x"
EDGE_PROBE = [(1, A_, 1), (1, A_, 9), (2, None, 3), (None, A_, 4), (None, None, 5), (7, X_, 6)]
EDGE_BUILD = [(1, A_, 8), (1, A_, 2), (1, B_, 3), (2, None, 12), (None, None, 10), (None, A_, 11)]
ALL_NULL = [(None, None, 1), (None, None, 2)]


@pytest.mark.parametrize("jt", ALL_TYPES)
def test_join_types_match_the_reference_oracle_on_edge_cases(jt):
    """The reference's edge cases (main.rs:232-290): NULL keys, duplicates on
    both sides, empty and all-NULL inputs, one and two keys, `=` and `<=>`,
    with and without the residual filter; hash and nested-loop forms."""
    inputs = [(EDGE_PROBE, EDGE_BUILD), (EDGE_PROBE, []), ([], EDGE_BUILD),
              (EDGE_PROBE, ALL_NULL), (ALL_NULL, EDGE_BUILD)]
    for keys in (1, 2):
        for s in (False, True):
            for filt in (False, True):
                if jt == kgo.JOIN_NULL_AWARE_ANTI and (keys > 1 or s or filt):
                    continue
                safe = [s] * keys
                for p, b in inputs:
                    want = ref_oracle(jt, keys, safe, filt, p, b)
                    for nested in (False, True):
                        assert _run(jt, keys, safe, filt, p, b, nested) == want, \
                            (jt, keys, s, filt, len(p), len(b), nested)


def test_not_in_follows_spark_null_semantics():
    """main.rs:300-321: an empty subquery keeps every row (NULL too); a NULL
    in it keeps none; otherwise a NULL probe value is dropped."""
    probe = [(1, None, 0), (2, None, 0), (None, None, 0)]
    jt = kgo.JOIN_NULL_AWARE_ANTI
    assert [p for p, _ in _run(jt, 1, [False], False, probe, [], False)] == [0, 1, 2]
    assert _run(jt, 1, [False], False, probe, [(1, None, 0), (None, None, 0)], False) == []
    assert _run(jt, 1, [False], False, probe, [(1, None, 0)], False) == [(1, -1)]


def test_random_joins_match_the_reference_oracle():
    """main.rs:356-380 in small: random rows with NULLs, keys in 0..6, values
    in 0..20, every join type, one or two keys, `=`/`<=>`, filter on/off."""
    rng = random.Random(7)

    def row():
        k1 = None if rng.random() < 0.2 else rng.randrange(6)
        k2 = None if rng.random() < 0.2 else rng.randrange(3)
        return (k1, k2, rng.randrange(20))

    for case in range(150):
        jt = ALL_TYPES[case % len(ALL_TYPES)]
        naaj = jt == kgo.JOIN_NULL_AWARE_ANTI
        keys = 1 if naaj else rng.choice([1, 2])
        safe = [False] if naaj else [rng.random() < 0.5 for _ in range(keys)]
        filt = (not naaj) and rng.random() < 0.5
        probe = [row() for _ in range(rng.randrange(0, 40))]
        build = [row() for _ in range(rng.randrange(0, 40))]
        want = ref_oracle(jt, keys, safe, filt, probe, build)
        assert _run(jt, keys, safe, filt, probe, build, False) == want, case


def test_residual_filter_three_valued_logic_and_arithmetic():
    """The filter program's operators on single pairs: wrapping arithmetic,
    % with the dividend's sign and NULL for a zero divisor, Kleene AND/OR,
    NOT, IS NULL, comparisons with NULL."""
    P = [kgo.Col(kgo.T_INT64, np.array([0], dtype=np.int64)),
         kgo.Col(kgo.T_INT64, np.array([-7], dtype=np.int64)),
         kgo.Col(kgo.T_INT64, np.array([0], dtype=np.int64), [False])]
    B = [kgo.Col(kgo.T_INT64, np.array([0], dtype=np.int64)),
         kgo.Col(kgo.T_INT64, np.array([3], dtype=np.int64))]
    X = kgo

    def passes(prog):
        op, _ = kgo.join_ex(kgo.JOIN_LEFT_SEMI, P[:1], B[:1], filt=prog, probe_cols=P,
                            build_cols=B)
        return len(op) == 1

    pm, bm, c = (X.X_PROBE, 1, 0), (X.X_BUILD, 1, 0), lambda v: (X.X_CONST, 0, v)
    pnull = (X.X_PROBE, 2, 0)
    assert passes([pm, bm, (X.X_MOD, 0, 0), c(-1), (X.X_EQ, 0, 0)])       # -7 % 3 == -1
    assert not passes([pm, c(0), (X.X_MOD, 0, 0), c(0), (X.X_EQ, 0, 0)])  # x % 0 is NULL
    assert passes([c(-(1 << 63)), c(-1), (X.X_ADD, 0, 0), c((1 << 63) - 1), (X.X_EQ, 0, 0)])
    assert passes([c(1 << 62), c(4), (X.X_MUL, 0, 0), c(0), (X.X_EQ, 0, 0)])
    assert not passes([pnull, c(1), (X.X_LT, 0, 0)])
    assert not passes([pnull, c(1), (X.X_LT, 0, 0), (X.X_NOT, 0, 0)])
    assert passes([pnull, (X.X_IS_NULL, 0, 0)])
    # NULL AND FALSE = FALSE; NULL OR TRUE = TRUE; NULL AND TRUE = NULL
    f, t, n = [c(1), c(2), (X.X_GT, 0, 0)], [c(1), c(2), (X.X_LT, 0, 0)], [pnull, c(1), (X.X_LT, 0, 0)]
    assert not passes(n + f + [(X.X_AND, 0, 0)])
    assert passes(n + f + [(X.X_AND, 0, 0), (X.X_NOT, 0, 0)])
    assert passes(n + t + [(X.X_OR, 0, 0)])
    assert not passes(n + t + [(X.X_AND, 0, 0)])
    assert not passes(n + t + [(X.X_AND, 0, 0), (X.X_NOT, 0, 0)])


def test_parallel_aggregate_equals_single_threaded_and_fsum():
    """The thread-pool form (kgo_pagg) returns the single-threaded
    aggregate's integer and decimal results bit for bit; with compensated
    float sums its SUM/AVG equal math.fsum within one unit in the last
    place, and the row-order double sums stay within 1e-12 of it."""
    n, nd = 300_000, 20_000
    cdf = kgo.zipf_cdf(nd, 1.1)
    key, dec, f, i32, _ = kgo.gen_agg(3, 0, n, cdf, nd)
    its = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0), (kgo.T_INT32, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0), (kgo.AGG_SUM, 0), (kgo.AGG_SUM, 1), (kgo.AGG_AVG, 1),
            (kgo.AGG_MIN, 2), (kgo.AGG_MAX, 1), (kgo.AGG_AVG, 0), (kgo.AGG_SUM, 2)]
    ins = lambda a, b: [kgo.Col(kgo.T_DECIMAL64, dec[a:b], precision=12, scale=2),
                        kgo.Col(kgo.T_FLOAT64, f[a:b]), kgo.Col(kgo.T_INT32, i32[a:b])]
    one = kgo.Agg((kgo.T_INT64, 0, 0), its, aggs)
    one.update(kgo.Col(kgo.T_INT64, key), ins(0, n))
    k1, _, r1 = one.finish(raw=True)
    par = kgo.AggPar((kgo.T_INT64, 0, 0), its, aggs, 6, exact_f64=True)
    for a in range(0, n, 70_000):
        par.update(kgo.Col(kgo.T_INT64, key[a:a + 70_000]), ins(a, a + 70_000))
    k2, _, r2 = par.finish(raw=True)
    o1, o2 = np.argsort(k1), np.argsort(k2)
    assert np.array_equal(k1[o1], k2[o2])
    for j, ((v1, ok1, t), (v2, ok2, _)) in enumerate(zip(r1, r2)):
        assert np.array_equal(ok1[o1], ok2[o2])
        if t[0] != kgo.T_FLOAT64:
            assert np.array_equal(np.asarray(v1)[o1], np.asarray(v2)[o2]), j
    # exact sums against math.fsum on the 200 largest groups and a sample
    order = np.argsort(key, kind="stable")
    ks, fs = key[order], f[order]
    starts = np.searchsorted(ks, k2)
    ends = np.searchsorted(ks, k2, side="right")
    for g in np.concatenate([np.argsort(ends - starts)[-200:], np.arange(0, len(k2), 97)]):
        exact = math.fsum(fs[starts[g]:ends[g]])
        assert abs(r2[2][0][g] - exact) <= 2 * math.ulp(exact), g
        assert abs(r2[3][0][g] - exact / (ends[g] - starts[g])) <= 4 * math.ulp(r2[3][0][g])
    naive = dict(zip(k1.tolist(), r1[2][0].tolist()))
    for g in range(0, len(k2), 13):
        assert math.isclose(naive[int(k2[g])], r2[2][0][g], rel_tol=1e-12)


def test_grace_factorization_equals_the_pair_walk():
    """Configuration 4's per-partition COUNT and SUM(probe.pay * build.pay),
    factorized per key, against every pair walked by the join oracle: the
    wrapping sum (mode 0), the sum mod 2^31 (mode 1), and the per-pair
    SUM(p * b % 2^31) that mode 0 equals when every product is below 2^31."""
    rng = np.random.default_rng(3)
    cdf = kgo.zipf_cdf(700, 0.8)
    bk = kgo.gen_agg(6, 0, 4000, cdf, 700)[0]
    pk = kgo.gen_agg(7, 0, 9000, cdf, 700)[0]
    bkv, pkv = rng.random(4000) > 0.02, rng.random(9000) > 0.02
    for hi in (256, 1 << 40):
        bp, pp = rng.integers(0, hi, 4000), rng.integers(-hi, hi, 9000)
        bpv, ppv = rng.random(4000) > 0.05, rng.random(9000) > 0.05
        op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk, pkv)],
                          [kgo.Col(kgo.T_INT64, bk, bkv)])
        pids = kgo.route([kgo.Col(kgo.T_INT64, pk[op])], 64)
        both = ppv[op] & bpv[ob]
        with np.errstate(over="ignore"):
            prod = pp[op].astype(np.uint64) * bp[ob].astype(np.uint64)
        s = np.zeros(64, dtype=np.uint64)
        np.add.at(s, pids[both], prod[both])
        v = np.bincount(pids[both], minlength=64) > 0
        for mode in (0, 1):
            c, got, gv = kgo.grace_factorized(kgo.Col(kgo.T_INT64, bk, bkv),
                                              kgo.Col(kgo.T_INT64, bp, bpv),
                                              kgo.Col(kgo.T_INT64, pk, pkv),
                                              kgo.Col(kgo.T_INT64, pp, ppv), 64, mode)
            assert np.array_equal(c, np.bincount(pids, minlength=64))
            want = s if mode == 0 else s & np.uint64(0x7FFFFFFF)
            assert np.array_equal(got.view(np.uint64), want) and np.array_equal(gv, v)
        if hi == 256:
            per_pair = np.zeros(64, dtype=np.int64)
            ok = both & (pp[op] >= 0)
            np.add.at(per_pair, pids[ok], (pp[op][ok] * bp[ob][ok]) % (1 << 31))
            c, got, _ = kgo.grace_factorized(kgo.Col(kgo.T_INT64, bk, bkv),
                                             kgo.Col(kgo.T_INT64, bp, bpv),
                                             kgo.Col(kgo.T_INT64, pk, pkv),
                                             kgo.Col(kgo.T_INT64, np.abs(pp), ppv), 64, 0)
            pa = np.abs(pp)
            per_pair = np.zeros(64, dtype=np.int64)
            np.add.at(per_pair, pids[both], (pa[op][both] * bp[ob][both]) % (1 << 31))
            assert np.array_equal(got, per_pair)
