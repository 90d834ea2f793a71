"""GPU parity of the routing hash H_kg, exchange routing and the stable hash
repartition against the oracle (bit-exact)."""
import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def _both(ctype, values, valid=None, p=0, s=0):
    return kgo.Col(ctype, values, valid, precision=p, scale=s), Col(ctype, values, valid, p, s)


def test_hash_every_key_type_matches_oracle():
    say("exchange: H_kg per type")
    rng = np.random.default_rng(1)
    n = 5000
    valid = rng.random(n) > 0.1
    cases = [
        (A.T_INT64, rng.integers(-(1 << 62), 1 << 62, n)),
        (A.T_INT32, rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)),
        (A.T_INT16, rng.integers(-30000, 30000, n).astype(np.int16)),
        (A.T_INT8, rng.integers(-128, 127, n).astype(np.int8)),
        (A.T_UINT64, rng.integers(0, 1 << 63, n).astype(np.uint64)),
        (A.T_DATE32, rng.integers(0, 20000, n).astype(np.int32)),
        (A.T_TIMESTAMP, rng.integers(0, 1 << 50, n)),
        (A.T_BOOL, rng.random(n) > 0.5),
    ]
    for t, v in cases:
        oc, gc = _both(t, v, valid)
        assert np.array_equal(kgo.hash_columns([oc]), kgpu.hash_columns(ctx(), [gc])), t
    f = rng.standard_normal(n)
    f[:4] = [0.0, -0.0, np.nan, -np.nan]
    for t, v in [(A.T_FLOAT64, f), (A.T_FLOAT32, f.astype(np.float32))]:
        oc, gc = _both(t, v, valid)
        assert np.array_equal(kgo.hash_columns([oc]), kgpu.hash_columns(ctx(), [gc])), t
    d = rng.integers(-(10 ** 17), 10 ** 17, n)
    oc, gc = _both(A.T_DECIMAL64, d, valid, 18, 2)
    assert np.array_equal(kgo.hash_columns([oc]), kgpu.hash_columns(ctx(), [gc]))
    big = [int(x) * 10 ** 19 + 7 for x in d]
    oc, gc = _both(A.T_DECIMAL128, big, valid, 38, 2)
    assert np.array_equal(kgo.hash_columns([oc]), kgpu.hash_columns(ctx(), [gc]))


def test_hash_goldens_on_device():
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "vectors.json")))["routing"]
    vals = g["int64"]["values"]
    h = kgpu.hash_columns(ctx(), [Col(A.T_INT64, vals + [0], [True] * len(vals) + [False])])
    assert list(h[:4]) == g["int64"]["hashes"] and int(h[4]) == g["null_hash"]


def test_multi_column_fold_and_route():
    say("exchange: multi-column fold and route")
    rng = np.random.default_rng(2)
    n = 20000
    a = rng.integers(0, 1000, n)
    b = rng.integers(-50, 50, n).astype(np.int32)
    va, vb = rng.random(n) > 0.05, rng.random(n) > 0.05
    oa, ga = _both(A.T_INT64, a, va)
    ob, gb = _both(A.T_INT32, b, vb)
    assert np.array_equal(kgo.hash_columns([oa, ob]), kgpu.hash_columns(ctx(), [ga, gb]))
    for parts in (1, 2, 8, 37, 64, 4096, 5000):
        assert np.array_equal(kgo.route([oa, ob], parts), kgpu.route(ctx(), [ga, gb], parts)), parts


@pytest.mark.parametrize("parts", [1, 7, 64, 1024, 4096])
def test_repartition_is_stable_and_complete(parts):
    say(f"exchange: repartition into {parts}")
    rng = np.random.default_rng(parts)
    n = 300_000
    key = rng.integers(-(1 << 40), 1 << 40, n)
    kv = rng.random(n) > 0.02
    pay = np.arange(n, dtype=np.int64)
    small = rng.integers(-100, 100, n).astype(np.int32)
    sv = rng.random(n) > 0.3
    flags = rng.random(n) > 0.5
    cols = [Col(A.T_INT64, key, kv), Col(A.T_INT64, pay), Col(A.T_INT32, small, sv),
            Col(A.T_BOOL, flags)]
    batch, offs = kgpu.repartition(ctx(), cols, [0], parts)
    pids = kgo.route([kgo.Col(kgo.T_INT64, key, kv)], parts)
    want_offs, want_rows = kgo.split(pids, parts)
    assert np.array_equal(offs, want_offs)
    got_pay, _ = batch.column(1)
    assert np.array_equal(got_pay, want_rows)  # stable: input order within a partition
    gk, gkv = batch.column(0)
    assert np.array_equal(gk[gkv], key[want_rows][kv[want_rows]])
    assert np.array_equal(gkv, kv[want_rows])
    gs, gsv = batch.column(2)
    assert np.array_equal(gsv, sv[want_rows])
    assert np.array_equal(gs[gsv], small[want_rows][sv[want_rows]])
    gf, _ = batch.column(3)
    assert np.array_equal(gf, flags[want_rows])


def test_repartition_device_columns_and_empty():
    ctxx = ctx()
    key = np.arange(1000, dtype=np.int64)
    kb = kgpu.DeviceBuffer.from_numpy(ctxx, key)
    batch, offs = kgpu.repartition(ctxx, [Col(A.T_INT64, kb, length=1000)], [0], 16)
    want_offs, want_rows = kgo.split(kgo.route([kgo.Col(kgo.T_INT64, key)], 16), 16)
    assert np.array_equal(offs, want_offs)
    assert np.array_equal(batch.column(0)[0], key[want_rows])
    batch, offs = kgpu.repartition(ctxx, [Col(A.T_INT64, key[:0])], [0], 16)
    assert batch.num_rows == 0 and np.all(offs == 0)


@pytest.mark.parametrize("depth", [0, 1, 3, 6, 7])
def test_join_partition_matches_oracle(depth):
    """The join's row-layout hash h and its grace partition (Hashing::partition)
    on the device equal the oracle's, over one and two key columns of several
    types with NULLs, NaN and -0.0."""
    say(f"exchange: join partition depth {depth}")
    rng = np.random.default_rng(40 + depth)
    n = 50_000
    i64 = rng.integers(-(1 << 62), 1 << 62, n)
    i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
    f = rng.standard_normal(n)
    f[::97] = np.nan
    f[::89] = -0.0
    ok = rng.random(n) > 0.05
    dec = [int(x) * (1 << 70) + int(y) for x, y in zip(rng.integers(-1000, 1000, n), rng.integers(0, 1 << 60, n))]
    cases = [
        ([kgo.Col(kgo.T_INT64, i64)], [Col(A.T_INT64, i64)]),
        ([kgo.Col(kgo.T_INT64, i64, ok), kgo.Col(kgo.T_INT32, i32)],
         [Col(A.T_INT64, i64, ok), Col(A.T_INT32, i32)]),
        ([kgo.Col(kgo.T_FLOAT64, f, ok)], [Col(A.T_FLOAT64, f, ok)]),
        ([kgo.Col(kgo.T_DECIMAL128, kgo.dec128_words(dec), precision=38, scale=2)],
         [Col(A.T_DECIMAL128, dec, precision=38, scale=2)]),
    ]
    L = kgo.lib()
    for ocols, gcols in cases:
        want_h = kgo.row_hash(ocols)
        for radix in (3, 6, 9):
            h, p = kgpu.join_partition(ctx(), gcols, radix, depth, 0xC0FFEE)
            assert np.array_equal(h, want_h)
            want_p = [L.kgo_join_partition(int(x), 0xC0FFEE, radix, depth) for x in want_h[:2000]]
            assert p[:2000].tolist() == want_p
