"""The device runtime KeyFilter against the oracle (oracle/kgo_filter.c):
the form a build publishes, its members, range and Bloom bits, its
versioned encoding byte for byte (runtime_filter/codec.rs), evaluation on
probe columns of the key type and of narrower types (member.rs:439-560),
decoding, builder and filter merges (filter.rs:478-510, merge.rs:35-169),
and the row stage's selections and switch-off decisions batch by batch
(operators/runtime_filter.rs:329-395)."""
import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col, DeviceBuffer
from kgpu import abi as A

pytestmark = pytest.mark.gpu

DEC = (kgo.T_DECIMAL64, kgo.T_DECIMAL128)


def oracle_opts(d: A.FilterOptions) -> kgo.FilterOpts:
    o = kgo.FilterOpts.default()
    for name, _ in A.FilterOptions._fields_:
        setattr(o, name, getattr(d, name))
    return o


def both(ctype, batches, opts=None, precision=0, scale=0, include=None):
    """The same keys into the device builder and the oracle's; both finished."""
    d_opts = opts or kgpu.KeyFilter.options()
    g = kgpu.KeyFilter(ctx(), (ctype, precision, scale), d_opts)
    o = kgo.KeyFilter(ctype, precision, scale, oracle_opts(d_opts))
    for b, (vals, valid) in enumerate(batches):
        inc = None if include is None else include[b]
        g.add(Col(ctype, vals, valid, precision, scale),
              None if inc is None else Col(A.T_BOOL, inc))
        o.add(kgo.Col(ctype, vals, valid, precision=precision, scale=scale), inc)
    return g.finish(), o.finish()


def same_filter(g, o):
    gi, oi = g.info(), o.info()
    for k in ("form", "has_range", "nulls_pass", "empty", "keys", "bloom_bytes", "values"):
        assert gi[k] == oi[k], (k, gi[k], oi[k])
    if gi["has_range"]:
        for k in ("lo", "hi", "flo", "fhi"):
            assert gi[k] == oi[k] or (gi[k] != gi[k] and oi[k] != oi[k]), (k, gi[k], oi[k])
    assert g.members() == o.members()
    assert g.encode() == o.encode()


def rows_of(g, col_g, n):
    out, cnt = g.evaluate(col_g, n)
    return out.to_numpy(np.int32, cnt) if cnt else np.zeros(0, np.int32)


def check_eval(g, o, ctype, probe, pv, precision=0, scale=0):
    want = np.flatnonzero(o.evaluate(kgo.Col(ctype, probe, pv, precision=precision, scale=scale)))
    got = rows_of(g, Col(ctype, probe, pv, precision, scale), len(probe))
    assert np.array_equal(got, want)


def keys_for(ctype, rng, n, spread):
    if ctype == kgo.T_INT8:
        return rng.integers(-128, 128, n).astype(np.int8)
    if ctype == kgo.T_INT16:
        return rng.integers(-spread, spread, n).astype(np.int16)
    if ctype in (kgo.T_INT32, kgo.T_DATE32):
        return rng.integers(-spread, spread, n).astype(np.int32)
    if ctype in (kgo.T_INT64, kgo.T_TIMESTAMP, kgo.T_DECIMAL64):
        return rng.integers(-spread, spread, n).astype(np.int64) * 1_000_003
    if ctype == kgo.T_FLOAT32:
        return (rng.integers(-spread, spread, n) / 8.0).astype(np.float32)
    if ctype == kgo.T_FLOAT64:
        return rng.integers(-spread, spread, n) / 8.0
    # decimal128: low words of 0 and wide values among them
    base = [int(x) * (10 ** 20) + (int(x) % 3) * 7 for x in rng.integers(-spread, spread, n)]
    return base


TYPES = [kgo.T_INT8, kgo.T_INT16, kgo.T_INT32, kgo.T_INT64, kgo.T_DATE32, kgo.T_TIMESTAMP,
         kgo.T_FLOAT32, kgo.T_FLOAT64, kgo.T_DECIMAL64, kgo.T_DECIMAL128]


@pytest.mark.parametrize("ctype", TYPES)
@pytest.mark.parametrize("distinct", [5, 700, 20000])
def test_forms_encodings_and_evaluation_match(ctype, distinct):
    say(f"filter: type {ctype}, {distinct} distinct")
    rng = np.random.default_rng(ctype * 100 + distinct)
    p, s = (30, 4) if ctype == kgo.T_DECIMAL128 else (15, 2) if ctype == kgo.T_DECIMAL64 else (0, 0)
    n = 60_000
    pool = keys_for(ctype, rng, distinct, 10 * distinct)
    idx = rng.integers(0, len(pool), n)
    vals = [pool[i] for i in idx] if ctype == kgo.T_DECIMAL128 else np.asarray(pool)[idx]
    valid = rng.random(n) > 0.01
    half = n // 2
    batches = [(vals[:half], valid[:half]), (vals[half:], valid[half:])]
    g, o = both(ctype, batches, precision=p, scale=s)
    same_filter(g, o)
    probe_pool = keys_for(ctype, rng, 3000, 10 * distinct)
    pidx = rng.integers(0, len(pool), 40_000)
    probe = ([pool[i] for i in pidx] + list(probe_pool) if ctype == kgo.T_DECIMAL128 else
             np.concatenate([np.asarray(pool)[pidx], np.asarray(probe_pool)]))
    pv = rng.random(len(probe)) > 0.05
    check_eval(g, o, ctype, probe, pv, p, s)


def test_dense_small_exact_and_range_forms():
    say("filter: the form cutoffs")
    rng = np.random.default_rng(1)
    cases = [np.arange(100, dtype=np.int64),                       # dense: the range
             np.array([5, 9, -3, 77], dtype=np.int64),              # small
             rng.choice(np.arange(7200), 1028, replace=False).astype(np.int64),  # bit set
             rng.integers(-(1 << 60), 1 << 60, 3000),               # hash set
             np.array([np.iinfo(np.int64).min, 0, 5], dtype=np.int64)]  # the empty-slot pattern
    forms = [A.FORM_RANGE, A.FORM_SMALL, A.FORM_DENSE, A.FORM_EXACT, A.FORM_SMALL]
    for keys, form in zip(cases, forms):
        g, o = both(kgo.T_INT64, [(keys, None)])
        assert g.info()["form"] == form
        same_filter(g, o)
        probe = np.concatenate([keys, rng.integers(-100, 8000, 5000)])
        check_eval(g, o, kgo.T_INT64, probe, None)


def test_bloom_rules_include_and_nulls():
    say("filter: Bloom keep rule, include column, NULL keys")
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 1 << 50, 200_000)
    valid = rng.random(len(keys)) > 0.001
    inc = rng.random(len(keys)) > 0.3
    for opts in (kgpu.KeyFilter.options(),
                 kgpu.KeyFilter.options(bloom_max_bytes=4096),        # fpp too high: range
                 kgpu.KeyFilter.options(bloom_max_items=100_000),     # hashes dropped
                 kgpu.KeyFilter.options(bloom_bytes=1 << 20),         # planned size
                 kgpu.KeyFilter.options(side=A.SIDE_UPPER),           # a bound: the range only
                 kgpu.KeyFilter.options(max_values=0, exact_max_values=0)):
        g, o = both(kgo.T_INT64, [(keys[:70_000], valid[:70_000]), (keys[70_000:], valid[70_000:])],
                    opts, include=[inc[:70_000], inc[70_000:]])
        same_filter(g, o)
        probe = np.concatenate([keys[:50_000], rng.integers(0, 1 << 50, 50_000)])
        check_eval(g, o, kgo.T_INT64, probe, rng.random(len(probe)) > 0.02)
    # an empty build: only NULL keys pass when a NULL key was kept
    g, o = both(kgo.T_INT64, [(np.array([1, 2], np.int64), np.array([False, False]))])
    assert g.info()["empty"] and g.info()["nulls_pass"]
    same_filter(g, o)
    check_eval(g, o, kgo.T_INT64, np.array([1, 2, 3]), np.array([True, False, True]))


def test_widened_probe_columns():
    say("filter: narrower probe types and rescaled decimals")
    rng = np.random.default_rng(3)
    for n_keys in (6, 900, 30_000):
        keys = rng.integers(-30_000, 30_000, n_keys)
        g, o = both(kgo.T_INT64, [(keys, None)])
        probe = rng.integers(-30_000, 30_000, 20_000)
        for t in (kgo.T_INT32, kgo.T_INT16):
            check_eval(g, o, t, probe, None)
        gf, of = both(kgo.T_FLOAT64, [(keys.astype(np.float64), None)])
        check_eval(gf, of, kgo.T_INT32, probe, None)
        check_eval(gf, of, kgo.T_FLOAT32, probe.astype(np.float32), None)
        gd, od = both(kgo.T_DECIMAL128, [([int(k) * 100 for k in keys], None)], precision=30,
                      scale=4)
        check_eval(gd, od, kgo.T_DECIMAL64, probe * 10, None, precision=12, scale=3)
        check_eval(gd, od, kgo.T_INT32, probe, None)
    with pytest.raises(kgpu.KgError, match="UNSUPPORTED"):
        g.evaluate(Col(A.T_FLOAT64, np.zeros(4)))


def test_builder_merge_and_filter_merge_match():
    say("filter: builder merges and filter merges")
    rng = np.random.default_rng(4)
    d_opts = kgpu.KeyFilter.options(bloom_bytes=64 << 10, exact_max_values=500)
    o_opts = oracle_opts(d_opts)
    parts = [rng.integers(0, 1 << 40, 300), rng.integers(0, 1 << 40, 300),
             rng.integers(0, 1 << 40, 20_000), np.arange(40, dtype=np.int64)]
    built = [both(kgo.T_INT64, [(p, None)], d_opts) for p in parts]
    # builder merge: one builder from three partials
    g = kgpu.KeyFilter(ctx(), (A.T_INT64, 0, 0), d_opts)
    o = kgo.KeyFilter(kgo.T_INT64, opts=o_opts)
    g.add(Col(A.T_INT64, parts[0]))
    o.add(kgo.Col(kgo.T_INT64, parts[0]))
    for p in parts[1:3]:
        gp = kgpu.KeyFilter(ctx(), (A.T_INT64, 0, 0), d_opts).add(Col(A.T_INT64, p))
        g.merge_builder(gp)
        o.merge_builder(kgo.KeyFilter(kgo.T_INT64, opts=o_opts).add(kgo.Col(kgo.T_INT64, p)))
    same_filter(g.finish(), o.finish())
    # filter merges: set + set, Bloom + Bloom, Bloom + set, range + set
    for a, b in ((0, 1), (2, 2), (2, 0), (3, 0), (0, 3)):
        gm = built[a][0].merge(built[b][0], d_opts)
        om = built[a][1].merge(built[b][1], o_opts)
        same_filter(gm, om)
        probe = np.concatenate([parts[a][:200], parts[b][:200], rng.integers(0, 1 << 40, 2000)])
        check_eval(gm, om, kgo.T_INT64, probe, None)
    # past the exact cap without a plan: the range; with a tight fpp: the range
    plain = kgpu.KeyFilter.options(exact_max_values=500)
    x = both(kgo.T_INT64, [(parts[0], None)], plain)
    y = both(kgo.T_INT64, [(parts[1] + (1 << 41), None)], plain)
    same_filter(x[0].merge(y[0], plain), x[1].merge(y[1], oracle_opts(plain)))
    tight = kgpu.KeyFilter.options(bloom_bytes=256, exact_max_values=10, bloom_max_fpp=0.05)
    x = both(kgo.T_INT64, [(parts[0], None)], tight)
    y = both(kgo.T_INT64, [(parts[1], None)], tight)
    same_filter(x[0].merge(y[0], tight), x[1].merge(y[1], oracle_opts(tight)))


def test_decode_of_oracle_bytes_and_refusals():
    say("filter: decode")
    rng = np.random.default_rng(5)
    for keys in (np.arange(20, dtype=np.int64) * 3, rng.integers(0, 1 << 60, 700),
                 rng.integers(0, 1 << 40, 9000), np.array([4, -4], np.int64)):
        o = kgo.KeyFilter(kgo.T_INT64).add(kgo.Col(kgo.T_INT64, keys)).finish()
        g = kgpu.KeyFilter.decode(ctx(), o.encode())
        same_filter(g, o)
        check_eval(g, o, kgo.T_INT64, rng.integers(-10, 100, 3000), None)
    e = o.encode()
    for cut in (0, 2, 5, len(e) - 1):
        with pytest.raises(kgpu.KgError):
            kgpu.KeyFilter.decode(ctx(), e[:cut])
    with pytest.raises(kgpu.KgError, match="trailing"):
        kgpu.KeyFilter.decode(ctx(), e + b"\0")


def test_row_stage_matches_the_oracle_batch_by_batch():
    say("filter: row stage, 12 batches of 50K rows")
    rng = np.random.default_rng(6)
    n = 50_000
    fb_keys = rng.integers(0, 1_000_000, 30_000)
    specs = [(kgo.T_INT64, fb_keys),                                   # Bloom on column 0
             (kgo.T_INT64, np.arange(0, 40, dtype=np.int64)),           # range on column 1
             (kgo.T_INT64, np.arange(0, 100, dtype=np.int64)),          # passes all of column 1
             (kgo.T_INT32, rng.choice(np.arange(5000), 800, replace=False).astype(np.int32))]
    gf, of = [], []
    for t, k in specs:
        g, o = both(t, [(k, None)])
        gf.append(g)
        of.append(o)
    gs = kgpu.RowStage(ctx(), 65_536, 90)
    os_ = kgo.RowStage(65_536, 0.9)
    for slot, col in ((0, 0), (1, 1), (2, 1), (3, 2)):
        gs.set(slot, gf[slot], col)
        os_.set(slot, of[slot], col)
    for b in range(12):
        cols_np = [rng.integers(0, 1_000_000, n), rng.integers(0, 100 if b < 6 else 45, n),
                   rng.integers(0, 5000, n).astype(np.int32)]
        valid = rng.random(n) > 0.01
        types = [kgo.T_INT64, kgo.T_INT64, kgo.T_INT32]
        want = os_.apply([kgo.Col(t, c, valid) for t, c in zip(types, cols_np)])
        out, cnt = gs.apply([Col(t, c, valid) for t, c in zip(types, cols_np)])
        if want is None:
            assert cnt == -1, b
        else:
            assert np.array_equal(out.to_numpy(np.int32, cnt), np.flatnonzero(want)), b
        for slot in range(4):
            assert gs.stats(slot) == os_.stats(slot), (b, slot)
        if b == 8:
            gs.start_partition()
            os_.start_partition()
        if b == 9:
            gs.set(2, gf[2], 1)  # a new generation
            os_.set(2, of[2], 1)
            gs.set(0, None, 0)   # pending
            os_.set(0, None, 0)
    assert any(gs.stats(s)[2] for s in range(4)) or gs.stats(2)[0] > 0


def test_filter_at_config3_size():
    """15M build keys into a Bloom of the planned size, 60M probe rows: the
    filter bytes equal the oracle's; the survivors equal a host recount."""
    say("filter: 15M-key Bloom, 60M probe rows")
    nb, npr = 15_000_000, 60_000_000
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << 62, nb)
    opts = kgpu.KeyFilter.options(bloom_max_items=1 << 26, bloom_max_bytes=64 << 20)
    g, o = both(kgo.T_INT64, [(keys[:nb // 2], None), (keys[nb // 2:], None)], opts)
    assert g.info()["form"] == A.FORM_BLOOM
    assert g.encode() == o.encode()
    probe = np.concatenate([keys[:5_000_000], rng.integers(0, 1 << 62, npr - 5_000_000)])
    dp = DeviceBuffer.from_numpy(ctx(), probe)
    out, cnt = g.evaluate(Col(A.T_INT64, dp, length=npr))
    rows = out.to_numpy(np.int32, cnt)
    assert np.all(rows[:5_000_000] == np.arange(5_000_000))  # every stored key passes
    sample = np.sort(rng.choice(npr, 200_000, replace=False))
    want = o.evaluate(kgo.Col(kgo.T_INT64, probe[sample]))
    got = np.isin(sample, rows)
    assert np.array_equal(got, want)
