"""Pins of the CPU oracle's runtime KeyFilter (oracle/kgo_filter.c) against
what the reference fixes independently of its formulas: brute-force
membership, the form cutoffs it states (runtime_filter/member.rs:1-60,
138-176), hand-assembled encodings of the codec layout (codec.rs:6-20,
104-180), the decoder's refusals (form_tests.rs:474-537), merges against the
filter of the union of the keys (merge.rs:1-22), and the row stage's
ordering, narrowing and switch-off (operators/runtime_filter.rs:1-45,
329-395) checked by counting."""
import struct

import numpy as np
import pytest

import kgo
from kgo import Col, KeyFilter, FilterOpts

I64 = kgo.T_INT64


def build(ctype, values, valid=None, opts=None, precision=0, scale=0):
    f = KeyFilter(ctype, precision, scale, opts)
    f.add(Col(ctype, values, valid, precision=precision, scale=scale))
    return f.finish()


# ---------------------------------------------------------------- forms
@pytest.mark.parametrize("n,form", [(1, kgo.FORM_RANGE), (2, kgo.FORM_SMALL), (8, kgo.FORM_SMALL),
                                    (9, kgo.FORM_EXACT), (4096, kgo.FORM_EXACT)])
def test_form_cutoffs_sparse_keys(n, form):
    keys = (np.arange(n, dtype=np.int64) * 1_000_003) ** 2 % (1 << 61)
    f = build(I64, keys)
    assert f.info()["form"] == form
    if form != kgo.FORM_RANGE:
        assert sorted(f.members()) == sorted(set(int(k) for k in keys))


def test_past_the_exact_cap_a_bloom_of_the_keys_seen():
    keys = np.arange(4097, dtype=np.int64) * 7919 + 11
    f = build(I64, keys)
    i = f.info()
    assert i["form"] == kgo.FORM_BLOOM and i["values"] == -1
    # bloom() (filter.rs:606-624): sized for the rows, the oracle's SBBF bits
    h = np.array([kgo.xxh64(struct.pack("<q", int(k))) for k in keys], dtype=np.uint64)
    kept, words = kgo.sbbf_build(h, 0, 0.01)
    assert kept and np.array_equal(f.bloom_words(), words)


def test_dense_integers_are_the_range():
    f = build(I64, np.arange(-5, 1000, dtype=np.int64))
    i = f.info()
    assert i["form"] == kgo.FORM_RANGE and (i["lo"], i["hi"]) == (-5, 999)


@pytest.mark.parametrize("nkeys,span,dense", [
    (1028, 7200, True),     # member.rs:52: q34's 1,028 of 7,200 in 904 B
    (314, 73049, True),     # member.rs:52: 314 dates of 73,049 in 9.1 KiB
    (20, 1_000_000, False),  # 125,000 B of bits against 20 * 9 * 16
    (300, (1 << 20) + 1, False),  # past DENSE_MAX_BITS
])
def test_dense_bit_set_size_rule(nkeys, span, dense):
    rng = np.random.default_rng(nkeys)
    inner = rng.choice(np.arange(1, span - 1), nkeys - 2, replace=False)
    keys = np.concatenate([[0, span - 1], inner]).astype(np.int64) + 12345
    f = build(I64, keys)
    assert f.info()["form"] == (kgo.FORM_DENSE if dense else kgo.FORM_EXACT)
    assert sorted(f.members()) == sorted(int(k) for k in keys)


def test_expected_fpp_rule_drops_a_small_bloom():
    opts = FilterOpts.default(bloom_max_bytes=1024)
    f = build(I64, np.arange(50_000, dtype=np.int64) * 3, opts=opts)
    assert f.info()["form"] == kgo.FORM_RANGE
    opts = FilterOpts.default(bloom_max_items=10_000)
    f = build(I64, np.arange(50_000, dtype=np.int64) * 3, opts=opts)
    assert f.info()["form"] == kgo.FORM_RANGE  # hashes dropped past the cap


# ---------------------------------------------------------------- evaluation
@pytest.mark.parametrize("ctype", [kgo.T_INT8, kgo.T_INT16, kgo.T_INT32, kgo.T_INT64,
                                   kgo.T_DATE32, kgo.T_FLOAT64, kgo.T_FLOAT32])
@pytest.mark.parametrize("nkeys", [5, 300, 6000])
def test_every_form_keeps_its_keys_and_exact_forms_nothing_else(ctype, nkeys):
    rng = np.random.default_rng(nkeys + ctype)
    lo, hi = (-120, 120) if ctype == kgo.T_INT8 else (-30000, 30000)
    if ctype == kgo.T_INT8:
        nkeys = min(nkeys, 200)
    keys = rng.choice(np.arange(lo, hi), nkeys, replace=False)
    if ctype in (kgo.T_FLOAT64, kgo.T_FLOAT32):
        keys = keys / 4.0
    valid = rng.random(nkeys) > 0.05
    f = build(ctype, keys, valid)
    probe = np.concatenate([keys, rng.integers(lo, hi, 5000) / (4.0 if ctype in (
        kgo.T_FLOAT64, kgo.T_FLOAT32) else 1)])
    pv = rng.random(len(probe)) > 0.1
    got = f.evaluate(Col(ctype, probe, pv))
    stored = set(np.asarray(keys)[valid].tolist())
    want = np.array([ok and (p in stored) for p, ok in zip(probe.tolist(), pv)])
    form = f.info()["form"]
    nulls_pass = f.info()["nulls_pass"]
    assert nulls_pass == (not valid.all())
    assert np.array_equal(got[~pv], np.full((~pv).sum(), nulls_pass))
    assert np.all(got[want])  # no false negative
    if form != kgo.FORM_BLOOM:
        assert np.array_equal(got[pv], want[pv])


def test_negative_zero_and_nan_keys():
    f = build(kgo.T_FLOAT64, np.array([-0.0, 1.5, np.nan]))
    i = f.info()
    assert not i["has_range"]  # a NaN key drops the range
    got = f.evaluate(Col(kgo.T_FLOAT64, np.array([0.0, -0.0, np.nan, 2.0, 1.5])))
    assert got.tolist() == [True, True, True, False, True]


def test_empty_build():
    f = build(I64, np.array([1, 2], dtype=np.int64), np.array([False, False]))
    i = f.info()
    assert i["empty"] and i["nulls_pass"] and i["keys"] == 0
    got = f.evaluate(Col(I64, np.array([1, 2, 3]), np.array([True, False, True])))
    assert got.tolist() == [False, True, False]


def test_widened_probe_hashes_at_the_key_width():
    """An int32 probe against an int64 key hashes 8 bytes (key.rs:290-300):
    a Bloom keeps every stored key."""
    keys = np.arange(10_000, dtype=np.int64) * 37 - 100_000
    f = build(I64, keys)
    assert f.info()["form"] == kgo.FORM_BLOOM
    assert f.evaluate(Col(kgo.T_INT32, keys.astype(np.int32))).all()
    assert f.evaluate(Col(kgo.T_FLOAT64, keys.astype(np.float64))) is None  # no widening


def test_decimal_keys_and_rescaled_probes():
    keys = [12345, -99999, 10 ** 20, 7]
    f = build(kgo.T_DECIMAL128, keys, precision=30, scale=2)
    assert f.info()["form"] == kgo.FORM_SMALL
    probe = Col(kgo.T_DECIMAL64, np.array([1234, 5, 1, -99999 // 10 - 1]), precision=10, scale=1)
    got = f.evaluate(probe)  # 123.4 == 123.40, 0.5 != 0.07
    assert got.tolist() == [False, False, False, False]
    probe = Col(kgo.T_DECIMAL128, [12345, 7, 8, 10 ** 20], precision=30, scale=2)
    assert f.evaluate(probe).tolist() == [True, True, False, True]
    probe = Col(kgo.T_DECIMAL64, np.array([12340, 70, 71]), precision=12, scale=3)
    assert f.evaluate(probe) is None  # a narrower scale does not widen


# ---------------------------------------------------------------- codec
def header(tag, extra=b""):
    return bytes([1, 1, 1, tag]) + extra


def test_encoding_of_a_small_int64_set_assembled_by_hand():
    f = build(I64, np.array([9, -3, 9, 5], dtype=np.int64), np.array([1, 1, 1, 0], bool))
    want = header(5)
    want += bytes([1])                      # nulls_pass
    want += struct.pack("<Q", 3)            # keys stored
    want += bytes([1]) + struct.pack("<qq", -3, 9)
    want += bytes([1]) + struct.pack("<I", 2)
    for v in (-3, 9):
        want += bytes([0]) + v.to_bytes(16, "little", signed=True)
    want += bytes([1]) + struct.pack("<I", 2) + struct.pack("<qq", -3, 9)
    assert f.encode() == want


def test_encoding_of_a_decimal_range_and_a_bloom_by_hand():
    opts = FilterOpts.default(max_values=0, exact_max_values=0)
    keys = [10 ** 30 + k * 977 for k in range(40)]
    f = build(kgo.T_DECIMAL128, keys, opts=opts, precision=38, scale=4)
    i = f.info()
    assert i["form"] == kgo.FORM_BLOOM
    bits = f.bloom_words().tobytes()
    want = header(8, bytes([38, 4])) + bytes([0]) + struct.pack("<Q", 40)
    want += bytes([2]) + min(keys).to_bytes(16, "little", signed=True)
    want += max(keys).to_bytes(16, "little", signed=True)
    want += bytes([0])                              # no IN list
    want += bytes([3]) + struct.pack("<I", len(bits)) + bits
    assert f.encode() == want


def test_encoding_of_a_float_range_and_timestamp_tags():
    f = build(kgo.T_FLOAT32, np.array([2.5, -1.0], dtype=np.float32),
              opts=FilterOpts.default(max_values=0))
    e = f.encode()
    assert e[:4] == header(6)
    assert e[5 + 8:5 + 8 + 17] == bytes([3]) + struct.pack("<dd", -1.0, 2.5)
    # members: float32 bit patterns as 8-byte integers (codec.rs:252-256)
    tail = bytes([1]) + struct.pack("<I", 2) + struct.pack("<qq", *sorted(
        struct.unpack("<I", struct.pack("<f", x))[0] for x in (2.5, -1.0)))
    assert e.endswith(tail)
    t = build(kgo.T_TIMESTAMP, np.array([1, 2], dtype=np.int64)).encode()
    assert t[:5] == header(10, bytes([0]))


def test_round_trips_and_dense_wire_form():
    rng = np.random.default_rng(5)
    for keys in (np.arange(20, dtype=np.int64) * 3,                  # dense bit set
                 rng.integers(-(1 << 60), 1 << 60, 700),              # hash set
                 rng.integers(0, 1 << 40, 9000),                      # Bloom
                 np.arange(100, dtype=np.int64)):                     # range
        f = build(I64, keys)
        e = f.encode()
        g = KeyFilter.decode(e)
        assert g.encode() == e
        assert g.info() == f.info()
        probe = Col(I64, rng.integers(-5, 200, 3000))
        assert np.array_equal(g.evaluate(probe), f.evaluate(probe))
    # a dense set travels as its members (codec.rs:268-273)
    f = build(I64, np.arange(20, dtype=np.int64) * 3)
    assert f.info()["form"] == kgo.FORM_DENSE
    assert f.encode().endswith(bytes([2]) + struct.pack("<I", 20)
                               + struct.pack("<20q", *range(0, 60, 3)))


def test_decode_refuses_what_this_format_did_not_write():
    f = build(I64, np.array([1, 5, 9], dtype=np.int64))
    e = f.encode()
    for n in range(len(e)):
        with pytest.raises(ValueError):
            KeyFilter.decode(e[:n])
    with pytest.raises(ValueError, match="trailing"):
        KeyFilter.decode(e + b"\0")
    for at in range(3):
        b = bytearray(e)
        b[at] ^= 0xFF
        with pytest.raises(ValueError, match="version"):
            KeyFilter.decode(bytes(b))
    # a small set of nine members (form_tests.rs:525-537)
    small = build(I64, np.arange(8, dtype=np.int64) * 5)
    b = bytearray(small.encode())
    cut = len(b) - (4 + 8 * 8)
    b = b[:cut] + struct.pack("<I", 9) + struct.pack("<9q", *range(0, 45, 5))
    with pytest.raises(ValueError, match="small set"):
        KeyFilter.decode(bytes(b))
    with pytest.raises(ValueError):
        KeyFilter.decode(b"")


# ---------------------------------------------------------------- merge
def test_merge_of_sets_is_the_filter_of_the_union():
    rng = np.random.default_rng(8)
    a = rng.integers(0, 1 << 50, 1500)
    b = rng.integers(0, 1 << 50, 1500)
    m = build(I64, a).merge(build(I64, b))
    u = build(I64, np.concatenate([a, b]))
    assert m.info()["form"] == u.info()["form"] == kgo.FORM_EXACT
    assert m.members() == u.members()
    assert m.info()["keys"] == 3000


def test_merge_of_planned_blooms_is_their_or():
    opts = FilterOpts.default(bloom_bytes=64 << 10, exact_max_values=100)
    rng = np.random.default_rng(9)
    a, b = rng.integers(0, 1 << 50, 5000), rng.integers(0, 1 << 50, 5000)
    fa, fb = build(I64, a, opts=opts), build(I64, b, opts=opts)
    m = fa.merge(fb, opts)
    assert m.info()["form"] == kgo.FORM_BLOOM
    assert np.array_equal(m.bloom_words(), fa.bloom_words() | fb.bloom_words())
    u = build(I64, np.concatenate([a, b]), opts=opts)
    assert np.array_equal(m.bloom_words(), u.bloom_words())
    # a set merged into a planned Bloom: its members' hashes added
    s = build(I64, np.array([3, 1 << 55], dtype=np.int64), opts=opts)
    ms = fa.merge(s, opts)
    assert ms.evaluate(Col(I64, np.array([3, 1 << 55]))).all()
    # unplanned Blooms do not merge: the range
    plain = FilterOpts.default(exact_max_values=100)
    m2 = build(I64, a, opts=plain).merge(build(I64, b, opts=plain), plain)
    assert m2.info()["form"] == kgo.FORM_RANGE and m2.info()["has_range"]


def test_merge_rules_empty_range_and_overfull_bloom():
    e = build(I64, np.array([0], dtype=np.int64), np.array([False]))
    f = build(I64, np.array([4, 6], dtype=np.int64))
    m = e.merge(f)
    assert m.members() == [4, 6] and m.info()["nulls_pass"] and m.info()["keys"] == 2
    r = build(I64, np.arange(10, dtype=np.int64))          # dense: the range form
    m = r.merge(f)
    assert m.info()["form"] == kgo.FORM_RANGE and (m.info()["lo"], m.info()["hi"]) == (0, 9)
    # a merged Bloom passing more than bloom_max_fpp falls back to the range
    opts = FilterOpts.default(bloom_bytes=256, exact_max_values=10, bloom_max_fpp=1.0)
    rng = np.random.default_rng(2)
    a = build(I64, rng.integers(0, 1 << 50, 200), opts=opts)
    b = build(I64, rng.integers(0, 1 << 50, 200), opts=opts)
    tight = FilterOpts.default(bloom_bytes=256, exact_max_values=10, bloom_max_fpp=0.05)
    assert a.merge(b, opts).info()["form"] == kgo.FORM_BLOOM
    assert a.merge(b, tight).info()["form"] == kgo.FORM_RANGE
    with pytest.raises(ValueError, match="different key types"):
        f.merge(build(kgo.T_INT32, np.array([1], dtype=np.int32)))


def test_pass_rate_matches_random_probes():
    rng = np.random.default_rng(3)
    top = np.iinfo(np.uint64).max
    _, words = kgo.sbbf_build(rng.integers(0, top, 3000, dtype=np.uint64), 4096, 0.01,
                              max_fpp=1.0)
    rate = kgo.sbbf_pass_rate(words)
    probes = rng.integers(0, top, 400_000, dtype=np.uint64)
    seen = kgo.sbbf_check_batch(words, probes).mean()
    assert abs(seen - rate) < 4 * np.sqrt(rate / 400_000) + 1e-4
    assert kgo.sbbf_pass_rate(np.full(64, 0xFFFFFFFF, np.uint32)) == 1.0
    assert kgo.sbbf_pass_rate(np.zeros(64, np.uint32)) == 0.0


def test_builder_merge_equals_one_builder():
    rng = np.random.default_rng(4)
    parts = [rng.integers(-1000, 1000, 3000) for _ in range(3)]
    one = build(I64, np.concatenate(parts))
    acc = KeyFilter(I64).add(Col(I64, parts[0]))
    for p in parts[1:]:
        acc.merge_builder(KeyFilter(I64).add(Col(I64, p)))
    acc.finish()
    assert acc.encode() == one.encode()


# ---------------------------------------------------------------- row stage
def test_row_stage_orders_narrows_and_switches_off():
    rng = np.random.default_rng(6)
    n = 40_000
    a = rng.integers(0, 1_000_000, n)          # column 0: a Bloom filter on it
    b = rng.integers(0, 100, n)                # column 1: a range filter on it
    bloom_keys = rng.integers(0, 1_000_000, 20_000)
    fb = build(I64, bloom_keys)
    assert fb.info()["form"] == kgo.FORM_BLOOM
    fr = build(I64, np.arange(0, 10, dtype=np.int64))     # dense: the range, cost 0.5
    stage = kgo.RowStage(min_rows=65_536, max_pass_rate=0.9)
    stage.set(0, fb, 0)
    stage.set(1, fr, 1)
    cols = [Col(I64, a), Col(I64, b)]
    got = stage.apply(cols)
    pb, pr = fb.evaluate(cols[0]), fr.evaluate(cols[1])
    assert np.array_equal(got, pb & pr)
    # the range ran first on every row; the Bloom saw only its ~10% survivors
    assert stage.stats(1) == (n, n - pr.sum(), False)
    assert stage.stats(0) == (pr.sum(), pr.sum() - (pb & pr).sum(), False)
    # a filter passing every row is switched off once a window is measured
    allkeys = build(I64, np.arange(0, 100, dtype=np.int64))
    s2 = kgo.RowStage(min_rows=65_536, max_pass_rate=0.9)
    s2.set(0, allkeys, 1)
    s2.set(1, fr, 1)
    for _ in range(2):
        s2.apply(cols)
    assert s2.stats(0)[2] and not s2.stats(1)[2]
    assert s2.stats(0)[0] == 2 * n
    s2.apply(cols)
    assert s2.stats(0)[0] == 2 * n  # off: not tested any more
    s2.start_partition()
    s2.apply(cols)
    assert s2.stats(0)[0] == 3 * n and not s2.stats(0)[2]
    # a new generation turns it back on
    s2.set(0, allkeys, 1)
    assert s2.apply([Col(I64, a[:10]), Col(I64, b[:10])]) is not None
    # nothing filtered: None
    s3 = kgo.RowStage()
    s3.set(0, allkeys, 1)
    assert s3.apply(cols) is None
    s3.set(1, None, 0)  # pending: applies nothing
    assert s3.apply(cols) is None
