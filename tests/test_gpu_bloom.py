"""GPU parity of the runtime Bloom filter: the device-built split-block bit
array is byte-identical to the oracle's (and to the Parquet specification
vector); the device probe selects exactly the rows the oracle's check
passes, in input order."""
import hashlib
import math
import json
import os
import struct

import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu
V = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                "vectors.json")))


def test_parquet_spec_vector_on_device():
    vals = np.array([i * 7919 - 3 for i in range(500)], dtype=np.int64)
    b = kgpu.RuntimeBloom(ctx(), 1024)
    b.insert(Col(A.T_INT64, vals))
    assert b.to_bytes() == bytes.fromhex(V["sbbf_1024"]["bytes_hex"])


def test_config3_filter_size_and_digest():
    say("bloom: 750K keys at fpp 0.03")
    nbytes = kgpu.RuntimeBloom.bytes_for(750000, 0.03)
    assert nbytes == 723968
    keys = np.arange(750000, dtype=np.int64) * 20 + 3
    b = kgpu.RuntimeBloom(ctx(), nbytes)
    b.insert(Col(A.T_INT64, keys))
    assert hashlib.sha256(b.to_bytes()).hexdigest() == V["sbbf_config3"]["sha256"]
    assert b.expected_fpp(750000) == kgo.lib().kgo_sbbf_expected_fpp(nbytes // 32, 750000)


def test_build_with_predicate_and_probe_matches_oracle():
    """Config-3 shape at 1/100: build keys kept by attr < 5, probe column
    over the full key range; survivors = the oracle's check, in order."""
    say("bloom: build with selection, probe 6M rows")
    nb, npr = 150_000, 6_000_000
    key, _ = kgo.gen_join_build(4, 0, nb, nb, with_val=False)
    attr = (np.arange(nb) * 7919 % 100).astype(np.int32)
    keep = attr < 5
    kept = key[keep]
    nbytes = kgpu.RuntimeBloom.bytes_for(len(kept), 0.03)
    b = kgpu.RuntimeBloom(ctx(), nbytes)
    b.insert(Col(A.T_INT64, key), Col(A.T_BOOL, keep))
    h = kgo.hash_columns([kgo.Col(kgo.T_INT64, kept)])
    ok, words = kgo.sbbf_build(h, nbytes=nbytes)
    assert b.to_bytes() == words.tobytes()
    fk, fkv, _ = kgo.gen_join_probe(5, 0, npr, nb, 10)
    rows, cnt = b.probe(Col(A.T_INT64, fk, fkv))
    got = rows.to_numpy(np.int32, cnt)
    ph = kgo.hash_columns([kgo.Col(kgo.T_INT64, fk)])
    L = kgo.lib()
    nblk = len(words) // 8
    # vectorized oracle check over the same words (the per-hash C check on a sample)
    blocks = ((ph >> np.uint64(32)) * np.uint64(nblk)) >> np.uint64(32)
    salt = np.array([0x47b6137b, 0x44974d91, 0x8824ad5b, 0xa2b7289d, 0x705495c7, 0x2df1424b,
                     0x9efc4947, 0x5c6bfb31], dtype=np.uint64)
    key32 = (ph & np.uint64(0xFFFFFFFF))
    bits = ((key32[:, None] * salt[None, :]) & np.uint64(0xFFFFFFFF)) >> np.uint64(27)
    wv = words.reshape(-1, 8)[blocks.astype(np.int64)].astype(np.uint64)
    passes = np.all((wv >> bits) & np.uint64(1), axis=1) & fkv
    for i in np.random.default_rng(0).integers(0, npr, 2000):
        assert bool(passes[i]) == (bool(fkv[i]) and bool(L.kgo_sbbf_check(
            words.ctypes.data, nblk, int(ph[i]))))
    want = np.nonzero(passes)[0]
    assert np.array_equal(got, want)
    # no false negatives: every probe key that is a kept build key passes
    kept_set = np.zeros(nb, dtype=bool)
    kept_set[kept] = True
    true_hits = np.nonzero(fkv & kept_set[fk])[0]
    assert np.all(np.isin(true_hits, got))


def test_probe_edge_sizes():
    b = kgpu.RuntimeBloom(ctx(), 4096)
    b.insert(Col(A.T_INT64, np.arange(0, 3000, 3, dtype=np.int64)))
    for n in (0, 1, 63, 64, 65, 8191, 8192, 8193, 20000):
        keys = np.arange(n, dtype=np.int64)
        rows, cnt = b.probe(Col(A.T_INT64, keys))
        got = rows.to_numpy(np.int32, cnt)
        assert np.all(np.diff(got) > 0)
        assert np.all(np.isin(np.arange(0, min(n, 3000), 3), got))


def test_config3_pipeline_small():
    """Configuration 3 at 1/100: build keys kept by attr < 5 (range row stage),
    an SBBF of the kept keys at fpp 0.03, the probe column pre-screened by it,
    the survivors joined (selection-vector probe) and SUM(payload), COUNT(*)
    aggregated on the device; equal to the oracle's join of every probe row."""
    say("bloom: config-3 pipeline, 150K build, 6M probe")
    nb, npr = 150_000, 6_000_000
    key, _ = kgo.gen_join_build(4, 0, nb, nb, with_val=False)
    attr = (np.random.default_rng(4).integers(0, 100, nb)).astype(np.int32)
    fk, fkv, pay = kgo.gen_join_probe(5, 0, npr, nb, 0)
    c = ctx()
    rows, cnt = kgpu.select_range(c, Col(A.T_INT32, attr), 0, 4)
    kept = kgpu.gather(c, [Col(A.T_INT64, key)], rows, cnt)
    assert cnt == int(np.sum(attr < 5))
    assert np.array_equal(kept.column(0)[0], key[attr < 5])
    bloom = kgpu.RuntimeBloom(c, kgpu.RuntimeBloom.bytes_for(cnt, 0.03))
    bloom.insert(kept.as_col(0), nrows=cnt)
    j = kgpu.HashJoin(c, A.JOIN_INNER, [0], [0], expected_build_rows=cnt)
    j.build_sink([kept.as_col(0)], cnt)
    j.build_finalize()
    fk_d = kgpu.DeviceBuffer.from_numpy(c, fk)
    pay_d = kgpu.DeviceBuffer.from_numpy(c, pay)
    probe_cols = [Col(A.T_INT64, fk_d, length=npr), Col(A.T_INT64, pay_d, length=npr)]
    surv, ns = bloom.probe(probe_cols[0])
    out = j.probe_selected(probe_cols, surv, ns, npr)
    g = kgpu.HashAggregate(c, None, [(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
    g.sink(None, [out.as_col(1)], out.num_rows)
    g.finalize()
    r = g.result()
    keep = np.zeros(nb, dtype=bool)
    keep[key[attr < 5]] = True
    hit = keep[fk]
    assert int(r.column(0)[0][0]) == int(pay[hit].sum())
    assert int(r.column(1)[0][0]) == int(hit.sum())
    assert ns >= hit.sum() and ns < 0.15 * npr  # 5 % true + about 3 % false positives


@pytest.mark.parametrize("t,width", [(A.T_INT64, 8), (A.T_INT32, 4)])
def test_spark_layout_filter_properties(t, width):
    """Spark's BloomFilterImpl layout (optional, property-only: no Spark-built
    filter bytes exist to pin the layout against).  Sizing follows Spark's
    closed forms (optimalNumOfBits, optimalNumOfHashFunctions); every kept
    non-NULL build key passes the probe (no false negatives), NULL probe keys
    never pass, survivors come back in input order, and the rate at which
    keys that were never inserted pass is within 30 % of (1 - e^(-kn/m))^k."""
    say(f"bloom: Spark layout, {width}-byte keys")
    rng = np.random.default_rng(width)
    nb, npr = 40_000, 2_000_000
    dt = np.int64 if width == 8 else np.int32
    lim = (1 << 40) if width == 8 else (1 << 30)
    key = np.unique(rng.integers(-lim, lim, nb).astype(dt))
    nb = len(key)
    kv = rng.random(nb) > 0.02
    keep = rng.random(nb) > 0.5
    n_in = int((kv & keep).sum())
    bits = kgpu.RuntimeBloom.spark_bits_for(n_in, 0.03)
    k = kgpu.RuntimeBloom.spark_hashes_for(n_in, bits)
    assert abs(bits - (-n_in * math.log(0.03) / math.log(2) ** 2)) < 1
    assert k == max(1, round(bits / n_in * math.log(2)))
    b = kgpu.RuntimeBloom(ctx(), spark_bits=bits, spark_hashes=k)
    b.insert(Col(t, key, kv), Col(A.T_BOOL, keep))
    words = np.frombuffer(b.to_bytes(), dtype=np.uint64)
    assert len(words) == (bits + 63) // 64
    members = key[kv & keep]
    others = np.setdiff1d(rng.integers(-lim, lim, npr).astype(dt), key)
    pk = np.concatenate([members, others])
    pkv = np.ones(len(pk), dtype=bool)
    pkv[::97] = False
    rows, cnt = b.probe(Col(t, pk, pkv))
    got = rows.to_numpy(np.int32, cnt)
    assert np.all(np.diff(got) > 0)
    assert not np.isin(np.nonzero(~pkv)[0], got).any()
    mem_rows = np.nonzero(pkv[:len(members)])[0]
    assert np.isin(mem_rows, got).all()
    fp = np.count_nonzero(got >= len(members)) / max(1, int(pkv[len(members):].sum()))
    want = (1 - math.exp(-k * n_in / (len(words) * 64))) ** k
    assert 0.7 * want < fp < 1.3 * want
