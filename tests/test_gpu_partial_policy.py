"""PartialAggregate's BypassPolicy and table flushes
(crates/krishiv-exec/src/aggregate/partial.rs:20-50, 200-262): whatever the
partial phase emits -- flushed tables, rows passed through as their own
states, the table left at the end -- the final phase merging all of it gives
the single-phase answer (the oracle's)."""
import numpy as np
import pytest

from gpu_util import ctx, say
from test_gpu_agg import ALL6, T_MAP, compare
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def _policy_two_phase(key, kvalid, ktype, cols, types, aggs, batches, policy):
    """The partial aggregate with `policy` over `batches` slices, every batch
    it emits (and its final table) merged by the final aggregate.  Returns
    (oracle result, merged result, (flushes, bypassed rows))."""
    o = kgo.Agg((ktype, 0, 0), types, aggs)
    o.update(kgo.Col(ktype, key, kvalid),
             [kgo.Col(t[0], v, ok, precision=t[1], scale=t[2]) for (v, ok), t in zip(cols, types)])
    ref = o.finish()
    gtypes = [(T_MAP[t[0]], t[1], t[2]) for t in types]
    p = kgpu.HashAggregate(ctx(), (T_MAP[ktype], 0, 0), gtypes, aggs, partial=True,
                           expected_groups=1 << 10)
    p.set_bypass(*policy)
    final = kgpu.HashAggregate(ctx(), (T_MAP[ktype], 0, 0), gtypes, aggs, merge=True)
    keep = []

    def merge(batches_out):
        for st in batches_out:
            keep.append(st)
            final.sink(st.as_col(0), [st.as_col(i) for i in range(1, st.num_columns)],
                       st.num_rows)

    bounds = np.linspace(0, len(key), batches + 1).astype(np.int64)
    for b in range(batches):
        lo, hi = int(bounds[b]), int(bounds[b + 1])
        kc = Col(T_MAP[ktype], key[lo:hi], None if kvalid is None else kvalid[lo:hi])
        ic = [Col(T_MAP[t[0]], v[lo:hi], None if ok is None else ok[lo:hi], t[1], t[2])
              for (v, ok), t in zip(cols, types)]
        p.sink(kc, ic, hi - lo)
        merge(p.partial_batches())
    p.finalize()
    merge(p.partial_batches())
    last = p.result()
    merge([last])
    final.finalize()
    return ref, final.result(), p.partial_stats()


def _columns(rng, n):
    i64 = rng.integers(-(1 << 40), 1 << 40, n)
    dec = rng.integers(-(10 ** 12) + 1, 10 ** 12, n)
    f = rng.random(n) - 0.5
    oks = [rng.random(n) > 0.1 for _ in range(3)]
    types = [(kgo.T_INT64, 0, 0), (kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0)]
    aggs = [(kgo.AGG_COUNT_STAR, 0)] + [(fn, c) for c in range(3) for fn in ALL6]
    return list(zip((i64, dec, f), oks)), types, aggs


def test_bypass_passes_rows_through():
    """Nearly every key distinct: after 50,000 rows the groups reach 80 per 100
    rows, the table is flushed once and the other 150,000 rows pass through
    (NULL keys among them)."""
    say("partial policy: bypass, 200K rows of distinct keys")
    rng = np.random.default_rng(31)
    n = 200_000
    key = rng.integers(0, 1 << 40, n)
    kvalid = rng.random(n) > 0.02
    cols, types, aggs = _columns(rng, n)
    ref, r, (flushes, bypassed) = _policy_two_phase(key, kvalid, kgo.T_INT64, cols, types, aggs,
                                                    8, (50_000, 80, 16 << 20))
    assert bypassed == 150_000 and flushes == 1, (flushes, bypassed)
    compare(ref, r, aggs)


def test_flushes_keep_the_table_small():
    """4,000 keys, a table capped at 64 KiB of records: the table is flushed
    many times and never bypassed (the reduction stays high)."""
    say("partial policy: flushes, 240K rows over 4K keys")
    rng = np.random.default_rng(32)
    n = 240_000
    key = rng.integers(-2000, 2000, n)
    cols, types, aggs = _columns(rng, n)
    ref, r, (flushes, bypassed) = _policy_two_phase(key, None, kgo.T_INT64, cols, types, aggs,
                                                    12, (100_000, 80, 64 << 10))
    assert bypassed == 0 and flushes >= 6, (flushes, bypassed)
    compare(ref, r, aggs)


def test_one_byte_key_bypass_at_once():
    """An int8 key with max_group_percent 0: the first batch switches the
    aggregate to pass-through; NULL keys pass through as NULL."""
    say("partial policy: int8 key, bypass after the first batch")
    rng = np.random.default_rng(33)
    n = 20_000
    key = rng.integers(-128, 128, n).astype(np.int8)
    kvalid = rng.random(n) > 0.1
    cols, types, aggs = _columns(rng, n)
    ref, r, (flushes, bypassed) = _policy_two_phase(key, kvalid, kgo.T_INT8, cols, types, aggs,
                                                    4, (1000, 0, 16 << 20))
    assert bypassed == n - n // 4 and flushes == 1, (flushes, bypassed)
    compare(ref, r, aggs)


def test_policy_refusals():
    c = ctx()
    t = [(A.T_INT64, 0, 0)]
    full = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), t, [(A.AGG_SUM, 0)])
    with pytest.raises(kgpu.KgError, match="partial"):
        full.set_bypass()
    glob = kgpu.HashAggregate(c, None, t, [(A.AGG_SUM, 0)], partial=True)
    with pytest.raises(kgpu.KgError, match="global"):
        glob.set_bypass()
    p = kgpu.HashAggregate(c, (A.T_INT64, 0, 0), t, [(A.AGG_SUM, 0)], partial=True)
    v = np.arange(10, dtype=np.int64)
    p.sink(Col(A.T_INT64, v), [Col(A.T_INT64, v)])
    with pytest.raises(kgpu.KgError, match="before the first sink"):
        p.set_bypass()
