"""GPU parity of the grace hash join fused with the per-partition aggregate
(configuration 4's plan): bit-exact against the oracle's pair walk on small
inputs and against the per-key factorization (checked here against the pair
walk) on larger ones; the spill to pinned host memory gives the same answer."""
import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def zipf_keys(seed, n, distinct, s):
    """Zipf(s) keys over `distinct` values; both sides of a join share the
    rank-to-key permutation (seed 6), so the same keys are hot on both."""
    return kgo.gen_zipf_keys(seed, 6, 0, n, kgo.zipf_cdf(distinct, s), distinct)


def factorized(bk, bp, bv, pk, pp, pv, parts):
    """Per partition: sum over keys of nb*np pairs and Sb*Sp (wrapping)."""
    out_c = np.zeros(parts, dtype=np.int64)
    out_s = np.zeros(parts, dtype=np.uint64)
    out_v = np.zeros(parts, dtype=bool)
    ub, inv_b = np.unique(bk, return_inverse=True)
    nb = np.bincount(inv_b, minlength=len(ub)).astype(np.int64)
    nbv = np.bincount(inv_b, weights=bv.astype(float), minlength=len(ub)).astype(np.int64)
    sb = np.zeros(len(ub), dtype=np.uint64)
    np.add.at(sb, inv_b[bv], bp[bv].astype(np.uint64))
    idx = np.searchsorted(ub, pk)
    idx[idx >= len(ub)] = 0
    hit = ub[idx] == pk
    pids = kgo.route([kgo.Col(kgo.T_INT64, pk)], parts)
    np.add.at(out_c, pids[hit], nb[idx[hit]])
    m = hit & pv
    with np.errstate(over="ignore"):
        np.add.at(out_s, pids[m], pp[m].astype(np.uint64) * sb[idx[m]])
    vc = np.zeros(parts, dtype=np.int64)
    np.add.at(vc, pids[m], nbv[idx[m]])
    out_v = vc > 0
    return out_c, out_s.view(np.int64), out_v


def test_factorization_equals_pair_walk():
    """The oracle's pair walk (kgo_join) pins the factorized reference."""
    rng = np.random.default_rng(1)
    bk = zipf_keys(6, 3000, 800, 0.8)
    pk = zipf_keys(7, 9000, 800, 0.8)
    bp = rng.integers(0, 256, len(bk))
    pp = rng.integers(0, 256, len(pk))
    bv = rng.random(len(bk)) > 0.05
    pv = rng.random(len(pk)) > 0.05
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk)], [kgo.Col(kgo.T_INT64, bk)])
    pids = kgo.route([kgo.Col(kgo.T_INT64, pk[op])], 64)
    c = np.bincount(pids, minlength=64)
    both = pv[op] & bv[ob]
    s = np.zeros(64, dtype=np.int64)
    np.add.at(s, pids[both], (pp[op][both] * bp[ob][both]) % (1 << 31))
    fc, fs, fv = factorized(bk, bp, bv, pk, pp, pv, 64)
    assert np.array_equal(c, fc) and np.array_equal(s, fs)
    # and the device
    g = kgpu.GraceJoinAggregate(ctx(), 64)
    g.build_sink(Col(A.T_INT64, bk), Col(A.T_INT64, bp, bv))
    g.probe_sink(Col(A.T_INT64, pk), Col(A.T_INT64, pp, pv))
    gc, gs, gv, _ = g.finish()
    assert np.array_equal(gc, c) and np.array_equal(gs, s) and np.array_equal(gv, fv)


@pytest.mark.parametrize("budget_mb", [0, 40])
def test_config4_shape_with_and_without_spill(budget_mb):
    """1/2000 of configuration 4 (750K build, 3M probe, Zipf 0.8 over 600K
    keys, 64 partitions, several sink batches); with a 40 MB HBM budget the
    later partitioned batches spill to pinned host memory."""
    say(f"grace: config-4 shape, budget {budget_mb} MB")
    nb, npr = 750_000, 3_000_000
    bk = zipf_keys(6, nb, 600_000, 0.8)
    pk = zipf_keys(7, npr, 600_000, 0.8)
    rng = np.random.default_rng(2)
    bp = rng.integers(0, 256, nb)
    pp = rng.integers(0, 256, npr)
    kv = rng.random(npr) > 0.01  # some NULL probe keys
    want_c, want_s, want_v = factorized(bk, bp, np.ones(nb, bool), pk[kv], pp[kv],
                                        np.ones(int(kv.sum()), bool), 64)
    c = kgpu.Context(0, budget_mb << 20) if budget_mb else ctx()
    g = kgpu.GraceJoinAggregate(c, 64, table_budget=8 << 20 if budget_mb else 8 << 30)
    for a in range(0, nb, 250_000):
        g.build_sink(Col(A.T_INT64, bk[a:a + 250_000]), Col(A.T_INT64, bp[a:a + 250_000]))
    for a in range(0, npr, 1_000_000):
        g.probe_sink(Col(A.T_INT64, pk[a:a + 1_000_000], kv[a:a + 1_000_000]),
                     Col(A.T_INT64, pp[a:a + 1_000_000]))
    gc, gs, gv, spilled = g.finish()
    assert np.array_equal(gc, want_c)
    assert np.array_equal(gs, want_s)
    if budget_mb:
        assert spilled > 0
    g.close()


def test_products_that_do_not_factor_are_refused():
    g = kgpu.GraceJoinAggregate(ctx(), 8)
    g.build_sink(Col(A.T_INT64, np.array([1, 2])), Col(A.T_INT64, np.array([1 << 20, 1])))
    g.probe_sink(Col(A.T_INT64, np.array([1, 2])), Col(A.T_INT64, np.array([1 << 12, 3])))
    with pytest.raises(kgpu.KgError, match="KRISHIV_UNSUPPORTED_FEATURE"):
        g.finish()


def test_sum_mod31_reading_takes_any_payloads():
    """Configuration 4's factorable reading: SUM = ((sum of p * b, wrapping)
    mod 2^31) for payloads of any sign and size, against the oracle's
    factorization (pinned by the pair walk in the oracle tests)."""
    say("grace: mod 2^31 reading, wide payloads")
    rng = np.random.default_rng(12)
    nb, npr = 200_000, 900_000
    bk = zipf_keys(6, nb, 150_000, 0.8)
    pk = zipf_keys(7, npr, 150_000, 0.8)
    bp = rng.integers(-(1 << 40), 1 << 40, nb)
    pp = rng.integers(-(1 << 40), 1 << 40, npr)
    bv, pv = rng.random(nb) > 0.02, rng.random(npr) > 0.02
    g = kgpu.GraceJoinAggregate(ctx(), 64, sum_mod31=True)
    g.build_sink(Col(A.T_INT64, bk), Col(A.T_INT64, bp, bv))
    g.probe_sink(Col(A.T_INT64, pk), Col(A.T_INT64, pp, pv))
    gc, gs, gv, _ = g.finish()
    wc, ws, wv = kgo.grace_factorized(kgo.Col(kgo.T_INT64, bk), kgo.Col(kgo.T_INT64, bp, bv),
                                      kgo.Col(kgo.T_INT64, pk), kgo.Col(kgo.T_INT64, pp, pv), 64, 1)
    assert np.array_equal(gc, wc) and np.array_equal(gs, ws) and np.array_equal(gv, wv)
    assert (gs >= 0).all() and (gs < (1 << 31)).all()
