"""GPU parity of the batched join-aggregate server (kg_server_*): every
request's (key, SUM(pay), COUNT(*)) equals the oracle's inner join of its
build and probe columns aggregated per build key, whatever batch it lands
in, with own and resident build tables, graph replay on or off, small
requests batched and large ones on the partitioned path."""
import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import DeviceBuffer

pytestmark = pytest.mark.gpu


def oracle_request(bkeys, fk, fkv, pay):
    """{key: (sum, count)}: the oracle's inner join, then its aggregate."""
    pi, bi = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, fk, fkv)],
                      [kgo.Col(kgo.T_INT64, bkeys)])
    if len(pi) == 0:
        return {}
    a = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)])
    a.update(kgo.Col(kgo.T_INT64, bkeys[bi]), [kgo.Col(kgo.T_INT64, pay[pi])])
    rk, rkv, res = a.finish()
    (sv, sok, _), (cv, _, _) = res
    return {int(k): (int(s), int(c)) for k, s, c in zip(rk, sv, cv)}


def got_request(batch):
    k, _ = batch.column(0)
    s, _ = batch.column(1)
    c, _ = batch.column(2)
    return {int(a): (int(b), int(d)) for a, b, d in zip(k, s, c)}


def make_request(rng, nb, dup=False):
    keys = rng.choice(np.int64(1) << 40, nb, replace=False).astype(np.int64) - (1 << 39)
    if dup:
        keys[: nb // 4] = keys[nb // 4: nb // 2]  # duplicate build keys: multiplicity
        keys[0] = np.iinfo(np.int64).min  # the table's sentinel value
    np_ = nb * 10
    span = np.concatenate([keys, rng.integers(-(1 << 39), 1 << 39, nb // 10 + 1)])
    fk = span[rng.integers(0, len(span), np_)]
    fkv = rng.random(np_) > 0.01
    pay = rng.integers(-(1 << 40), 1 << 40, np_)
    return keys, fk, fkv, pay


@pytest.mark.parametrize("use_graph,workers", [(True, 1), (False, 1), (True, 2)])
def test_mixed_batch_own_and_resident_tables(use_graph, workers, monkeypatch):
    say(f"serve: mixed sizes, resident tables, graph={use_graph}, large workers={workers}")
    rng = np.random.default_rng(8 + use_graph)
    c = ctx()
    monkeypatch.setenv("KG_SERVE_LARGE_WORKERS", str(workers))
    srv = kgpu.JoinAggServer(c, arena_bytes=1 << 30, max_batch=16, use_graph=use_graph)
    # two resident build tables
    res_keys, res_dev = [], []
    for nb in (3000, 50000):
        k = rng.choice(np.int64(1) << 36, nb, replace=False).astype(np.int64)
        d = DeviceBuffer.from_numpy(c, k)
        res_keys.append(k)
        res_dev.append((srv.add_table(d, nb), d))
    keep, want = [], {}
    # requests 9, 11, 12, 14 and 17 probe 4.2M to 5M rows: they run through
    # the partitioned fused aggregate on the large-request workers (own
    # builds and resident tables; 11 and 17 share resident table 1, so the
    # workers must take them one after the other)
    sizes = [1000, 4096, 1, 20000, 250, 70000, 9000, 1000, 3, 450_000, 2000, 500_000,
             480_000, 100, 420_000, 7, 30, 460_000]
    for rid, nb in enumerate(sizes):
        if rid % 3 == 2:
            t = rid % 2
            bkeys = res_keys[t]
            np_ = 10 * max(nb, 1)
            fk = bkeys[rng.integers(0, len(bkeys), np_)]
            fk[::7] += 1 << 37  # misses
            fkv = rng.random(np_) > 0.02
            pay = rng.integers(-(1 << 40), 1 << 40, np_)
            dk, dv, dp = (DeviceBuffer.from_numpy(c, fk),
                          DeviceBuffer.from_numpy(c, np.packbits(fkv, bitorder="little")),
                          DeviceBuffer.from_numpy(c, pay))
            keep += [dk, dv, dp]
            srv.submit(rid, dk, dp, np_, table=res_dev[t][0], probe_valid=dv)
        else:
            bkeys, fk, fkv, pay = make_request(rng, nb, dup=(rid == 3))
            db = DeviceBuffer.from_numpy(c, bkeys)
            dk, dv, dp = (DeviceBuffer.from_numpy(c, fk),
                          DeviceBuffer.from_numpy(c, np.packbits(fkv, bitorder="little")),
                          DeviceBuffer.from_numpy(c, pay))
            keep += [db, dk, dv, dp]
            srv.submit(rid, dk, dp, len(fk), build_keys=db, build_rows=nb, probe_valid=dv)
        want[rid] = oracle_request(bkeys, fk, fkv, pay)
        if rid == 4:  # some requests finish before the rest arrive
            done = srv.progress(wait=True)
            assert all(r <= 4 for r in done)
    while srv.pending():
        srv.progress(wait=True)
    for rid in range(len(sizes)):
        got = got_request(srv.result(rid))
        assert got == want[rid], rid
    srv.close()


def test_admission_splits_batches_and_refuses_oversize():
    say("serve: admission by arena bytes")
    rng = np.random.default_rng(3)
    c = ctx()
    srv = kgpu.JoinAggServer(c, arena_bytes=64 << 20, max_batch=64)
    keep, want = [], {}
    for rid in range(12):
        bkeys, fk, fkv, pay = make_request(rng, 40000)
        bufs = [DeviceBuffer.from_numpy(c, x) for x in
                (bkeys, fk, np.packbits(fkv, bitorder="little"), pay)]
        keep += bufs
        srv.submit(rid, bufs[1], bufs[3], len(fk), build_keys=bufs[0], build_rows=40000,
                   probe_valid=bufs[2])
        want[rid] = oracle_request(bkeys, fk, fkv, pay)
    done = []
    while len(done) < 12:
        done += srv.progress(wait=True)
    for rid in range(12):
        assert got_request(srv.result(rid)) == want[rid]
    big = DeviceBuffer(c, 8 * 8)
    with pytest.raises(kgpu.KgError, match="arena"):
        srv.submit(99, big, big, 8, build_keys=big, build_rows=4_000_000)
    srv.close()


def test_every_id_reported_once_through_a_one_id_buffer():
    """A batch retires more requests than the caller's id buffer holds: the
    ids past it are reported by the next calls, none is lost, and the
    server counts them as pending until they are."""
    say("serve: ids through a one-id buffer")
    rng = np.random.default_rng(5)
    c = ctx()
    srv = kgpu.JoinAggServer(c, arena_bytes=256 << 20, max_batch=16)
    keep, want = [], {}
    for rid in range(10):
        bkeys, fk, fkv, pay = make_request(rng, 2000)
        bufs = [DeviceBuffer.from_numpy(c, x) for x in
                (bkeys, fk, np.packbits(fkv, bitorder="little"), pay)]
        keep += bufs
        srv.submit(rid, bufs[1], bufs[3], len(fk), build_keys=bufs[0], build_rows=2000,
                   probe_valid=bufs[2])
        want[rid] = oracle_request(bkeys, fk, fkv, pay)
    seen = []
    for _ in range(200):
        if not srv.pending():
            break
        ids = srv.progress(wait=True, max_ids=1)
        assert len(ids) <= 1
        seen += ids
    assert sorted(seen) == list(range(10)), seen
    assert srv.pending() == 0
    for rid in range(10):
        assert got_request(srv.result(rid)) == want[rid]
    srv.close()
