"""The join / fused-aggregate shape autotuner (kg_tune_*, kg_autotune_joinagg):
every candidate shape (table load factor, staged-scatter rows per lane,
merge workgroup) gives the oracle's aggregate bit for bit, the tuner times
all of them and keeps one, and malformed shapes are refused."""
import numpy as np
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu

NB, NPR = 300_000, 1_300_000  # a probe batch of the partitioned path


def _oracle(bk, pk, pkv, amt):
    op, ob = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, pk, pkv)], [kgo.Col(kgo.T_INT64, bk)])
    o = kgo.Agg((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)])
    o.update(kgo.Col(kgo.T_INT64, bk[ob]), [kgo.Col(kgo.T_INT64, amt[op])])
    rk, _, res = o.finish()
    (sv, _, _), (cv, _, _) = res
    return {int(k): (int(s), int(c)) for k, s, c in zip(rk, sv, cv)}


def _gpu(bk, pk, pkv, amt):
    j = kgpu.HashJoin(ctx(), A.JOIN_INNER, [0], [0], expected_build_rows=len(bk))
    j.build_sink([Col(A.T_INT64, bk)])
    j.build_finalize()
    j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
    j.probe_aggregate([Col(A.T_INT64, pk, pkv), Col(A.T_INT64, amt)], [1])
    r = j.aggregate_result()
    k, _ = r.column(0)
    s, _ = r.column(1)
    c, _ = r.column(2)
    out = {int(a): (int(b), int(d)) for a, b, d in zip(k, s, c)}
    j.close()
    return out


def test_every_shape_gives_the_oracle_result():
    say("tune: every shape, same aggregate")
    rng = np.random.default_rng(21)
    bk = rng.permutation(NB).astype(np.int64) * 7 - 1000
    pk = bk[rng.integers(0, NB, NPR)]
    pk[::9] += 3  # misses
    pkv = rng.random(NPR) > 0.01
    amt = rng.integers(-(1 << 40), 1 << 40, NPR)
    want = _oracle(bk, pk, pkv, amt)
    c = ctx()
    try:
        for load in (25, 50, 75):
            for rows in (4, 12):
                for lanes in (512, 1024):
                    say(f"  load {load} %, {rows} rows a lane, {lanes}-lane merge")
                    kgpu.tune_set(c, NB, load, rows, lanes)
                    assert _gpu(bk, pk, pkv, amt) == want, (load, rows, lanes)
    finally:
        kgpu.tune_set(c, NB)  # the class's default again for the other tests


def test_autotune_times_every_candidate_and_keeps_one():
    say("tune: autotune")
    c = ctx()
    try:
        best, ms, n = kgpu.autotune_joinagg(c, 200_000, 3_000_000)
        assert n == 18 and ms > 0
        assert best["load_factor_pct"] in (25, 50, 75)
        assert best["scatter_rows"] in (4, 8, 12)
        assert best["merge_lanes"] in (512, 1024)
        assert kgpu.tune_get(c, 200_000) == best
        assert kgpu.tune_get(c, 150_000) == best  # the same size class (2^18)
    finally:
        kgpu.tune_set(c, 200_000)
    # back to the untuned default of the class (2^18 rows: 25 %), not a tuned 50 %
    assert kgpu.tune_get(c, 200_000) == {"load_factor_pct": 25, "scatter_rows": 0,
                                         "merge_lanes": 1024}
    for bad in ({"load_factor_pct": 5}, {"scatter_rows": 6}, {"merge_lanes": 256}):
        with pytest.raises(kgpu.KgError):
            kgpu.tune_set(c, 1000, **bad)


def test_tuned_shapes_save_and_load(tmp_path):
    say("tune: save and load")
    c = ctx()
    path = str(tmp_path / "shapes.txt")
    try:
        kgpu.tune_set(c, 5000, 25, 8, 512)
        kgpu.tune_set(c, 3_000_000, 75, 4, 1024)
        kgpu.tune_save(c, path)
        kgpu.tune_set(c, 5000)
        kgpu.tune_set(c, 3_000_000)
        assert kgpu.tune_load(c, path) >= 2
        assert kgpu.tune_get(c, 5000) == {"load_factor_pct": 25, "scatter_rows": 8,
                                          "merge_lanes": 512}
        assert kgpu.tune_get(c, 3_000_000) == {"load_factor_pct": 75, "scatter_rows": 4,
                                               "merge_lanes": 1024}
        bad = tmp_path / "bad.txt"
        bad.write_text("13 50 12 1024\n14 50 twelve 1024\n")
        with pytest.raises(kgpu.KgError, match="malformed"):
            kgpu.tune_load(c, str(bad))
        bad.write_text("13 50 7 1024\n")
        with pytest.raises(kgpu.KgError):
            kgpu.tune_load(c, str(bad))
        # a class returned to its default is not saved as tuned
        kgpu.tune_set(c, 5000)
        kgpu.tune_save(c, path)
        classes = [int(line.split()[0]) for line in open(path) if line.strip()]
        assert 13 not in classes and 22 in classes, classes
    finally:
        kgpu.tune_set(c, 5000)
        kgpu.tune_set(c, 3_000_000)


def test_untuned_classes_take_the_measured_load_factors():
    say("tune: untuned defaults")
    c = ctx()
    for rows, load in ((100, 25), (7_000_000, 75), (40_000_000, 50)):
        assert kgpu.tune_get(c, rows) == {"load_factor_pct": load, "scatter_rows": 0,
                                          "merge_lanes": 1024}, rows
