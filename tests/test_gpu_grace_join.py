"""The grace hash join (kg_grace_join, crates/krishiv-exec/src/join/grace.rs)
against the oracle's nested-loop-pinned join (kgo.join_ex): with a table
budget far below the build, every join type spills its partitions to host
memory and still returns the in-memory join's rows (the reference's own
check, tests/hash_join/main.rs:383-418); partitions that do not fit are
split again (levels 1..6) and small-probe semi/anti partitions are joined
with roles reversed."""
import numpy as np
import pytest

from gpu_util import ctx, say
from test_gpu_join import ALL_TYPES, VLESS
import kgo
import kgpu
from kgpu import Col
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def side(rng, n, key_range, null_frac):
    k = rng.integers(0, key_range, n)
    ok = rng.random(n) >= null_frac
    v = rng.integers(-1000, 1000, n)
    return [(k, ok), (v, np.ones(n, bool))]


def matrix(cols):
    """Rows as an int64 matrix (a validity column and a value column per
    output column), sorted."""
    if not cols or len(cols[0][0]) == 0:
        return np.zeros((0, 2 * len(cols)), np.int64)
    m = []
    for v, ok in cols:
        ok = np.asarray(ok, bool)
        m.append(ok.astype(np.int64))
        m.append(np.where(ok, np.asarray(v).astype(np.int64), 0))
    a = np.stack(m, axis=1)
    return a[np.lexsort(a.T[::-1])]


def device_rows(batches, ncols):
    cols = [([], []) for _ in range(ncols)]
    for b in batches:
        if b.num_rows == 0:
            continue
        assert b.num_columns == ncols
        for i in range(ncols):
            v, ok = b.column(i)
            cols[i][0].append(np.asarray(v).astype(np.int64))
            cols[i][1].append(np.asarray(ok, bool))
    return matrix([(np.concatenate(v) if v else np.zeros(0, np.int64),
                    np.concatenate(ok) if ok else np.zeros(0, bool)) for v, ok in cols])


def expected_rows(jt, pairs, P, B):
    op, ob = pairs

    def take(sidecols, idx):
        out = []
        for v, ok in sidecols:
            safe = np.where(idx >= 0, idx, 0)
            out.append((v[safe], ok[safe] & (idx >= 0)))
        return out

    if jt in (A.JOIN_LEFT_SEMI, A.JOIN_LEFT_ANTI, A.JOIN_NULL_AWARE_ANTI):
        cols = take(P, op)
    elif jt == A.JOIN_EXISTENCE:
        cols = take(P, op) + [((ob != 0).astype(np.int64), np.ones(len(ob), bool))]
    elif jt in (A.JOIN_RIGHT_SEMI, A.JOIN_RIGHT_ANTI):
        cols = take(B, ob)
    else:
        cols = take(P, op) + take(B, ob)
    return matrix(cols)


def ncols_of(jt):
    return {A.JOIN_LEFT_SEMI: 2, A.JOIN_LEFT_ANTI: 2, A.JOIN_NULL_AWARE_ANTI: 2,
            A.JOIN_EXISTENCE: 3, A.JOIN_RIGHT_SEMI: 2, A.JOIN_RIGHT_ANTI: 2}.get(jt, 4)


def run_grace(jt, P, B, budget, build_batches=3, probe_batches=3, filt=False, max_rows=50_000):
    g = kgpu.GraceJoin(ctx(), jt, [0], [0], budget)
    if filt:
        g.set_filter(VLESS_2)
    nb, npr = len(B[0][0]), len(P[0][0])
    types = [A.T_INT64, A.T_INT64]
    for a, b in zip(*[np.linspace(0, nb, build_batches + 1).astype(int)[s:] for s in (0, 1)]):
        if b > a:
            g.build_sink([Col(t, v[a:b], ok[a:b]) for t, (v, ok) in zip(types, B)], b - a)
    g.build_finalize()
    out = []
    for a, b in zip(*[np.linspace(0, npr, probe_batches + 1).astype(int)[s:] for s in (0, 1)]):
        if b > a:
            out += g.probe([Col(t, v[a:b], ok[a:b]) for t, (v, ok) in zip(types, P)], b - a,
                           max_rows)
    out += g.finish(max_rows)
    assert all(o.num_rows <= max_rows for o in out)
    return out, g.stats()


# the residual filter over column 1 of each side: (p + b) % 7 != 0 AND p < b
VLESS_2 = [(op, 1 if op in (A.X_PROBE, A.X_BUILD) else a, imm) for op, a, imm in VLESS]


def oracle_pairs(jt, P, B, filt=False):
    kc = lambda s: [kgo.Col(kgo.T_INT64, v, ok) for v, ok in s]
    return kgo.join_ex(jt, kc(P)[:1], kc(B)[:1], [0], filt=(
        [(op, 1 if op in (A.X_PROBE, A.X_BUILD) else a, imm) for op, a, imm in VLESS]
        if filt else None), probe_cols=kc(P), build_cols=kc(B))


@pytest.mark.parametrize("jt", ALL_TYPES)
def test_every_join_type_spilled_equals_the_oracle(jt):
    say(f"grace join: type {jt}, 150K x 200K rows, 256 KiB budget")
    rng = np.random.default_rng(jt + 10)
    B = side(rng, 150_000, 120_000, 0.01)
    P = side(rng, 200_000, 150_000, 0.02)
    if jt == A.JOIN_NULL_AWARE_ANTI:
        B[0][1][:] = True  # NOT IN: a NULL build key would leave no row
    out, st = run_grace(jt, P, B, 256 << 10)
    assert st["spilled"] == 1 and st["partitions_joined"] >= 64
    got = device_rows(out, ncols_of(jt))
    want = expected_rows(jt, oracle_pairs(jt, P, B), P, B)
    assert got.shape == want.shape and np.array_equal(got, want)


def test_in_memory_under_the_budget():
    say("grace join: under the budget")
    rng = np.random.default_rng(3)
    B, P = side(rng, 50_000, 40_000, 0.01), side(rng, 80_000, 50_000, 0.01)
    for jt in (A.JOIN_INNER, A.JOIN_FULL_OUTER):
        out, st = run_grace(jt, P, B, 1 << 30)
        assert st["spilled"] == 0
        assert np.array_equal(device_rows(out, 4), expected_rows(jt, oracle_pairs(jt, P, B), P, B))


def test_recursive_split_and_filter():
    """A budget below a level-0 partition: every partition is split on the
    next 3 bits of g (depth 1), with the residual filter."""
    say("grace join: recursive split, inner and left outer with the filter")
    rng = np.random.default_rng(4)
    B, P = side(rng, 600_000, 500_000, 0.0), side(rng, 300_000, 500_000, 0.01)
    for jt in (A.JOIN_INNER, A.JOIN_LEFT_OUTER):
        out, st = run_grace(jt, P, B, 96 << 10, filt=True)
        assert st["repartitions"] >= 64 and st["max_depth"] >= 1
        want = expected_rows(jt, oracle_pairs(jt, P, B, filt=True), P, B)
        assert np.array_equal(device_rows(out, 4), want)


def test_role_reversal_for_semi_and_anti():
    say("grace join: role reversal, 20K probe rows against 800K build rows")
    rng = np.random.default_rng(5)
    B, P = side(rng, 800_000, 600_000, 0.01), side(rng, 20_000, 700_000, 0.05)
    for jt in (A.JOIN_LEFT_SEMI, A.JOIN_LEFT_ANTI):
        for filt in (False, True):
            out, st = run_grace(jt, P, B, 512 << 10, filt=filt)
            assert st["reversed"] > 0
            want = expected_rows(jt, oracle_pairs(jt, P, B, filt=filt), P, B)
            assert np.array_equal(device_rows(out, 2), want)


def test_not_in_with_a_null_build_key_spilled():
    say("grace join: NOT IN, a NULL build key, spilled")
    rng = np.random.default_rng(6)
    B, P = side(rng, 100_000, 90_000, 0.001), side(rng, 50_000, 100_000, 0.02)
    out, st = run_grace(A.JOIN_NULL_AWARE_ANTI, P, B, 128 << 10)
    assert st["spilled"] == 1 and sum(o.num_rows for o in out) == 0
    want = expected_rows(A.JOIN_NULL_AWARE_ANTI, oracle_pairs(A.JOIN_NULL_AWARE_ANTI, P, B), P, B)
    assert want.shape[0] == 0
