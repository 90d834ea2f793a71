"""Benchmark of the MI355X hash join / aggregate / Bloom / grace / served path.

With no --workload every configuration runs and one JSON line is printed:
the headline is configuration 2 (GROUP BY an int64 key drawn Zipf(s=1.1)
over 10M values, 200M rows, decimal(12,2) / float64 / int32 columns, COUNT(*)
and SUM/MIN/MAX/AVG of each: 13 aggregates), and "configs" holds one record
per other configuration, each with its roofline and its CPU baseline:

  config1_join_sum     1M-row build x 8M-row probe, SELECT key, SUM(amt),
                       COUNT(*) ... GROUP BY key (the fused join aggregate)
  config2_agg_int8     the 200M-row aggregation over a 4-value int8 key
  config3_bloom_join   15M-row build, 5 % kept, Bloom filter at fpp 0.03,
                       600M probe rows screened, joined and summed
  config4_grace        grace join + per-partition COUNT/SUM, 1.5B x 6B rows
                       Zipf(0.8) over 1.2B keys, 64 partitions, 8 GB table
                       budget (--scale shrinks it)
  config5_served       10,000 join+aggregate requests, batched/served form
                       against the plain one-launch-sequence-per-request form

Inputs are generated into HBM (synthetic, seeded) before each timed region.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload all|agg|...]

Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "krishiv-engine_amd"))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--rows", type=int, default=200_000_000)
    p.add_argument("--distinct", type=int, default=10_000_000)
    p.add_argument("--key", choices=["int64", "int8"], default="int64")
    p.add_argument("--cpu-rows", type=int, default=20_000_000)
    p.add_argument("--host-batch-rows", type=int, default=65536,
                   help="aggregation: also time the rows handed over from host memory in "
                        "batches of this many rows (0: skip)")
    p.add_argument("--bloom-probe-rows", type=int, default=600_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--workload", choices=["all", "agg", "joinagg", "bloom", "serve", "grace",
                                          "calib"],
                   default="all",
                   help="all: every configuration, one line (default); agg: the hash "
                        "aggregation; joinagg: the join of "
                        "SELECT key, SUM(amt), COUNT(*) ... GROUP BY key over 1M build rows; "
                        "bloom: the runtime Bloom filter feeding a join (15M build, 600M probe); "
                        "serve: the batched/served join-aggregate mix against the plain form; "
                        "grace: the grace join + per-partition aggregate at --scale of "
                        "configuration 4")
    p.add_argument("--scale", type=float, default=1.0,
                   help="grace: fraction of configuration 4 (1.5B build, 6B probe rows)")
    p.add_argument("--grace-hbm-budget", type=int, default=0,
                   help="grace: context HBM budget in bytes (partitions past it go to pinned host)")
    p.add_argument("--requests", type=int, default=10_000)
    p.add_argument("--serve-max-build", type=float, default=1e7,
                   help="serve: builds log-uniform in [1K, this] rows")
    p.add_argument("--qps", type=float, default=400.0)
    p.add_argument("--autotune", action="store_true",
                   help="joinagg: run the shape autotuner for this build size before timing; "
                        "serve: for the size classes of the large requests' builds")
    p.add_argument("--large-workers", type=int, default=0,
                   help="serve: large-request worker threads (KG_SERVE_LARGE_WORKERS; 0: the default)")
    p.add_argument("--latency-out", default="",
                   help="serve: write each request's probe rows and latencies (npz) here")
    p.add_argument("--probe-rows", type=int, default=8_000_000,
                   help="joinagg: probe rows (configuration 1: 8M)")
    p.add_argument("--build-rows", type=int, default=1_000_000,
                   help="joinagg: build rows (keys a permutation of 0..n-1; fk uniform over 1.2n); "
                        "4096 or fewer take the direct index")
    p.add_argument("--bloom-layout", choices=["sbbf", "spark"], default="sbbf",
                   help="bloom: the runtime filter's split-block layout (the reference's), or "
                        "Spark's BloomFilterImpl layout over xxhash64 items (murmur3 bit layout)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 1 GiB device-to-device copy (to scale FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def copy_calibration():
    """A plain 1 GiB device copy: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    its counters give the scale of those counters for known traffic."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 1 << 30}), flush=True)


def cpu_threads() -> int:
    """Host threads of the CPU baselines: the GPU box gives one GPU 16 cores."""
    return max(1, min(16, os.cpu_count() or 1))


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import kgo
    return kgo


def cpu_baseline(args, aggs, key="int64"):
    """The oracle (the C restatement of the reference's CPU aggregate) on a
    bounded sample of the same workload, run as the reference's pool runs
    it: thread-local tables in 16 radix partitions over contiguous shares of
    the rows, then a parallel merge per partition (aggregate/mod.rs:11-21,
    rowtable.rs:1805; kgo_pagg_*)."""
    kgo = oracle()
    n = args.cpu_rows
    threads = cpu_threads()
    cdf = kgo.zipf_cdf(args.distinct, 1.1)
    keyv, dec, f, i32, k8 = kgo.gen_agg(3, 0, n, cdf, args.distinct)
    kt, kv = (kgo.T_INT64, keyv) if key == "int64" else (kgo.T_INT8, k8)
    types = [(kgo.T_DECIMAL64, 12, 2), (kgo.T_FLOAT64, 0, 0), (kgo.T_INT32, 0, 0)]
    t0 = time.perf_counter()
    a = kgo.AggPar((kt, 0, 0), types, aggs, threads)
    a.update(kgo.Col(kt, kv), [kgo.Col(kgo.T_DECIMAL64, dec, precision=12, scale=2),
                               kgo.Col(kgo.T_FLOAT64, f), kgo.Col(kgo.T_INT32, i32)])
    a.finish(raw=True)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"first {n} rows of the same generator ({key} key); the oracle's aggregate "
                      f"on {threads} threads (thread-local tables in 16 partitions, parallel "
                      f"merge per partition); {dt:.2f} s"}


def cpu_join_pairs(kgo, probe_key, probe_valid, build_key, threads):
    """The oracle's hash join of probe chunks on `threads` threads (ctypes
    releases the GIL inside each C call): (probe rows, build rows)."""
    import concurrent.futures as cf
    n = len(probe_key)
    cuts = np.linspace(0, n, threads + 1).astype(np.int64)
    bcol = [kgo.Col(kgo.T_INT64, build_key)]

    def part(t):
        lo, hi = int(cuts[t]), int(cuts[t + 1])
        pv = None if probe_valid is None else probe_valid[lo:hi]
        pi, bi = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, probe_key[lo:hi], pv)], bcol,
                          cap=(hi - lo) + 16)
        return pi + lo, bi

    with cf.ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(part, range(threads)))
    return np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])


def cpu_joinagg():
    """The oracle on configuration 1 exactly (1M build rows, 8M probe rows):
    the join of probe shares on the pool, then SUM(amt), COUNT(*) GROUP BY
    key in the pool-shaped aggregate."""
    kgo = oracle()
    threads = cpu_threads()
    nb, n = 1_000_000, 8_000_000
    bk, _ = kgo.gen_join_build(1, 0, nb, nb)
    fk, ok, amt = kgo.gen_join_probe(2, 0, n, 1_200_000, 10)
    t0 = time.perf_counter()
    pi, bi = cpu_join_pairs(kgo, fk, ok, bk, threads)
    a = kgo.AggPar((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                   [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)], threads)
    a.update(kgo.Col(kgo.T_INT64, bk[bi]), [kgo.Col(kgo.T_INT64, amt[pi])])
    a.finish(raw=True)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"configuration 1's exact shape (1M build, 8M probe rows, same generators); "
                      f"the oracle's hash join on {threads} threads, then its pool-shaped "
                      f"aggregate; {dt:.2f} s"}


def cpu_bloom(nbytes):
    """The oracle on the Bloom workload's first 20M probe rows: the filter
    built from the same 15M build rows (val < 50000), H_kg and the SBBF check
    of each probe key, the survivors' join and SUM(amt), COUNT(*), on the
    pool (probe shares per thread)."""
    import concurrent.futures as cf
    kgo = oracle()
    threads = cpu_threads()
    nb, n = 15_000_000, 20_000_000
    bk, bv = kgo.gen_join_build(4, 0, nb, nb)
    fk, ok, amt = kgo.gen_join_probe(5, 0, n, nb, 0)
    t0 = time.perf_counter()
    kept = bk[bv < 50_000]
    _, words = kgo.sbbf_build(kgo.hash_columns([kgo.Col(kgo.T_INT64, kept)]), nbytes=nbytes)
    cuts = np.linspace(0, n, threads + 1).astype(np.int64)

    def part(t):
        lo, hi = int(cuts[t]), int(cuts[t + 1])
        hit = kgo.sbbf_check_batch(words, kgo.hash_columns([kgo.Col(kgo.T_INT64, fk[lo:hi])]))
        sel = np.nonzero(hit)[0] + lo
        pi, _ = kgo.join(kgo.JOIN_INNER, [kgo.Col(kgo.T_INT64, fk[sel])],
                         [kgo.Col(kgo.T_INT64, kept)])
        return int(amt[sel][pi].sum()), len(pi)

    with cf.ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(part, range(threads)))
    total = sum(p[0] for p in parts)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"the same 15M build rows and the first 20M probe rows; the oracle's H_kg, "
                      f"SBBF build and check, join and SUM/COUNT on {threads} threads; "
                      f"{dt:.2f} s", "check_sum": total}


def cpu_grace(scale):
    """The oracle on configuration 4 at 1/1000 of its size (1.5M build x 6M
    probe rows, Zipf(0.8) over 1.2M keys, payloads in [1, 1e4]): the
    per-partition COUNT and SUM factorized per key (kgo_grace_factorized;
    the reference's own pair walk over ~1e11 pairs at this size is out of
    reach of a bounded sample), one thread."""
    kgo = oracle()
    nb, npr, nd = 1_500_000, 6_000_000, 1_200_000
    cdf = kgo.zipf_cdf(nd, 0.8)
    bk = kgo.gen_zipf_keys(6, 6, 0, nb, cdf, nd)
    pk = kgo.gen_zipf_keys(7, 6, 0, npr, cdf, nd)
    _, _, bp = kgo.gen_join_probe(8, 0, nb, 1, 0)
    _, _, pp = kgo.gen_join_probe(9, 0, npr, 1, 0)
    t0 = time.perf_counter()
    c, _, _ = kgo.grace_factorized(kgo.Col(kgo.T_INT64, bk), kgo.Col(kgo.T_INT64, bp),
                                   kgo.Col(kgo.T_INT64, pk), kgo.Col(kgo.T_INT64, pp), 64, 0)
    dt = time.perf_counter() - t0
    return {"value": (nb + npr) / dt, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"configuration 4 at 1/1000 (1.5M x 6M rows, {int(c.sum())} pairs); the "
                      f"oracle's per-key factorized grace aggregate, one thread; {dt:.2f} s"}


def joinagg(args):
    """SELECT key, SUM(amt), COUNT(*) FROM build JOIN probe ON key = fk GROUP
    BY key: build 1M rows (key a seeded permutation of 0..1M-1), probe rows
    with fk uniform in [0, 1.2M) (1 % NULL) and amt in [1, 1e4], generated in
    HBM; a step builds the table, probes every row through the join fused
    with the aggregate and materializes the groups."""
    import kgpu
    from kgpu import Col, DeviceBuffer
    from kgpu import abi as A
    ctx = kgpu.Context(0)
    nb, n = args.build_rows, args.probe_rows
    L = kgpu.lib()
    bk = DeviceBuffer(ctx, nb * 8)
    bv = DeviceBuffer(ctx, nb * 8)
    A.check(L.kg_gen_join_build(ctx.h, 1, 0, nb, nb, bk.ptr, bv.ptr))
    fk = DeviceBuffer(ctx, n * 8)
    fkv = DeviceBuffer(ctx, (n + 7) // 8 + 8)
    amt = DeviceBuffer(ctx, n * 8)
    A.check(L.kg_gen_join_probe(ctx.h, 2, 0, n, nb * 12 // 10, 10, fk.ptr, fkv.ptr, amt.ptr))
    bcol = Col(A.T_INT64, bk, length=nb)
    pcols = [Col(A.T_INT64, fk, fkv, length=n), Col(A.T_INT64, amt, length=n)]
    groups = [0]
    tuned = None
    if args.autotune:
        t0 = time.perf_counter()
        shape, tms, ncand = kgpu.autotune_joinagg(ctx, nb, n)
        tuned = {"shape": shape, "est_ms": tms, "candidates": ncand,
                 "tuning_s": time.perf_counter() - t0}

    def step():
        j = kgpu.HashJoin(ctx, A.JOIN_INNER, [0], [0], expected_build_rows=nb)
        j.build_sink([bcol], nb)
        j.build_finalize()
        j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
        j.probe_aggregate(pcols, [1], n)
        r = j.aggregate_result()
        groups[0] = r.num_rows
        del r
        j.close()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kgpu.prof_reset()
    kgpu.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kgpu.prof_enable(False)
    # the roofline kernel: the staged scatter, or the direct-index kernel
    # for a tiny build (which reads the inputs once and writes nothing)
    rk_name, fam = ("k_jp_scatter_staged", "jp_scatter")
    if kgpu.prof_get("joinagg_direct")[1]:
        rk_name, fam = ("k_joinagg_direct", "joinagg_direct")
    sc_ms, sc_n = kgpu.prof_get(fam)
    host = (joinagg_host_batches(ctx, A, bk, fk, fkv, amt, nb, n, args.host_batch_rows, groups[0])
            if args.host_batch_rows > 0 else None)
    in_bytes = n * 16 + n // 8  # fk, amt, fk validity: each read once
    out_bytes = groups[0] * 24 + nb * 8  # (key, sum, count) rows; the build keys
    avg = sc_ms / max(sc_n, 1)
    per_launch = in_bytes / max(sc_n / args.steps, 1)
    achieved = per_launch / (avg / 1e3) if sc_ms > 0 else None
    step_ms = dt * 1e3 / args.steps
    return {
        "metric": f"join probe rows/s (fused GROUP BY build key; {nb} build rows)",
        "value": n * args.steps / dt, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": f"SELECT key, SUM(amt), COUNT(*) join-aggregate, {nb} build rows, "
                               f"{n} probe rows, fk uniform [0, {nb * 12 // 10}), 1% NULL",
                   "groups": groups[0], "shape": kgpu.tune_get(ctx, nb),
                   **({"autotune": tuned} if tuned else {})},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK if achieved else None,
                     "traffic": traffic_of(f"{rk_name}_joinagg_{n // 1_000_000}M_traffic",
                                           int(per_launch // 16.125) if sc_n else -1),
                     "kernel": rk_name, "kernel_avg_ms": avg,
                     "launches_per_step": sc_n / args.steps,
                     "algorithmic_bytes_per_step": in_bytes},
        "step_hbm_frac": ((in_bytes + out_bytes) / (step_ms / 1e3)) / HBM_PEAK,
        "phases": phases(["join_build", "jp_hist", "jp_scatter", "jp_merge", "joinagg_direct",
                          "agg_output"]),
        **({"host_batches": host} if host else {}),
        **({} if args.no_cpu or nb != 1_000_000 or n != 8_000_000
           else {"cpu_baseline": cpu_joinagg()}),
    }


def joinagg_host_batches(ctx, A, bk, fk, fkv, amt, nb, n, batch, groups):
    """The drop-in path of the join aggregate: the same build and probe rows
    handed over from pageable host memory in batches of `batch` rows (one
    build_sink or probe call each, as operators receive Arrow batches), then
    the result; PCIe and host copies included.  One untimed pass, then one
    timed pass."""
    import ctypes as C
    import kgpu
    L = kgpu.lib()
    hb = np.empty(nb, dtype=np.int64)
    hf = np.empty(n, dtype=np.int64)
    ha = np.empty(n, dtype=np.int64)
    hv = np.empty((n + 7) // 8, dtype=np.uint8)
    for h, d, nbytes in ((hb, bk, nb * 8), (hf, fk, n * 8), (ha, amt, n * 8), (hv, fkv, hv.nbytes)):
        A.check(L.kg_memcpy_d2h(ctx.h, h.ctypes.data, d.ptr, nbytes))
    bc = (A.Column * 1)()
    bc[0] = kgpu.Col(A.T_INT64, hb).c()
    pc = (A.Column * 2)()
    pc[0] = kgpu.Col(A.T_INT64, hf).c()
    pc[0].validity = hv.ctypes.data
    pc[1] = kgpu.Col(A.T_INT64, ha).c()
    ic = (C.c_int32 * 1)(1)
    times = []
    for _ in range(2):
        ctx.synchronize()
        t0 = time.perf_counter()
        j = kgpu.HashJoin(ctx, A.JOIN_INNER, [0], [0], expected_build_rows=nb)
        for off in range(0, nb, batch):
            bc[0].offset = off
            A.check(L.kg_join_build_sink(j.h, 1, bc, min(batch, nb - off)))
        j.build_finalize()
        j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
        for off in range(0, n, batch):
            pc[0].offset = pc[1].offset = off
            A.check(L.kg_joinagg_probe(j.h, 2, pc, min(batch, n - off), ic))
        r = j.aggregate_result()
        got = r.num_rows
        times.append(time.perf_counter() - t0)
        del r
        j.close()
        if got != groups:
            raise RuntimeError(f"host-batch join aggregate gave {got} groups, not {groups}")
    dt = times[-1]
    return {"batch_rows": batch, "rows_per_s": n / dt, "ms": dt * 1e3,
            "note": "build and probe rows from pageable host memory, one call a batch; "
                    "PCIe-inclusive, not `value`"}


def bloom(args):
    """Runtime Bloom filter feeding a join: build 15M rows (key unique, val
    uniform in [0, 1e6)); the predicate val < 50000 keeps 5 % (the same
    selectivity as attr < 5 over attr uniform in [0, 100)); a split-block
    Bloom filter of the kept keys sized for 750K items at fpp 0.03; 600M probe
    fk uniform over the 15M key range pre-screened by it; the survivors
    joined (selection-vector probe) and SUM(amt), COUNT(*) aggregated."""
    import kgpu
    from kgpu import Col, DeviceBuffer
    from kgpu import abi as A
    ctx = kgpu.Context(0)
    nb, n = 15_000_000, args.bloom_probe_rows
    L = kgpu.lib()
    bk = DeviceBuffer(ctx, nb * 8)
    bv = DeviceBuffer(ctx, nb * 8)
    A.check(L.kg_gen_join_build(ctx.h, 4, 0, nb, nb, bk.ptr, bv.ptr))
    fk = DeviceBuffer(ctx, n * 8)
    fkv = DeviceBuffer(ctx, (n + 7) // 8 + 8)
    amt = DeviceBuffer(ctx, n * 8)
    A.check(L.kg_gen_join_probe(ctx.h, 5, 0, n, nb, 0, fk.ptr, fkv.ptr, amt.ptr))
    nbytes = kgpu.RuntimeBloom.bytes_for(750_000, 0.03)
    sbits = kgpu.RuntimeBloom.spark_bits_for(750_000, 0.03)
    shash = kgpu.RuntimeBloom.spark_hashes_for(750_000, sbits)
    spark = args.bloom_layout == "spark"
    surv = DeviceBuffer(ctx, n * 4 + 64)
    stats = {}

    def step():
        rows, kept = kgpu.select_range(ctx, Col(A.T_INT64, bv, length=nb), 0, 49_999)
        keys = kgpu.gather(ctx, [Col(A.T_INT64, bk, length=nb)], rows, kept)
        b = (kgpu.RuntimeBloom(ctx, spark_bits=sbits, spark_hashes=shash) if spark
             else kgpu.RuntimeBloom(ctx, nbytes))
        b.insert(keys.as_col(0), nrows=kept)
        j = kgpu.HashJoin(ctx, A.JOIN_INNER, [0], [0], expected_build_rows=kept)
        j.build_sink([keys.as_col(0)], kept)
        j.build_finalize()
        pc = [Col(A.T_INT64, fk, length=n), Col(A.T_INT64, amt, length=n)]
        _, ns = b.probe(pc[0], out_rows=surv)
        out = j.probe_selected(pc, surv, ns, n)
        g = kgpu.HashAggregate(ctx, None, [(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
        g.sink(None, [out.as_col(1)], out.num_rows)
        g.finalize()
        r = g.result()
        stats.update(kept=kept, survivors=ns, joined=out.num_rows, sum=int(r.column(0)[0][0]))
        del r, out
        g.close()
        j.close()
        b.close()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kgpu.prof_reset()
    kgpu.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kgpu.prof_enable(False)
    fams = {f: kgpu.prof_get(f) for f in ("bloom_probe", "bloom_compact", "bloom_insert",
                                          "join_probe", "join_gather", "join_build",
                                          "agg_update", "select_range", "gather")}
    bp_ms, bp_n = fams["bloom_probe"]
    bc_ms, _ = fams["bloom_compact"]
    bp_ms += bc_ms
    avg = bp_ms / max(bp_n, 1)
    per_launch = n * 8 / max(bp_n / args.steps, 1)  # the probe key column
    achieved = per_launch / (avg / 1e3) if bp_ms > 0 else None
    return {
        "metric": f"Bloom-screened probe rows/s (600M probe, 15M build, fpp 0.03, "
                  f"{'Spark murmur3' if spark else 'split-block'} layout)",
        "value": n * args.steps / dt, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": f"Bloom filter feeding a join, {nb} build rows, {n} probe rows",
                   **stats},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK if achieved else None,
                     "traffic": traffic_of("k_bloom_probe_bits_bloom_600M_traffic",
                                           n // max(int(bp_n / args.steps), 1)),
                     "kernel": "bloom probe (k_bloom_probe_bits + expand)", "kernel_avg_ms": avg},
        "random_roofline": random_ceiling(
            ctx, (sbits + 7) // 8 if spark else nbytes, 8 if spark else 32,
            n / max(bp_n / args.steps, 1) / (avg / 1e3) if bp_ms > 0 else 0.0,
            "bloom probe (k_bloom_probe_bits + expand)",
            "probe rows per second against random 8-byte reads (Spark layout: up to k words "
            "a row, the first zero bit ends it)" if spark else
            "probe rows per second against random 32-byte reads (one filter block a row)"),
        "families_ms_per_step": {f: v[0] / args.steps for f, v in fams.items()},
        **({} if args.no_cpu or spark else {"cpu_baseline": cpu_bloom(nbytes)}),
    }


def grace(args):
    """Configuration 4 at --scale: build rows (int64 key Zipf(0.8) over 0.8x
    the build rows' distinct keys, int64 payload in [1, 1e4]) and 4x as many
    probe rows from the same Zipf with payloads in [1, 1e4]; SELECT p,
    COUNT(*), SUM(probe.pay * build.pay % 2^31) GROUP BY p = partition_of(
    H_kg(key), 64), through the grace join (64 partitions in HBM, pinned host
    past the context budget) fused with the per-partition aggregate.  A step
    sinks every batch and finishes."""
    import kgpu
    from kgpu import Col, DeviceBuffer
    from kgpu import abi as A
    ctx = kgpu.Context(0, args.grace_hbm_budget) if args.grace_hbm_budget else kgpu.Context(0)
    L = kgpu.lib()
    nb = int(1_500_000_000 * args.scale)
    npr = 4 * nb
    nd = int(1_200_000_000 * args.scale)
    t_gen = time.perf_counter()
    # the Zipf table of 1.2B ranks is built on the device (the host loop of
    # kg_zipf_cdf over 1.2B terms would take tens of seconds)
    cdf_d = DeviceBuffer(ctx, nd * 8)
    A.check(L.kg_zipf_cdf_device(ctx.h, nd, 0.8, cdf_d.ptr))
    bk, bp = DeviceBuffer(ctx, nb * 8), DeviceBuffer(ctx, nb * 8)
    pk, pp = DeviceBuffer(ctx, npr * 8), DeviceBuffer(ctx, npr * 8)
    # one key distribution on both sides: the permutation of ranks to keys is
    # shared (seed 6), the draws differ
    A.check(L.kg_gen_zipf_keys(ctx.h, 6, 6, 0, nb, cdf_d.ptr, nd, bk.ptr))
    A.check(L.kg_gen_zipf_keys(ctx.h, 7, 6, 0, npr, cdf_d.ptr, nd, pk.ptr))
    cdf_d.free()
    # payloads: the amt column of the probe generator, uniform in [1, 1e4]
    A.check(L.kg_gen_join_probe(ctx.h, 8, 0, nb, 1, 0, None, None, bp.ptr))
    A.check(L.kg_gen_join_probe(ctx.h, 9, 0, npr, 1, 0, None, None, pp.ptr))
    ctx.synchronize()
    t_gen = time.perf_counter() - t_gen
    batch = 1 << 26
    res = {}

    def step():
        g = kgpu.GraceJoinAggregate(ctx, 64, table_budget=int((8 << 30) * args.scale))
        for a in range(0, nb, batch):
            m = min(batch, nb - a)
            g.build_sink(Col(A.T_INT64, bk, length=m, offset=a),
                         Col(A.T_INT64, bp, length=m, offset=a), m)
        for a in range(0, npr, batch):
            m = min(batch, npr - a)
            g.probe_sink(Col(A.T_INT64, pk, length=m, offset=a),
                         Col(A.T_INT64, pp, length=m, offset=a), m)
        c, s_, v, spilled = g.finish()
        res.update(pairs=int(c.sum()), spilled=spilled)
        g.close()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kgpu.prof_reset()
    kgpu.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kgpu.prof_enable(False)
    rows = nb + npr
    step_ms = dt * 1e3 / args.steps
    in_bytes = rows * 16
    ph = phases(["hash_partition", "scatter", "grace_build", "grace_probe", "agg_update",
                 "agg_partition", "agg_merge", "agg_output"])
    # a partition's table: 2x its build rows rounded up to a power of two, 32-byte slots
    tab = (1 << ((2 * nb // 64) - 1).bit_length()) * 32
    gp = ph.get("grace_probe")
    rr = (random_ceiling(ctx, tab, 32, (npr / 64) / (gp["avg_ms"] / 1e3), "k_grace_probe",
                         "probe rows per second of one partition (1/64 of the probe rows) "
                         "against random 32-byte reads from a table of a mean partition's size")
          if gp else None)
    sc_ms, sc_n = kgpu.prof_get("scatter")
    sc_avg = sc_ms / max(sc_n, 1)
    sc_rows = rows * args.steps / max(sc_n, 1)  # rows of one partition-pass launch
    # a partition pass reads each row's key and payload (16 B) and writes them
    # once into the row's partition (16 B)
    sc_ach = sc_rows * 32 / (sc_avg / 1e3) if sc_ms > 0 else None
    return {
        "metric": "grace join + per-partition aggregate rows/s (build + probe rows)",
        "value": rows * args.steps / dt, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": f"configuration 4 at scale {args.scale}: {nb} build rows, {npr} probe "
                               f"rows, Zipf(0.8) over {nd} keys, 64 partitions, table budget "
                               f"{int((8 << 30) * args.scale)} B",
                   "pairs": res.get("pairs"), "spilled_bytes": res.get("spilled")},
        "roofline": {"bound": "hbm", "achieved": sc_ach / 1e9 if sc_ach else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": sc_ach / HBM_PEAK if sc_ach else None,
                     "traffic": traffic_of("k_gp_scatter_grace_traffic", int(sc_rows)),
                     "kernel": "k_gp_scatter (the partition pass: key and payload read once, "
                               "written once)",
                     "kernel_avg_ms": sc_avg, "launches_per_step": sc_n / args.steps,
                     "algorithmic_bytes_per_step": in_bytes},
        "step_hbm_frac": (in_bytes / (step_ms / 1e3)) / HBM_PEAK,
        "random_roofline": rr,
        "phases": ph,
        "generation_s": t_gen,
        **({} if args.no_cpu else {"cpu_baseline": cpu_grace(args.scale)}),
    }


def serve(args):
    """The batched/served mix: requests arrive as Poisson at --qps (seed 8);
    each is SELECT key, SUM(pay), COUNT(*) FROM build JOIN probe ON key = fk
    GROUP BY key with a build of log-uniform [1K, 10M] unique int64 keys and
    a probe 10x as long (fk uniform over 1.1x the key range, int64 payload);
    30 % probe one of 8 resident builds.  Build sizes are snapped to 64
    geometric size classes whose inputs (and the resident tables' probe
    inputs) are generated in HBM once before either form runs (timed apart),
    so all requests arrive as one continuous Poisson stream.  The plain form
    runs one join-aggregate launch sequence per request in arrival order on
    one stream; the batched form submits each request to the server as it
    arrives.  Latency = completion - arrival."""
    import kgpu
    from kgpu import Col, DeviceBuffer
    from kgpu import abi as A
    ctx = kgpu.Context(0)
    L = kgpu.lib()
    rng = np.random.default_rng(8)
    R = args.requests
    top = args.serve_max_build
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(top), R)).astype(np.int64)
    resident = rng.random(R) < 0.3
    table_of = rng.integers(0, 8, R)
    res_sizes = np.exp(rng.uniform(np.log(1e3), np.log(top), 8)).astype(np.int64)
    gaps = rng.exponential(1.0 / args.qps, R)
    # resident builds (keys 0..n-1 permuted), built once for both forms
    res_keys = []
    for t in range(8):
        k = DeviceBuffer(ctx, int(res_sizes[t]) * 8)
        A.check(L.kg_gen_join_build(ctx.h, 1000 + t, 0, int(res_sizes[t]), int(res_sizes[t]), k.ptr, None))
        res_keys.append(k)
    plain_tables = []
    for t in range(8):
        j = kgpu.HashJoin(ctx, A.JOIN_INNER, [0], [0], expected_build_rows=int(res_sizes[t]))
        j.build_sink([Col(A.T_INT64, res_keys[t], length=int(res_sizes[t]))], int(res_sizes[t]))
        j.build_finalize()
        plain_tables.append(j)
    if args.large_workers:
        os.environ["KG_SERVE_LARGE_WORKERS"] = str(args.large_workers)
    tuned = {}
    if args.autotune:
        # the size classes of the large requests' builds (4M probe rows and
        # more: builds of 400K rows and more), tuned before either form runs
        t0 = time.perf_counter()
        for lg in range(19, 25):
            shape, _, _ = kgpu.autotune_joinagg(ctx, 1 << lg, 10 << lg)
            tuned[f"2^{lg}"] = shape
        tuned["tuning_s"] = time.perf_counter() - t0
    srv = kgpu.JoinAggServer(ctx, arena_bytes=64 << 30, max_batch=256, use_graph=True)
    srv_tables = [srv.add_table(res_keys[t], int(res_sizes[t])) for t in range(8)]

    # each request's build size snapped to one of 64 geometric size classes
    # over [1K, top]; the inputs of every class (and of every resident
    # table's probes) are generated once, before either form runs, so the
    # whole stream of requests arrives as one Poisson process with nothing
    # else on the GPU
    grid = np.unique(np.exp(np.linspace(np.log(1e3), np.log(top), 64)).astype(np.int64))
    sizes = grid[np.abs(np.log(sizes)[:, None] - np.log(grid)[None, :]).argmin(axis=1)]

    def nb_of(r):
        return int(res_sizes[table_of[r]]) if resident[r] else int(sizes[r])

    t0 = time.perf_counter()
    pool = {}
    for nb in sorted({nb_of(r) for r in range(R)}):
        npr = 10 * nb
        bk = DeviceBuffer(ctx, nb * 8)
        A.check(L.kg_gen_join_build(ctx.h, 2 * nb + 1, 0, nb, nb, bk.ptr, None))
        fk = DeviceBuffer(ctx, npr * 8)
        pay = DeviceBuffer(ctx, npr * 8)
        A.check(L.kg_gen_join_probe(ctx.h, 2 * nb + 2, 0, npr, int(nb * 1.1), 0, fk.ptr, None,
                                    pay.ptr))
        pool[nb] = (bk, fk, pay, nb, npr)
    ctx.synchronize()
    gen_s = time.perf_counter() - t0
    arrive_at = np.cumsum(gaps) - gaps[0]

    def run_form(batched, saturated=False):
        lat = np.zeros(R)
        arrive = np.zeros(R) if saturated else arrive_at
        # a quiet 0.3 s before each pass (far above any inter-arrival gap)
        # marks the passes' boundaries in a kernel trace
        ctx.synchronize()
        time.sleep(0.3)
        start = time.perf_counter()
        if not batched:
            for r in range(R):
                wait = arrive[r] - (time.perf_counter() - start)
                if wait > 0:
                    time.sleep(wait)
                bk, fk, pay, nb, npr = pool[nb_of(r)]
                if resident[r]:
                    j = plain_tables[table_of[r]]
                else:
                    j = kgpu.HashJoin(ctx, A.JOIN_INNER, [0], [0], expected_build_rows=nb)
                    j.build_sink([Col(A.T_INT64, bk, length=nb)], nb)
                    j.build_finalize()
                j.aggregate([(A.T_INT64, 0, 0)], [(A.AGG_SUM, 0), (A.AGG_COUNT_STAR, 0)])
                j.probe_aggregate([Col(A.T_INT64, fk, length=npr), Col(A.T_INT64, pay, length=npr)],
                                  [1], npr)
                res = j.aggregate_result()
                lat[r] = time.perf_counter() - start - arrive[r]
                del res
                if not resident[r]:
                    j.close()
        else:
            nxt, done = 0, 0
            while done < R:
                now = time.perf_counter() - start
                while nxt < R and arrive[nxt] <= now:
                    bk, fk, pay, nb, npr = pool[nb_of(nxt)]
                    if resident[nxt]:
                        srv.submit(nxt, fk, pay, npr, table=srv_tables[table_of[nxt]])
                    else:
                        srv.submit(nxt, fk, pay, npr, build_keys=bk, build_rows=nb)
                    nxt += 1
                ids = srv.progress(wait=nxt >= R)
                t = time.perf_counter() - start
                for r in ids:
                    lat[r] = t - arrive[r]
                    srv.result(r)
                    done += 1
        return lat, time.perf_counter() - start, gen_s

    out = {}
    dump = {"probe_rows": np.array([10 * nb_of(r) for r in range(R)]), "resident": resident}
    for name, batched in (("plain", False), ("batched", True)):
        m0 = ctx.mem_stats()
        lat_s, busy_s, _ = run_form(batched, saturated=True)
        lat, busy, gen = run_form(batched)
        m1 = ctx.mem_stats()
        dump[name + "_saturated"], dump[name + "_at_qps"] = lat_s, lat
        out[name] = {"queries_per_s": R / busy_s,
                     "saturated_p50_ms": float(np.percentile(lat_s, 50) * 1e3),
                     "at_qps": {"qps": args.qps, "achieved_qps": R / busy,
                                "p50_ms": float(np.percentile(lat, 50) * 1e3),
                                "p99_ms": float(np.percentile(lat, 99) * 1e3)},
                     "wall_s_excl_gen": busy_s, "gen_s": gen,
                     "allocator": {k: m1[k] - m0[k] for k in ("hip_mallocs", "hip_frees",
                                                              "cache_hits", "releases")}}
    if args.latency_out:
        np.savez(args.latency_out, **dump)
    # algorithmic bytes of the mix: each request's build keys (8 B a row; the
    # resident ones once), probe keys and payloads (16 B a row), result rows
    # (24 B a group, at most the build rows)
    build_rows = sum(nb_of(r) for r in range(R) if not resident[r]) + int(res_sizes.sum())
    probe_rows = sum(10 * nb_of(r) for r in range(R))
    mix_bytes = build_rows * 8 + probe_rows * 16 + sum(nb_of(r) for r in range(R)) * 24
    bb = mix_bytes / out["batched"]["wall_s_excl_gen"]
    return {
        "metric": "batched vs plain join-aggregate queries/s (saturated) and p50/p99 latency at a fixed QPS",
        "value": out["batched"]["queries_per_s"], "unit": "queries/s", "n_gpus": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded counter-based generator; one input set per build size class, "
                "generated in HBM before the stream)",
        "config": {"workload": f"{R} join+aggregate requests, Poisson {args.qps} q/s (seed 8), build "
                               f"log-uniform [1K, {int(top):,}], probe 10x, 30 % on 8 resident builds, "
                               "arena 64 GB", "size_classes": int(len(grid)),
                   "generation_s": gen_s,
                   "large_workers": int(os.environ.get("KG_SERVE_LARGE_WORKERS", "1")),
                   **({"autotuned_classes": tuned} if tuned else {})},
        "plain": out["plain"], "batched": out["batched"],
        "batched_over_plain_qps": out["batched"]["queries_per_s"] / out["plain"]["queries_per_s"],
        "roofline": {"bound": "hbm", "achieved": bb / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": bb / HBM_PEAK, "traffic": None,
                     "kernel": "the whole batched form, saturated (every kernel of the mix)",
                     "algorithmic_bytes": mix_bytes},
        **({} if args.no_cpu else {"cpu_baseline": cpu_serve(R)}),
    }


def cpu_serve(R, budget_s=10.0):
    """The oracle running the served mix serially, in arrival order (the same
    seeded sizes and generators): each request's join of probe shares on the
    pool, then its pool-shaped SUM/COUNT GROUP BY key, until about budget_s
    seconds of work; queries/s and the p50 of the per-request times."""
    kgo = oracle()
    threads = cpu_threads()
    rng = np.random.default_rng(8)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e7), R)).astype(np.int64)
    resident = rng.random(R) < 0.3
    table_of = rng.integers(0, 8, R)
    res_sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e7), 8)).astype(np.int64)
    res_keys = {}
    times = []
    spent = 0.0
    for r in range(R):
        if resident[r]:
            t = int(table_of[r])
            nb = int(res_sizes[t])
            if t not in res_keys:
                res_keys[t] = kgo.gen_join_build(1000 + t, 0, nb, nb, with_val=False)[0]
            bk = res_keys[t]
        else:
            nb = int(sizes[r])
            bk = kgo.gen_join_build(2 * r + 1, 0, nb, nb, with_val=False)[0]
        fk, _, pay = kgo.gen_join_probe(2 * r + 2, 0, 10 * nb, int(nb * 1.1), 0)
        t0 = time.perf_counter()
        pi, bi = cpu_join_pairs(kgo, fk, None, bk, threads)
        a = kgo.AggPar((kgo.T_INT64, 0, 0), [(kgo.T_INT64, 0, 0)],
                       [(kgo.AGG_SUM, 0), (kgo.AGG_COUNT_STAR, 0)], threads)
        a.update(kgo.Col(kgo.T_INT64, bk[bi]), [kgo.Col(kgo.T_INT64, pay[pi])])
        a.finish(raw=True)
        dt = time.perf_counter() - t0
        times.append(dt)
        spent += dt
        if spent >= budget_s:
            break
    return {"value": len(times) / spent, "unit": "queries/s", "cores": threads, "kind": "port",
            "p50_ms": float(np.percentile(times, 50) * 1e3),
            "sample": f"the first {len(times)} requests of the same mix in arrival order, one at a "
                      f"time; the oracle's join on {threads} threads and its pool-shaped "
                      f"aggregate; {spent:.2f} s"}


def random_ceiling(ctx, table_bytes: int, read_bytes: int, achieved_per_s: float, kernel: str,
                   note: str):
    """A probe phase's random reads per second against the measured ceiling
    of random reads of read_bytes from a device table of (the next power of
    two of) table_bytes, kg_calib_random_read on this GPU."""
    import kgpu
    tb = 1 << max(12, (int(table_bytes) - 1).bit_length())
    peak = kgpu.calib_random_read(ctx, tb, read_bytes, 1 << 27)
    return {"bound": "random reads", "kernel": kernel, "table_bytes": int(table_bytes),
            "calib_table_bytes": tb, "read_bytes": read_bytes, "achieved": achieved_per_s / 1e9,
            "peak": peak / 1e9, "unit": "G reads/s", "frac": achieved_per_s / peak, "note": note}


def calib(args):
    """Measured ceilings of the random-access phases on this GPU: random
    reads of 8, 16 and 32 bytes at hashed entries of device tables of 1 MB to
    4 GB (the probe phases' roofline at their table size), and no-return
    64-bit global atomic adds over 1 to 2^28 words (global-table updates)."""
    import kgpu
    ctx = kgpu.Context(0)
    reads = []
    for mb in (1, 4, 16, 64, 256, 1024, 4096):
        for b in (8, 16, 32):
            r = kgpu.calib_random_read(ctx, mb << 20, b, 1 << 28)
            reads.append({"table_mb": mb, "read_bytes": b, "g_reads_per_s": r / 1e9,
                          "gb_per_s": r * b / 1e9})
    atomics = []
    for a in (1, 64, 4096, 1 << 16, 1 << 20, 1 << 24, 1 << 28):
        r = kgpu.calib_atomic_add(ctx, a, (1 << 22) if a < 4096 else (1 << 27))
        atomics.append({"addresses": a, "g_ops_per_s": r / 1e9})
    ref = [x for x in reads if x["table_mb"] == 256 and x["read_bytes"] == 32][0]
    print(json.dumps({
        "metric": "random-read and global-atomic ceilings (calibration)",
        "value": ref["g_reads_per_s"] * 1e9, "unit": "random 32-byte reads/s from a 256 MB table",
        "n_gpus": 1, "higher_is_better": True, "dtype": "u64", "data": "synthetic (hashed indexes)",
        "config": {"workload": "kg_calib_random_read over 2^28 reads a launch, "
                               "kg_calib_atomic_add over 2^27 adds (2^22 below 4096 addresses)"},
        "random_reads": reads, "atomic_adds": atomics}), flush=True)


def traffic_of(summary: str, launch_rows: int):
    """HBM bytes per launch from a committed counter summary
    (profiles/<summary>.md: `key: value` lines rows_per_launch, fetch_kib and
    write_kib per dispatch from two rocprofv3 --pmc runs of the same workload,
    fetch_scale and write_scale from what those counters read for a 1 GiB
    device copy), or None when absent or taken at another size."""
    path = os.path.join(ROOT, "profiles", summary + ".md")
    if not os.path.exists(path):
        return None
    t = {}
    with open(path) as f:
        for line in f:
            parts = line.strip().split(": ")
            if len(parts) == 2 and parts[0].isidentifier():
                try:
                    t[parts[0]] = float(parts[1])
                except ValueError:
                    pass
    if int(t.get("rows_per_launch", -1)) != launch_rows:
        return None
    return (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024


def phases(names):
    """Average ms per launch of the library's timed regions (HIP events on
    the stream the kernels run on)."""
    import kgpu
    out = {}
    for name in names:
        ms, cnt = kgpu.prof_get(name)
        if cnt:
            out[name] = {"avg_ms": ms / cnt, "launches": cnt}
    return out


def agg(args, key="int64"):
    """Configuration 2: the hash aggregation of 200M rows over the int64 Zipf
    key, or (key int8) the 4-value int8 key."""
    import kgpu
    from kgpu import Col, DeviceBuffer
    from kgpu import abi as A

    ctx = kgpu.Context(0)
    n, nd = args.rows, args.distinct
    # inputs resident in HBM
    cdf = np.zeros(nd)
    A.check(kgpu.lib().kg_zipf_cdf(nd, 1.1, cdf.ctypes.data))
    cdf_d = DeviceBuffer.from_numpy(ctx, cdf)
    key_d = DeviceBuffer(ctx, n * 8)
    k8_d = DeviceBuffer(ctx, n)
    dec_d = DeviceBuffer(ctx, n * 8)
    f_d = DeviceBuffer(ctx, n * 8)
    i32_d = DeviceBuffer(ctx, n * 4)
    A.check(kgpu.lib().kg_gen_agg(ctx.h, 3, 0, n, cdf_d.ptr, nd,
                                  key_d.ptr if key == "int64" else None, dec_d.ptr, f_d.ptr,
                                  i32_d.ptr, k8_d.ptr if key == "int8" else None))
    if key == "int64":
        kcol = Col(A.T_INT64, key_d, length=n)
        key_bytes = 8
    else:
        kcol = Col(A.T_INT8, k8_d, length=n)
        key_bytes = 1
    cols = [Col(A.T_DECIMAL64, dec_d, precision=12, scale=2, length=n),
            Col(A.T_FLOAT64, f_d, length=n), Col(A.T_INT32, i32_d, length=n)]
    types = [(A.T_DECIMAL64, 12, 2), (A.T_FLOAT64, 0, 0), (A.T_INT32, 0, 0)]
    aggs = [(A.AGG_COUNT_STAR, 0)]
    for c in range(3):
        aggs += [(A.AGG_SUM, c), (A.AGG_MIN, c), (A.AGG_MAX, c), (A.AGG_AVG, c)]
    # no group-count estimate: the table is sized at the drain from the
    # sketch of the spilled keys (the one-byte key's table is fixed)
    expected = 0

    groups = [0]

    def step():
        g = kgpu.HashAggregate(ctx, (kcol.type, 0, 0), types, aggs, expected_groups=expected)
        g.sink(kcol, cols, n)
        g.finalize()
        r = g.result()
        groups[0] = r.num_rows
        del r
        g.close()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kgpu.prof_reset()
    kgpu.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kgpu.prof_enable(False)
    upd_ms, upd_n = kgpu.prof_get("agg_update")
    ph = phases(["agg_update", "agg_partition", "agg_merge", "agg_output"])
    ms_per_step = dt * 1e3 / args.steps
    rows_per_s = n * args.steps / dt
    in_bytes = n * (key_bytes + 8 + 8 + 4)  # every input column read once
    # the update kernel's algorithmic bytes per launch: its rows x 21-28 B;
    # summed over a step's launches that is in_bytes
    upd_avg_ms = upd_ms / max(upd_n, 1)
    launches_per_step = upd_n / args.steps
    bytes_per_launch = in_bytes / max(launches_per_step, 1)
    achieved = bytes_per_launch / (upd_avg_ms / 1e3) if upd_ms > 0 else None
    kname = "k_agg_pass1" if key == "int64" else "k_agg_small"
    out = {
        "metric": "hash aggregation rows/s (200M rows, Zipf 1.1 over 10M int64 keys, 13 aggregates)"
                  if key == "int64" else
                  "hash aggregation rows/s (200M rows, 4-value int8 key, 13 aggregates)",
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/decimal(12,2)/f64/int32",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": f"hash aggregation, {n} rows, key {key} "
                               f"{'Zipf(1.1) over %d' % nd if key == 'int64' else '4 values'}, "
                               "COUNT(*) + SUM/MIN/MAX/AVG of decimal(12,2), float64, int32",
                   "groups": groups[0]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9 if achieved else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK if achieved else None,
                     "traffic": traffic_of(f"{kname}_hash_agg_200M{'' if key == 'int64' else '_int8'}"
                                           "_traffic", int(bytes_per_launch // (key_bytes + 20))),
                     "kernel": kname,
                     "kernel_avg_ms": upd_avg_ms,
                     "launches_per_step": launches_per_step,
                     "algorithmic_bytes_per_step": in_bytes},
        # whole operator: inputs read once + result rows written once (key,
        # COUNT(*), and SUM 16 B, MIN, MAX, AVG of decimal(12,2), float64, int32:
        # 112 B a group)
        "step_hbm_frac": ((in_bytes + groups[0] * (112 if key == "int64" else 105))
                          / (ms_per_step / 1e3)) / HBM_PEAK,
        "phases": ph,
    }
    if args.host_batch_rows > 0:
        out["host_batches"] = agg_host_batches(ctx, A, kcol, cols, types, aggs, n, key_bytes,
                                               args.host_batch_rows, groups[0])
    del key_d, k8_d, dec_d, f_d, i32_d, cdf_d
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, [(1, 0)] + [(f, c) for c in range(3)
                                                             for f in (3, 4, 5, 6)], key)
    return out


def agg_host_batches(ctx, A, kcol, cols, types, aggs, n, key_bytes, batch, groups):
    """The drop-in path: the same rows handed over from pageable host memory
    in batches of `batch` rows (one kg_agg_sink call each, as an operator
    receives Arrow batches), then finalize and result; PCIe and host copies
    included.  One untimed pass, then one timed pass."""
    import ctypes as C
    import kgpu
    L = kgpu.lib()
    arrs = []
    for dc, w, dt in zip([kcol] + cols, [key_bytes, 8, 8, 4],
                         [np.int64 if key_bytes == 8 else np.int8, np.int64, np.float64, np.int32]):
        a = np.empty(n, dtype=dt)
        A.check(L.kg_memcpy_d2h(ctx.h, a.ctypes.data, dc.vptr, n * w))
        arrs.append(a)
    kc = kgpu.Col(kcol.type, arrs[0]).c()
    ca = (A.Column * 3)()
    for i in range(3):
        ca[i] = kgpu.Col(cols[i].type, arrs[i + 1], precision=cols[i].precision,
                         scale=cols[i].scale).c()
    times = []
    for _ in range(2):
        g = kgpu.HashAggregate(ctx, (kcol.type, 0, 0), types, aggs, expected_groups=0)
        ctx.synchronize()
        t0 = time.perf_counter()
        for off in range(0, n, batch):
            m = min(batch, n - off)
            kc.offset = off
            for i in range(3):
                ca[i].offset = off
            A.check(L.kg_agg_sink(g.h, C.byref(kc), ca, m))
        g.finalize()
        r = g.result()
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
        got = r.num_rows
        del r
        g.close()
        if got != groups:
            raise RuntimeError(f"host-batch aggregation gave {got} groups, not {groups}")
    dt = times[-1]
    return {"batch_rows": batch, "batches": -(-n // batch), "rows_per_s": n / dt,
            "ms": dt * 1e3, "host_bytes": n * (key_bytes + 20),
            "note": "pageable host columns, one sink call a batch; PCIe-inclusive, not `value`"}


def sub_record(rec):
    """A configuration's record inside the headline line (the keys of a bench
    line without the ones the headline already carries)."""
    drop = {"n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"}
    return {k: v for k, v in rec.items() if k not in drop}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the multi-GPU path is design only)"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    one = {"joinagg": lambda: joinagg(args), "bloom": lambda: bloom(args),
           "serve": lambda: serve(args), "grace": lambda: grace(args),
           "agg": lambda: agg(args, args.key)}
    if args.workload == "calib":
        calib(args)
        return
    if args.workload != "all":
        print(json.dumps(one[args.workload]()), flush=True)
        return
    import gc
    head = agg(args, "int64")
    subs = {}
    runs = [("config1_join_sum", lambda: joinagg(args)),
            ("config2_agg_int8", lambda: agg(args, "int8")),
            ("config3_bloom_join", lambda: bloom(args)),
            ("config5_served", lambda: serve(args)),
            # the same mix with builds of at most 100K rows: requests whose
            # cost is the per-query launch and host work the batched form saves
            ("config5_served_small_builds", lambda: serve(argparse.Namespace(
                **{**vars(args), "serve_max_build": 1e5, "no_cpu": True}))),
            ("config4_grace", lambda: grace(args))]
    for name, fn in runs:
        gc.collect()
        t0 = time.perf_counter()
        try:
            subs[name] = sub_record(fn())
        except Exception as e:  # a refused configuration is reported, not fatal
            subs[name] = {"error": f"{type(e).__name__}: {e}"}
        subs[name]["wall_s"] = time.perf_counter() - t0
        print(f"[bench] {name} done in {subs[name]['wall_s']:.1f} s", file=sys.stderr,
              flush=True)
    head["configs"] = subs
    print(json.dumps(head), flush=True)


if __name__ == "__main__":
    main()
