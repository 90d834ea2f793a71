// kg_tune.hip -- the shape autotuner of the hash join and its fused
// aggregate.  Per size class of the build side (ceil log2 of its rows) a
// context keeps one shape: the table's load factor (capacity >= rows * 100 /
// load_pct, a power of two), the staged partition scatter's rows per lane
// (its tile: 512 lanes x 4, 8 or 12 rows) and the partition merge's
// workgroup (1,024 lanes holding the whole LDS, or two of 512 with half the
// window each).  kg_autotune_joinagg times every candidate on synthetic
// inputs of the requested shape (unique build keys, probe keys uniform over
// 1.2x their range, 1 % NULL, SELECT key, SUM(pay), COUNT(*)) and stores the
// fastest; kg_tune_get/kg_tune_set read and force a class's shape.  Results
// do not depend on the shape, only the time does.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>

#include "kg_common.hpp"

namespace kg {

static void tck(int rc) {
  if (rc != KG_OK) throw Error{rc, std::string(kg_last_error())};
}

static kg_column tune_col(const void *p, int64_t n, const uint8_t *valid = nullptr) {
  kg_column c{};
  c.type = KG_T_INT64;
  c.flags = KG_COL_DEVICE;
  c.length = n;
  c.values = p;
  c.validity = valid;
  return c;
}

static void check_cfg(const kg_tune_config &c) {
  if (c.load_factor_pct < 10 || c.load_factor_pct > 90)
    fail(KG_ERR_UNSUPPORTED, "load_factor_pct must be 10 to 90");
  if (c.scatter_rows != 0 && c.scatter_rows != 4 && c.scatter_rows != 8 && c.scatter_rows != 12)
    fail(KG_ERR_UNSUPPORTED, "scatter_rows must be 0, 4, 8 or 12");
  if (c.merge_lanes != 512 && c.merge_lanes != 1024)
    fail(KG_ERR_UNSUPPORTED, "merge_lanes must be 512 or 1024");
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace kg

using namespace kg;

extern "C" {

int kg_tune_get(kg_ctx *ctx, int64_t build_rows, kg_tune_config *out) {
  return guarded([&] {
    if (!ctx || !out) fail(KG_ERR_INTERNAL, "no context or output");
    const TuneCfg c = tune_for(ctx, build_rows);
    *out = kg_tune_config{c.load_pct, c.scatter_rows, c.merge_lanes, 0};
  });
}

int kg_tune_set(kg_ctx *ctx, int64_t build_rows, const kg_tune_config *cfg) {
  return guarded([&] {
    if (!ctx || !cfg) fail(KG_ERR_INTERNAL, "no context or config");
    if (cfg->load_factor_pct == 0) {  // back to the class's untuned default
      tune_clear(ctx, build_rows);
      return;
    }
    check_cfg(*cfg);
    TuneCfg c;
    c.load_pct = cfg->load_factor_pct;
    c.scatter_rows = cfg->scatter_rows;
    c.merge_lanes = cfg->merge_lanes;
    tune_store(ctx, build_rows, c);
  });
}

// A context's tuned shapes as text, one line per size class: "class
// load_pct scatter_rows merge_lanes" (class = ceil log2 of the build rows).
int kg_tune_save(kg_ctx *ctx, const char *path) {
  return guarded([&] {
    if (!ctx || !path) fail(KG_ERR_INTERNAL, "no context or path");
    std::map<int, TuneCfg> t;
    {
      std::lock_guard<std::mutex> g(ctx->tune_mu);
      t = ctx->tune;
    }
    FILE *f = fopen(path, "w");
    if (!f) fail(KG_ERR_INTERNAL, std::string("cannot write ") + path);
    for (const auto &kv : t)
      fprintf(f, "%d %d %d %d\n", kv.first, kv.second.load_pct, kv.second.scatter_rows,
              kv.second.merge_lanes);
    if (fclose(f) != 0) fail(KG_ERR_INTERNAL, std::string("cannot write ") + path);
  });
}

int kg_tune_load(kg_ctx *ctx, const char *path, int32_t *classes) {
  return guarded([&] {
    if (!ctx || !path) fail(KG_ERR_INTERNAL, "no context or path");
    FILE *f = fopen(path, "r");
    if (!f) fail(KG_ERR_INTERNAL, std::string("cannot read ") + path);
    std::map<int, TuneCfg> t;
    int cls, load, rows, lanes, got;
    while ((got = fscanf(f, "%d %d %d %d", &cls, &load, &rows, &lanes)) == 4) {
      const kg_tune_config c{load, rows, lanes, 0};
      try {
        check_cfg(c);
      } catch (...) {
        fclose(f);
        throw;
      }
      if (cls < 0 || cls > 62) {
        fclose(f);
        fail(KG_ERR_UNSUPPORTED, "a size class outside 0 to 62");
      }
      t[cls] = TuneCfg{load, rows, lanes};
    }
    const bool clean = got == EOF;
    fclose(f);
    if (!clean) fail(KG_ERR_UNSUPPORTED, std::string("malformed tuning file ") + path);
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    for (const auto &kv : t) ctx->tune[kv.first] = kv.second;
    if (classes) *classes = (int32_t)t.size();
  });
}

int kg_autotune_joinagg(kg_ctx *ctx, int64_t build_rows, int64_t probe_rows, kg_tune_config *best,
                        double *best_ms, int32_t *candidates) {
  return guarded([&] {
    if (!ctx || !best || !best_ms) fail(KG_ERR_INTERNAL, "no context or output");
    if (build_rows < 1 || probe_rows < 1) fail(KG_ERR_UNSUPPORTED, "empty tuning shape");
    if (build_rows > ((int64_t)1 << 31) - 16)
      fail(KG_ERR_UNSUPPORTED, "a build side of 2^31 rows or more");
    // the probe side is timed on at most 2^26 rows (its time is linear in them)
    const int64_t nb = build_rows, np = std::min<int64_t>(probe_rows, (int64_t)1 << 26);
    StreamScope kg_scope(ctx->stream);
    DevBuf bk(ctx, (uint64_t)nb * 8), fk(ctx, (uint64_t)np * 8),
        fv(ctx, (uint64_t)(np + 31) / 32 * 4 + 8), amt(ctx, (uint64_t)np * 8);
    tck(kg_gen_join_build(ctx, 0x7E57ull, 0, nb, nb, bk.get<int64_t>(), nullptr));
    tck(kg_gen_join_probe(ctx, 0x7E58ull, 0, np, nb + nb / 5, 10, fk.get<int64_t>(),
                          fv.get<uint8_t>(), amt.get<int64_t>()));
    const int loads[3] = {50, 25, 75}, rows[3] = {12, 8, 4}, lanes[2] = {1024, 512};
    TuneCfg win;
    double win_ms = -1.0;
    int n = 0;
    // a failed candidate leaves the class's shape as it was (untuned stays
    // untuned)
    const TuneCfg saved = tune_for(ctx, nb);
    const bool had = tune_has(ctx, nb);
    struct Restore {
      kg_ctx *ctx;
      int64_t nb;
      TuneCfg c;
      bool had;
      bool armed = true;
      ~Restore() {
        if (!armed) return;
        if (had) tune_store(ctx, nb, c);
        else tune_clear(ctx, nb);
      }
    } restore{ctx, nb, saved, had};
    for (int li = 0; li < 3; li++)
      for (int ri = 0; ri < 3; ri++)
        for (int mi = 0; mi < 2; mi++) {
          TuneCfg c;
          c.load_pct = loads[li];
          c.scatter_rows = rows[ri];
          c.merge_lanes = lanes[mi];
          tune_store(ctx, nb, c);
          kg_join_spec spec{};
          spec.join_type = KG_JOIN_INNER;
          spec.num_keys = 1;
          kg_hash_join *j = nullptr;
          tck(kg_join_create(ctx, &spec, (uint64_t)nb, &j));
          std::unique_ptr<kg_hash_join, void (*)(kg_hash_join *)> guard(j, kg_join_destroy);
          KG_HIP(hipDeviceSynchronize());
          const double t0 = now_ms();
          kg_column k = tune_col(bk.get(), nb);
          tck(kg_join_build_sink(j, 1, &k, nb));
          tck(kg_join_build_finalize(j));
          KG_HIP(hipDeviceSynchronize());
          const double build_ms = now_ms() - t0;
          kg_column it = tune_col(nullptr, 0);
          kg_agg_spec aggs[2] = {{KG_AGG_SUM, 0}, {KG_AGG_COUNT_STAR, 0}};
          tck(kg_joinagg_create(j, 1, &it, 2, aggs, 0));
          kg_column cols[2] = {tune_col(fk.get(), np, fv.get<uint8_t>()), tune_col(amt.get(), np)};
          int32_t ic[1] = {1};
          tck(kg_joinagg_probe(j, 2, cols, np, ic));  // untimed: first use of the shape
          KG_HIP(hipDeviceSynchronize());
          const double t1 = now_ms();
          for (int rep = 0; rep < 2; rep++) tck(kg_joinagg_probe(j, 2, cols, np, ic));
          KG_HIP(hipDeviceSynchronize());
          // build once per probe side of probe_rows rows
          const double ms = build_ms + (now_ms() - t1) / 2 * ((double)probe_rows / np);
          n++;
          if (win_ms < 0 || ms < win_ms) {
            win_ms = ms;
            win = c;
          }
        }
    restore.armed = false;
    tune_store(ctx, nb, win);
    *best = kg_tune_config{win.load_pct, win.scatter_rows, win.merge_lanes, 0};
    *best_ms = win_ms;
    if (candidates) *candidates = n;
  });
}

}  // extern "C"
