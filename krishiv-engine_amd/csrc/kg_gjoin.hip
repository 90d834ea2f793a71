// kg_gjoin.hip -- the grace hash join (crates/krishiv-exec/src/join/grace.rs)
// over the device hash join: while the build side fits its HBM budget it is
// the in-memory join; past the budget every build row and then every probe
// row is partitioned on the device (the row hash h's top bits,
// Hashing::partition level 0) and its partition's slice copied to pinned
// host memory.  After the last probe batch each partition is uploaded and
// joined by a fresh in-memory join; one whose build rows do not fit is
// split 8 ways on the next 3 bits of g (levels 1..6), and a semi or anti
// partition with few probe rows is joined with roles reversed (its probe
// rows indexed, its build rows streamed past them).
#include <hip/hip_runtime.h>

#include <deque>
#include <memory>
#include <vector>

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

constexpr uint32_t kSpillRadix = 6;   // level-0 partitions of a spilled build: 64
constexpr uint32_t kSpillParts = 1u << kSpillRadix;
constexpr uint32_t kSplitWays = 8;    // PARTITIONS (hash.rs:60-62)
constexpr uint32_t kMaxDepth = 6;     // MAX_DEPTH (join/keys.rs:59)
constexpr uint64_t kReverseRatio = 8; // REVERSE_RATIO (grace.rs:498)
constexpr uint64_t kGraceSeed = 0x5851F42D4C957F2DULL;  // seeds g of the split levels
constexpr uint64_t kIndexBytesPerRow = 16;  // directory and chain a build row

// Pinned host memory (hipHostMalloc): a spilled slice of one column.
class Pinned {
 public:
  Pinned() = default;
  explicit Pinned(size_t n) : n_(n) {
    if (n) KG_HIP(hipHostMalloc(&p_, n, hipHostMallocDefault));
  }
  Pinned(Pinned &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  Pinned &operator=(Pinned &&o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  Pinned(const Pinned &) = delete;
  ~Pinned() {
    if (p_) (void)hipHostFree(p_);
  }
  void *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

struct ColType {
  int32_t type, precision, scale;
};

// Rows of one partition in host memory: elements [0, off + rows) of each
// column, the rows starting at element `off` (< 8, so bitmaps are copied
// from a byte boundary).
struct Chunk {
  int64_t rows = 0, off = 0;
  std::vector<Pinned> vals, valid;  // valid[c] empty: no NULLs
  std::vector<kg_column> cols(const std::vector<ColType> &t) const {
    std::vector<kg_column> out(t.size());
    for (size_t c = 0; c < t.size(); c++) {
      kg_column &k = out[c];
      k = kg_column{};
      k.type = t[c].type;
      k.precision = t[c].precision;
      k.scale = t[c].scale;
      k.length = rows;
      k.offset = off;
      k.values = vals[c].get();
      k.validity = valid[c].size() ? static_cast<const uint8_t *>(valid[c].get()) : nullptr;
    }
    return out;
  }
};

struct Side {
  std::vector<Chunk> chunks;
  uint64_t rows = 0, bytes = 0;
};

struct Task {
  Side build, probe;
  uint32_t depth = 0;
};

// Columns of the given types with no rows (a join's probe schema for the
// NULL padding of its tail).
static std::vector<kg_column> type_cols(const std::vector<ColType> &t) {
  static const uint64_t zero[2] = {0, 0};
  std::vector<kg_column> c(t.size());
  for (size_t i = 0; i < t.size(); i++) {
    c[i] = kg_column{};
    c[i].type = t[i].type;
    c[i].precision = t[i].precision;
    c[i].scale = t[i].scale;
    c[i].values = zero;
  }
  return c;
}

static uint64_t col_bytes(int type, int64_t rows) {
  const int w = type_width(type);
  return (w == 0 ? (uint64_t)ceil_div(rows, 8) : (uint64_t)rows * w) + (uint64_t)ceil_div(rows, 8);
}

static uint64_t batch_bytes(const std::vector<ColType> &t, int64_t rows) {
  uint64_t b = (uint64_t)rows * kIndexBytesPerRow;
  for (const ColType &c : t) b += col_bytes(c.type, rows);
  return b;
}

__global__ void k_iota32(int32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

// NOT IN past the budget: a probe row with a NULL key goes nowhere (with a
// non-empty build it matches nothing and passes nothing).
__global__ void k_drop_null_keys(DCol key, int64_t n, uint8_t *pids, uint8_t drop) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (!dev::valid(key, i)) pids[i] = drop;
}

__global__ void k_count_invalid(DCol key, int64_t n, unsigned long long *cnt) {
  unsigned long long mine = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    mine += dev::valid(key, i) ? 0 : 1;
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(cnt, mine);
}

static void check_rc(int rc) {
  if (rc != KG_OK) fail(rc, kg_last_error());
}

}  // namespace kg

using namespace kg;

struct kg_grace_join {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  kg_join_spec spec{};
  uint64_t budget = 0;
  std::vector<kg_expr> filter;
  std::vector<ColType> build_types, probe_types;
  bool finalized = false, finished = false, spilled = false;
  // build batches held in HBM while under the budget
  std::vector<kg_batch *> staged;
  uint64_t staged_bytes = 0;
  // in memory
  kg_hash_join *mem = nullptr;
  kg_join_local *loc = nullptr;
  bool pending = false;
  // spilled
  std::vector<Task> level0;
  bool build_null = false;  // NOT IN: a NULL build key
  std::deque<Task> tasks;
  // the partition being joined
  kg_hash_join *cur = nullptr;
  kg_join_local *cur_loc = nullptr;
  Task task;
  bool reversed = false, cur_pending = false;
  size_t next_chunk = 0;
  kg_grace_join_stats stats{};
  ~kg_grace_join() {
    for (kg_batch *b : staged) kg_batch_release(b);
    if (cur_loc) kg_join_local_destroy(cur_loc);
    if (cur) kg_join_destroy(cur);
    if (loc) kg_join_local_destroy(loc);
    if (mem) kg_join_destroy(mem);
  }
};

namespace kg {

static bool is_naaj(const kg_grace_join *g) { return g->spec.join_type == KG_JOIN_NULL_AWARE_ANTI; }

static std::vector<ColType> types_of(int32_t ncols, const kg_column *cols) {
  std::vector<ColType> t(ncols);
  for (int c = 0; c < ncols; c++) t[c] = ColType{cols[c].type, cols[c].precision, cols[c].scale};
  return t;
}

static void check_types(const std::vector<ColType> &want, int32_t ncols, const kg_column *cols) {
  if ((int32_t)want.size() != ncols) fail(KG_ERR_INTERNAL, "the column count changed");
  for (int c = 0; c < ncols; c++)
    if (cols[c].type != want[c].type) fail(KG_ERR_INTERNAL, "a column type changed");
}

// A device copy of a batch (the take of every row).
static kg_batch *copy_batch(kg_grace_join *g, int32_t ncols, const kg_column *cols, int64_t n) {
  DevBuf rows(g->ctx, (uint64_t)std::max<int64_t>(n, 1) * 4);
  k_iota32<<<stream_grid(n, 256), 256, 0, g->ctx->stream>>>(rows.get<int32_t>(), n);
  KG_HIP(hipGetLastError());
  kg_batch *b = nullptr;
  check_rc(kg_gather(g->ctx, ncols, cols, rows.get<int32_t>(), n, &b));
  return b;
}

static std::vector<kg_column> batch_cols(const kg_batch *b) {
  std::vector<kg_column> c(b->cols.size());
  for (size_t i = 0; i < c.size(); i++) check_rc(kg_batch_column(b, (int32_t)i, &c[i]));
  return c;
}

// Rows [r0, r0 + n) of a device batch copied to pinned host memory.
static Chunk download(kg_grace_join *g, const kg_batch *b, int64_t r0, int64_t n) {
  Chunk ch;
  ch.rows = n;
  ch.off = r0 % 8;
  const int64_t e0 = r0 - ch.off, e1 = r0 + n;
  for (const auto &c : b->cols) {
    const int w = type_width(c.type);
    const uint64_t lo = w == 0 ? (uint64_t)e0 / 8 : (uint64_t)e0 * w;
    const uint64_t hi = w == 0 ? (uint64_t)ceil_div(e1, 8) : (uint64_t)e1 * w;
    Pinned v(hi - lo);
    KG_HIP(hipMemcpyAsync(v.get(), c.values.get<uint8_t>() + lo, hi - lo, hipMemcpyDeviceToHost,
                          g->s));
    ch.vals.push_back(std::move(v));
    if (c.validity.bytes()) {
      const uint64_t blo = (uint64_t)e0 / 8, bhi = (uint64_t)ceil_div(e1, 8);
      Pinned m(bhi - blo);
      KG_HIP(hipMemcpyAsync(m.get(), c.validity.get<uint8_t>() + blo, bhi - blo,
                            hipMemcpyDeviceToHost, g->s));
      ch.valid.push_back(std::move(m));
    } else {
      ch.valid.emplace_back();
    }
  }
  return ch;
}

// The rows of `cols` partitioned (level `depth`, `ways` partitions) and
// each partition's slice appended to dest[p] in host memory.
static void spill(kg_grace_join *g, int32_t ncols, const kg_column *cols, int64_t n, bool probe,
                  uint32_t depth, uint32_t ways, const std::vector<Side *> &dest) {
  if (n <= 0) return;
  const int32_t *kc = probe ? g->spec.probe_keys : g->spec.build_keys;
  std::vector<kg_column> keys(g->spec.num_keys);
  for (int k = 0; k < g->spec.num_keys; k++) keys[k] = cols[kc[k]];
  DevBuf pids(g->ctx, (uint64_t)n);
  check_rc(kg_join_partition(g->ctx, g->spec.num_keys, keys.data(), n, kSpillRadix, depth,
                             kGraceSeed, nullptr, pids.get<uint8_t>(), KG_COL_DEVICE));
  uint32_t parts = ways;
  if (probe && is_naaj(g)) {
    Staged st(g->ctx, g->ctx->stream);
    const DCol key = st.add(keys[0], n);
    k_drop_null_keys<<<stream_grid(n, 256), 256, 0, g->ctx->stream>>>(key, n, pids.get<uint8_t>(),
                                                                      (uint8_t)ways);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(g->ctx->stream));
    parts = ways + 1;  // the last one is dropped
  }
  std::vector<int64_t> offs(parts + 1);
  kg_batch *b = nullptr;
  check_rc(kg_split_partitions(g->ctx, ncols, cols, n, pids.get<uint8_t>(), parts, offs.data(), &b));
  std::unique_ptr<kg_batch, void (*)(kg_batch *)> hold(b, kg_batch_release);
  const std::vector<ColType> &t = probe ? g->probe_types : g->build_types;
  for (uint32_t p = 0; p < ways; p++) {
    const int64_t rows = offs[p + 1] - offs[p];
    if (rows == 0) continue;
    dest[p]->chunks.push_back(download(g, b, offs[p], rows));
    dest[p]->rows += (uint64_t)rows;
    const uint64_t bytes = batch_bytes(t, rows);
    dest[p]->bytes += bytes;
    g->stats.spill_bytes += bytes - (uint64_t)rows * kIndexBytesPerRow;
  }
  KG_HIP(hipStreamSynchronize(g->s));
}

static std::vector<Side *> level0_sides(kg_grace_join *g, bool probe) {
  std::vector<Side *> d(kSpillParts);
  for (uint32_t p = 0; p < kSpillParts; p++) d[p] = probe ? &g->level0[p].probe : &g->level0[p].build;
  return d;
}

static void note_build_nulls(kg_grace_join *g, int32_t ncols, const kg_column *cols, int64_t n) {
  if (!is_naaj(g) || g->build_null || n <= 0) return;
  Staged st(g->ctx, g->ctx->stream);
  const DCol key = st.add(cols[g->spec.build_keys[0]], n);
  if (!key.validity) return;
  DevBuf cnt(g->ctx, 8);
  KG_HIP(hipMemsetAsync(cnt.get(), 0, 8, g->ctx->stream));
  k_count_invalid<<<stream_grid(n, 256), 256, 0, g->ctx->stream>>>(key, n,
                                                                  cnt.get<unsigned long long>());
  KG_HIP(hipGetLastError());
  unsigned long long c = 0;
  KG_HIP(hipMemcpyAsync(&c, cnt.get(), 8, hipMemcpyDeviceToHost, g->ctx->stream));
  KG_HIP(hipStreamSynchronize(g->ctx->stream));
  if (c) g->build_null = true;
}

// Past the budget: the staged batches go to their level-0 partitions.
static void start_spilling(kg_grace_join *g) {
  g->spilled = true;
  g->stats.spilled = 1;
  g->level0.resize(kSpillParts);
  const auto dest = level0_sides(g, false);
  for (kg_batch *b : g->staged) {
    const std::vector<kg_column> c = batch_cols(b);
    note_build_nulls(g, (int32_t)c.size(), c.data(), b->nrows);
    spill(g, (int32_t)c.size(), c.data(), b->nrows, false, 0, kSpillParts, dest);
    kg_batch_release(b);
  }
  g->staged.clear();
  g->staged_bytes = 0;
}

static kg_hash_join *new_join(kg_grace_join *g, const kg_join_spec &spec, uint64_t rows,
                              bool swap_filter) {
  kg_hash_join *j = nullptr;
  check_rc(kg_join_create(g->ctx, &spec, rows, &j));
  if (!g->filter.empty()) {
    std::vector<kg_expr> f = g->filter;
    if (swap_filter)
      for (kg_expr &e : f) {
        if (e.op == KG_X_PROBE) e.op = KG_X_BUILD;
        else if (e.op == KG_X_BUILD) e.op = KG_X_PROBE;
      }
    const int rc = kg_join_set_filter(j, (int32_t)f.size(), f.data());
    if (rc != KG_OK) {
      kg_join_destroy(j);
      check_rc(rc);
    }
  }
  return j;
}

// An empty build still fixes the build schema.
static void sink_empty(kg_hash_join *j, const std::vector<ColType> &t) {
  const std::vector<kg_column> c = type_cols(t);
  check_rc(kg_join_build_sink(j, (int32_t)c.size(), c.data(), 0));
}

// Whether a spilled partition is joined with roles reversed (grace.rs:501-514).
static bool reversible(const kg_grace_join *g, const Task &t) {
  const int jt = g->spec.join_type;
  return (jt == KG_JOIN_LEFT_SEMI || jt == KG_JOIN_LEFT_ANTI) && t.probe.rows > 0 &&
         t.probe.rows * kReverseRatio <= t.build.rows && t.probe.bytes <= g->budget;
}

// Takes the next task: split it again, or start its in-memory join.
static void start_task(kg_grace_join *g, Task t) {
  const bool rev = reversible(g, t);
  if (!rev && t.build.bytes > g->budget && t.depth < kMaxDepth) {
    // repartition on the next bits (grace.rs:1481-1522), build and probe alike
    std::vector<Task> sub(kSplitWays);
    std::vector<Side *> db(kSplitWays), dp(kSplitWays);
    for (uint32_t p = 0; p < kSplitWays; p++) {
      sub[p].depth = t.depth + 1;
      db[p] = &sub[p].build;
      dp[p] = &sub[p].probe;
    }
    for (Chunk &ch : t.build.chunks) {
      const std::vector<kg_column> c = ch.cols(g->build_types);
      spill(g, (int32_t)c.size(), c.data(), ch.rows, false, t.depth + 1, kSplitWays, db);
      ch = Chunk{};  // freed as it goes
    }
    for (Chunk &ch : t.probe.chunks) {
      const std::vector<kg_column> c = ch.cols(g->probe_types);
      spill(g, (int32_t)c.size(), c.data(), ch.rows, true, t.depth + 1, kSplitWays, dp);
      ch = Chunk{};
    }
    g->stats.repartitions++;
    for (Task &s : sub)
      if (s.build.rows || s.probe.rows) g->tasks.push_front(std::move(s));
    return;
  }
  kg_join_spec spec = g->spec;
  if (is_naaj(g)) spec.join_type = KG_JOIN_LEFT_ANTI;  // no NULL key on either side here
  const Side *bs = &t.build;
  const std::vector<ColType> *bt = &g->build_types;
  if (rev) {
    spec.join_type = g->spec.join_type == KG_JOIN_LEFT_SEMI ? KG_JOIN_RIGHT_SEMI : KG_JOIN_RIGHT_ANTI;
    for (int k = 0; k < spec.num_keys; k++) std::swap(spec.probe_keys[k], spec.build_keys[k]);
    bs = &t.probe;
    bt = &g->probe_types;
    g->stats.reversed++;
  }
  kg_hash_join *j = new_join(g, spec, bs->rows, rev);
  try {
    if (bs->chunks.empty()) sink_empty(j, *bt);
    for (const Chunk &ch : bs->chunks) {
      const std::vector<kg_column> c = ch.cols(*bt);
      check_rc(kg_join_build_sink(j, (int32_t)c.size(), c.data(), ch.rows));
    }
    check_rc(kg_join_build_finalize(j));
    check_rc(kg_join_local_create(j, &g->cur_loc));
  } catch (...) {
    kg_join_destroy(j);
    throw;
  }
  g->cur = j;
  g->task = std::move(t);
  g->reversed = rev;
  g->next_chunk = 0;
  g->cur_pending = false;
  g->stats.partitions_joined++;
  g->stats.max_depth = std::max(g->stats.max_depth, g->task.depth);
}

static void end_task(kg_grace_join *g) {
  kg_join_local_destroy(g->cur_loc);
  kg_join_destroy(g->cur);
  g->cur_loc = nullptr;
  g->cur = nullptr;
  g->task = Task{};
}

// The spilled partitions' next output batch (grace.rs:836-1031 `step`).
static void spilled_next(kg_grace_join *g, int64_t max_rows, kg_batch **out, int32_t *done) {
  for (;;) {
    if (g->cur) {
      if (g->cur_pending) {
        kg_batch *b = nullptr;
        int32_t d = 0;
        check_rc(kg_join_probe_next(g->cur_loc, max_rows, &b, &d));
        if (d) g->cur_pending = false;
        if (b && b->nrows > 0) {
          *out = b;
          return;
        }
        if (b) kg_batch_release(b);
        continue;
      }
      const Side &ps = g->reversed ? g->task.build : g->task.probe;
      const std::vector<ColType> &pt = g->reversed ? g->build_types : g->probe_types;
      if (g->next_chunk < ps.chunks.size()) {
        const Chunk &ch = ps.chunks[g->next_chunk++];
        const std::vector<kg_column> c = ch.cols(pt);
        check_rc(kg_join_probe_push(g->cur_loc, (int32_t)c.size(), c.data(), ch.rows, nullptr, 0));
        g->cur_pending = true;
        continue;
      }
      const std::vector<kg_column> ptc = type_cols(pt);
      kg_batch *b = nullptr;
      int32_t d = 0;
      check_rc(kg_join_tail_next(g->cur, (int32_t)ptc.size(), ptc.data(), max_rows, &b, &d));
      if (d) end_task(g);
      if (b && b->nrows > 0) {
        *out = b;
        return;
      }
      if (b) kg_batch_release(b);
      continue;
    }
    if (g->tasks.empty()) {
      *done = 1;
      return;
    }
    Task t = std::move(g->tasks.front());
    g->tasks.pop_front();
    start_task(g, std::move(t));
  }
}

}  // namespace kg

extern "C" {

int kg_grace_join_create(kg_ctx *ctx, const kg_join_spec *spec, uint64_t table_budget,
                         kg_grace_join **out) {
  return guarded([&] {
    *out = nullptr;
    auto g = std::make_unique<kg_grace_join>();
    g->ctx = ctx;
    g->spec = *spec;
    g->budget = table_budget;
    if (spec->num_keys < 1 || spec->num_keys > 4) fail(KG_ERR_UNSUPPORTED, "1 to 4 join keys");
    g->s = make_stream(ctx->device);
    *out = g.release();
  });
}

int kg_grace_join_set_filter(kg_grace_join *g, int32_t nops, const kg_expr *ops) {
  return guarded([&] {
    if (g->finalized) fail(KG_ERR_INTERNAL, "the residual filter is set before the build ends");
    if (is_naaj(g)) fail(KG_ERR_UNSUPPORTED, "a residual filter with NOT IN");
    g->filter.assign(ops, ops + nops);
  });
}

int kg_grace_join_build_sink(kg_grace_join *g, int32_t ncols, const kg_column *cols,
                             int64_t nrows) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    if (g->finalized) fail(KG_ERR_INTERNAL, "build sink after finalize");
    if (g->build_types.empty()) g->build_types = types_of(ncols, cols);
    else check_types(g->build_types, ncols, cols);
    for (int c = 0; c < ncols; c++) check_column(cols[c], nrows, "build column");
    if (nrows <= 0) return;
    const uint64_t bytes = batch_bytes(g->build_types, nrows);
    if (!g->spilled && g->staged_bytes + bytes > g->budget) start_spilling(g);
    if (g->spilled) {
      note_build_nulls(g, ncols, cols, nrows);
      spill(g, ncols, cols, nrows, false, 0, kSpillParts, level0_sides(g, false));
      return;
    }
    g->staged.push_back(copy_batch(g, ncols, cols, nrows));
    g->staged_bytes += bytes;
  });
}

int kg_grace_join_build_finalize(kg_grace_join *g) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    if (g->finalized) return;
    if (g->build_types.empty()) fail(KG_ERR_INTERNAL, "a build side with no schema");
    if (!g->spilled) {
      uint64_t rows = 0;
      for (kg_batch *b : g->staged) rows += (uint64_t)b->nrows;
      g->mem = new_join(g, g->spec, rows, false);
      if (g->staged.empty()) sink_empty(g->mem, g->build_types);
      for (kg_batch *b : g->staged) {
        const std::vector<kg_column> c = batch_cols(b);
        check_rc(kg_join_build_sink(g->mem, (int32_t)c.size(), c.data(), b->nrows));
      }
      check_rc(kg_join_build_finalize(g->mem));
      for (kg_batch *b : g->staged) kg_batch_release(b);
      g->staged.clear();
      check_rc(kg_join_local_create(g->mem, &g->loc));
    }
    g->finalized = true;
  });
}

int kg_grace_join_probe_push(kg_grace_join *g, int32_t ncols, const kg_column *cols,
                             int64_t nrows) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    if (!g->finalized) fail(KG_ERR_INTERNAL, "probe before the build is finalized");
    if (g->finished) fail(KG_ERR_INTERNAL, "probe after finish");
    if (g->probe_types.empty()) g->probe_types = types_of(ncols, cols);
    else check_types(g->probe_types, ncols, cols);
    if (!g->spilled) {
      if (g->pending) fail(KG_ERR_INTERNAL, "the previous probe batch's output is not taken");
      check_rc(kg_join_probe_push(g->loc, ncols, cols, nrows, nullptr, 0));
      g->pending = true;
      return;
    }
    for (int c = 0; c < ncols; c++) check_column(cols[c], nrows, "probe column");
    if (is_naaj(g) && g->build_null) return;  // NOT IN against a NULL: no row passes
    spill(g, ncols, cols, nrows, true, 0, kSpillParts, level0_sides(g, true));
  });
}

int kg_grace_join_finish(kg_grace_join *g) {
  return guarded([&] {
    if (!g->finalized) fail(KG_ERR_INTERNAL, "finish before the build is finalized");
    if (g->pending) fail(KG_ERR_INTERNAL, "the last probe batch's output is not taken");
    if (g->finished) return;
    g->finished = true;
    if (g->spilled && !(is_naaj(g) && g->build_null)) {
      for (Task &t : g->level0)
        if (t.build.rows || t.probe.rows) g->tasks.push_back(std::move(t));
      g->level0.clear();
    }
  });
}

int kg_grace_join_next(kg_grace_join *g, int64_t max_rows, kg_batch **out, int32_t *done) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    *out = nullptr;
    *done = 0;
    if (max_rows <= 0) fail(KG_ERR_INTERNAL, "max_rows must be positive");
    if (!g->finished) {
      if (!g->pending) {
        *done = 1;
        return;
      }
      check_rc(kg_join_probe_next(g->loc, max_rows, out, done));
      if (*done) g->pending = false;
      return;
    }
    if (!g->spilled) {
      const std::vector<kg_column> ptc = type_cols(g->probe_types);
      check_rc(kg_join_tail_next(g->mem, (int32_t)ptc.size(), ptc.data(), max_rows, out, done));
      return;
    }
    spilled_next(g, max_rows, out, done);
  });
}

int kg_grace_join_stats_get(const kg_grace_join *g, kg_grace_join_stats *out) {
  return guarded([&] { *out = g->stats; });
}

void kg_grace_join_destroy(kg_grace_join *g) {
  if (!g) return;
  hipStream_t s = g->s;
  delete g;
  if (s) destroy_stream(s);
}

}  // extern "C"
