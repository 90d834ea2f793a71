// kg_runtime.hip -- context, HBM budget, errors, batches and profiling of
// libkgpu (the C ABI declared in include/kgpu.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "kg_common.hpp"

namespace kg {

static thread_local std::string g_last_error;

const char *class_of(int code) {
  switch (code) {
    case KG_ERR_UNSUPPORTED: return "KRISHIV_UNSUPPORTED_FEATURE";
    case KG_ERR_OVERFLOW: return "ARITHMETIC_OVERFLOW";
    case KG_ERR_OUT_OF_RANGE: return "NUMERIC_VALUE_OUT_OF_RANGE.WITH_SUGGESTION";
    case KG_ERR_MEMORY: return "UNABLE_TO_ACQUIRE_MEMORY";
    default: return "INTERNAL_ERROR";
  }
}

void set_last_error(int code, const std::string &msg) {
  g_last_error = std::string(class_of(code)) + ": " + msg;
}

// Device memory with a per-context cache of freed blocks; sizes are rounded
// (powers of two up to 2 MiB, then multiples of 2 MiB) so blocks are reused
// across operators of the same shapes.  The HBM budget counts live bytes.
//
// A block may be handed out again only when no kernel still uses it.  A
// buffer freed by an operator whose work all runs on one stream waits, as a
// pending block, for an event recorded on that stream at the free, and for
// the use fences then outstanding (reads of caller-owned device columns,
// which the caller may free once the call returns, as the reference's
// `Sink::sink` borrows its batch).  Other frees synchronize the device,
// as hipFree would.
Fence::Fence(hipStream_t s) {
  if (hipEventCreateWithFlags(&e_, hipEventDisableTiming) != hipSuccess) {
    e_ = nullptr;
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();  // no event: order by a full synchronize
    return;
  }
  if (hipEventRecord(e_, s) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipEventDestroy(e_);
    e_ = nullptr;
    (void)hipDeviceSynchronize();
  }
}

Fence::~Fence() {
  if (e_) (void)hipEventDestroy(e_);
}

bool Fence::done() const {
  if (!e_) return true;
  hipError_t q = hipEventQuery(e_);
  if (q == hipErrorNotReady) return false;
  if (q != hipSuccess) (void)hipGetLastError();
  return true;
}

static thread_local hipStream_t g_alloc_stream = nullptr;

hipStream_t alloc_stream() { return g_alloc_stream; }

StreamScope::StreamScope(hipStream_t s) : prev_(g_alloc_stream) { g_alloc_stream = s; }
StreamScope::~StreamScope() { g_alloc_stream = prev_; }

// Size classes: powers of two up to 2 MB; above, eight classes per power of
// two (at most 12.5 % over the request) in whole 2 MB pages, so that blocks
// freed by one request are reused by the next of a similar size.  A class
// size maps to itself.
static uint64_t round_alloc(uint64_t b) {
  if (b <= (2u << 20)) {
    uint64_t r = 256;
    while (r < b) r <<= 1;
    return r;
  }
  const int e = 63 - __builtin_clzll(b - 1);  // 2^e < b <= 2^(e+1)
  const uint64_t step = std::max<uint64_t>(2u << 20, 1ull << (e - 3));
  return (b + step - 1) / step * step;
}

// Freed bytes kept for reuse, at most: hipFree waits for every stream of
// the device, so blocks are returned to HIP only past this or when hipMalloc
// fails (which releases the whole cache and retries).  Seven eighths of HBM:
// an operator that cycles through more than a quarter of it (the grace
// join's partition pieces at 1.5B x 6B rows, 120 GB a step) would otherwise
// free and allocate half of them again every step.
static uint64_t cache_cap(const kg_ctx *ctx) { return ctx->total_hbm / 8 * 7; }

// (cache_mu held) moves the pending blocks whose events have completed (all
// of them with `all`, after a device synchronize) into the cache, keeping
// at most cache_cap bytes cached.
static void reclaim_locked(kg_ctx *ctx, bool all) {
  size_t keep = 0;
  for (size_t i = 0; i < ctx->pending.size(); i++) {
    auto &pb = ctx->pending[i];
    bool ready = all;
    if (!ready) {
      ready = true;
      for (auto &f : pb.waits)
        if (!f->done()) {
          ready = false;
          break;
        }
    }
    if (!ready) {
      if (keep != i) ctx->pending[keep] = std::move(pb);
      keep++;
      continue;
    }
    if (ctx->cached + pb.rb > cache_cap(ctx)) {
      (void)hipFree(pb.p);
      ctx->n_frees++;
    } else {
      ctx->cache.emplace(pb.rb, pb.p);
      ctx->cached += pb.rb;
    }
  }
  ctx->pending.resize(keep);
  size_t k = 0;
  for (size_t i = 0; i < ctx->use_fences.size(); i++)
    if (!ctx->use_fences[i]->done()) ctx->use_fences[k++] = ctx->use_fences[i];
  ctx->use_fences.resize(k);
}

static void cache_release_all(kg_ctx *ctx) {
  (void)hipDeviceSynchronize();
  std::lock_guard<std::mutex> g(ctx->cache_mu);
  reclaim_locked(ctx, true);
  for (auto &kv : ctx->cache) (void)hipFree(kv.second);
  ctx->n_frees += ctx->cache.size();
  ctx->n_releases++;
  ctx->cache.clear();
  ctx->cached = 0;
}

void note_column_use(kg_ctx *ctx, hipStream_t s) {
  if (!ctx) return;
  auto f = std::make_shared<Fence>(s);
  std::lock_guard<std::mutex> g(ctx->cache_mu);
  if (ctx->use_fences.size() >= 64) {  // drop the completed ones
    size_t k = 0;
    for (size_t i = 0; i < ctx->use_fences.size(); i++)
      if (!ctx->use_fences[i]->done()) ctx->use_fences[k++] = ctx->use_fences[i];
    ctx->use_fences.resize(k);
  }
  ctx->use_fences.push_back(std::move(f));
}

void *dev_alloc(kg_ctx *ctx, uint64_t bytes) {
  if (!ctx) fail(KG_ERR_INTERNAL, "no context");
  const uint64_t rb = round_alloc(bytes);
  uint64_t prev = ctx->in_use.fetch_add(rb);
  if (ctx->budget && prev + rb > ctx->budget) {
    ctx->in_use.fetch_sub(rb);
    fail(KG_ERR_MEMORY, "HBM budget of " + std::to_string(ctx->budget) + " bytes exceeded (" +
                            std::to_string(prev) + " held, " + std::to_string(rb) +
                            " requested)");
  }
  {
    std::lock_guard<std::mutex> g(ctx->cache_mu);
    auto it = ctx->cache.find(rb);
    if (it == ctx->cache.end() && !ctx->pending.empty()) {
      reclaim_locked(ctx, false);
      it = ctx->cache.find(rb);
    }
    if (it != ctx->cache.end()) {
      void *p = it->second;
      ctx->cache.erase(it);
      ctx->cached -= rb;
      ctx->n_hits++;
      return p;
    }
  }
  void *p = nullptr;
  ctx->n_mallocs++;
  hipError_t e = hipMalloc(&p, rb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    cache_release_all(ctx);
    ctx->n_mallocs++;
    e = hipMalloc(&p, rb);
  }
  if (e != hipSuccess) {
    ctx->in_use.fetch_sub(rb);
    (void)hipGetLastError();
    fail(KG_ERR_MEMORY, "hipMalloc of " + std::to_string(rb) + " bytes: " + hipGetErrorString(e));
  }
  return p;
}

void dev_free(kg_ctx *ctx, void *p, uint64_t bytes, hipStream_t stream) {
  if (!p) return;
  if (!ctx) {
    (void)hipFree(p);
    return;
  }
  const uint64_t rb = round_alloc(bytes);
  ctx->in_use.fetch_sub(rb);
  if (!stream) {
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> g(ctx->cache_mu);
    reclaim_locked(ctx, true);
    if (ctx->cached + rb > cache_cap(ctx)) {
      (void)hipFree(p);
      ctx->n_frees++;
      return;
    }
    ctx->cache.emplace(rb, p);
    ctx->cached += rb;
    return;
  }
  auto f = std::make_shared<Fence>(stream);
  std::lock_guard<std::mutex> g(ctx->cache_mu);
  kg_ctx::Pending pb{rb, p, {std::move(f)}};
  for (auto &u : ctx->use_fences)
    if (!u->done()) pb.waits.push_back(u);
  ctx->pending.push_back(std::move(pb));
}

static int size_class(int64_t rows) {
  int b = 0;
  while (b < 62 && (1ll << b) < rows) b++;
  return b;
}

// A class nobody tuned: the load factors kg_autotune_joinagg picked on an
// MI355X for the join aggregate (builds up to 2^20 keys: 25 %, to 2^23: 75 %,
// larger: 50 %), the other fields at their defaults.
static TuneCfg untuned(int cls) {
  TuneCfg c;
  c.load_pct = cls <= 20 ? 25 : cls <= 23 ? 75 : 50;
  return c;
}

TuneCfg tune_for(kg_ctx *ctx, int64_t rows) {
  const int cls = size_class(rows);
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  auto it = ctx->tune.find(cls);
  return it == ctx->tune.end() ? untuned(cls) : it->second;
}

void tune_store(kg_ctx *ctx, int64_t rows, const TuneCfg &c) {
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  ctx->tune[size_class(rows)] = c;
}

void tune_clear(kg_ctx *ctx, int64_t rows) {
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  ctx->tune.erase(size_class(rows));
}

bool tune_has(kg_ctx *ctx, int64_t rows) {
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  return ctx->tune.count(size_class(rows)) != 0;
}

int type_width(int t) {
  switch (t) {
    case KG_T_BOOL: return 0;
    case KG_T_INT8:
    case KG_T_UINT8: return 1;
    case KG_T_INT16:
    case KG_T_UINT16: return 2;
    case KG_T_INT32:
    case KG_T_UINT32:
    case KG_T_FLOAT32:
    case KG_T_DATE32: return 4;
    case KG_T_INT64:
    case KG_T_UINT64:
    case KG_T_FLOAT64:
    case KG_T_TIMESTAMP:
    case KG_T_DECIMAL64: return 8;
    case KG_T_DECIMAL128: return 16;
    default: return -1;
  }
}
bool is_integral(int t) {
  return t == KG_T_INT8 || t == KG_T_INT16 || t == KG_T_INT32 || t == KG_T_INT64;
}
bool is_float(int t) { return t == KG_T_FLOAT32 || t == KG_T_FLOAT64; }
bool is_decimal(int t) { return t == KG_T_DECIMAL64 || t == KG_T_DECIMAL128; }
bool is_int_like(int t) {
  int w = type_width(t);
  return w >= 0 && w <= 8 && !is_float(t);
}

void check_column(const kg_column &c, int64_t nrows, const char *what) {
  if (type_width(c.type) < 0) fail(KG_ERR_UNSUPPORTED, std::string(what) + ": column type " + std::to_string(c.type));
  if (c.length < nrows) fail(KG_ERR_INTERNAL, std::string(what) + ": column shorter than the batch");
  if (c.offset < 0) fail(KG_ERR_INTERNAL, std::string(what) + ": negative offset");
  if (nrows > 0 && !c.values) fail(KG_ERR_INTERNAL, std::string(what) + ": no values buffer");
}

namespace {
std::mutex g_lds_mu;
std::map<std::pair<int, const void *>, size_t> g_lds_set;
}  // namespace

hipError_t set_dynamic_lds(const void *fn, size_t bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(g_lds_mu);
  auto it = g_lds_set.find({dev, fn});
  if (it != g_lds_set.end() && it->second >= bytes) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) g_lds_set[{dev, fn}] = bytes;
  return e;
}

// Each calling thread keeps one 256 KB pinned buffer for the life of the
// process (operators are driven by a few long-lived threads).
namespace {
constexpr size_t kPinnedReadBytes = 256u << 10;
thread_local uint8_t *t_pinned = nullptr;
thread_local size_t t_pinned_used = 0;
}  // namespace

PinnedReads::PinnedReads() : mark_(t_pinned_used) {
  if (!t_pinned) KG_HIP(hipHostMalloc((void **)&t_pinned, kPinnedReadBytes, hipHostMallocDefault));
}

PinnedReads::~PinnedReads() { t_pinned_used = mark_; }

void *PinnedReads::take_bytes(size_t bytes) {
  const size_t at = (t_pinned_used + 63) & ~(size_t)63;
  if (at + bytes > kPinnedReadBytes) fail(KG_ERR_INTERNAL, "pinned readback space exhausted");
  t_pinned_used = at + bytes;
  return t_pinned + at;
}

DCol Staged::add(const kg_column &c, int64_t nrows) {
  DCol d{};
  d.type = c.type;
  d.precision = c.precision;
  d.scale = c.scale;
  d.width = type_width(c.type);
  if (c.flags & KG_COL_DEVICE) {
    d.offset = c.offset;
    d.values = c.values;
    d.validity = c.validity;
    device_ = true;
    return d;
  }
  // Host column: copy the rows [offset, offset + nrows) (bit-packed buffers
  // keep their bit offset within the first byte).
  d.offset = (d.width == 0 || c.validity) ? (c.offset & 7) : 0;
  int64_t first = d.width == 0 ? c.offset >> 3 : c.offset - d.offset;
  uint64_t vbytes = d.width == 0 ? (uint64_t)ceil_div(nrows + d.offset, 8)
                                 : (uint64_t)(nrows + d.offset) * (uint64_t)d.width;
  if (nrows > 0) {
    bufs_.emplace_back(ctx_, vbytes);
    const uint8_t *src = static_cast<const uint8_t *>(c.values) +
                         (d.width == 0 ? first : first * d.width);
    KG_HIP(hipMemcpyAsync(bufs_.back().get(), src, vbytes, hipMemcpyHostToDevice, s_));
    d.values = bufs_.back().get();
  }
  if (c.validity && nrows > 0) {
    uint64_t nb = (uint64_t)ceil_div(nrows + d.offset, 8);
    bufs_.emplace_back(ctx_, nb);
    KG_HIP(hipMemcpyAsync(bufs_.back().get(), c.validity + (c.offset >> 3), nb,
                          hipMemcpyHostToDevice, s_));
    d.validity = bufs_.back().get<uint8_t>();
  }
  // The caller may free its buffers once we return.
  KG_HIP(hipStreamSynchronize(s_));
  return d;
}

// Streams are pooled per device and never destroyed: a freed buffer may
// still order its reuse by an event on the stream of an operator that is
// gone, so a stream handle stays valid for the life of the process.
namespace {
std::mutex g_stream_mu;
std::map<int, std::vector<hipStream_t>> g_stream_pool;
std::map<hipStream_t, int> g_stream_device;
}  // namespace

hipStream_t make_stream(int device) {
  KG_HIP(hipSetDevice(device));
  {
    std::lock_guard<std::mutex> g(g_stream_mu);
    auto &v = g_stream_pool[device];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s;
  KG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::lock_guard<std::mutex> g(g_stream_mu);
  g_stream_device[s] = device;
  return s;
}

void destroy_stream(hipStream_t s) {
  if (!s) return;
  (void)hipStreamSynchronize(s);
  std::lock_guard<std::mutex> g(g_stream_mu);
  auto it = g_stream_device.find(s);
  if (it == g_stream_device.end()) return;
  g_stream_pool[it->second].push_back(s);
}

// ---------------------------------------------------------------- profiling
namespace {
struct ProfEntry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  int64_t launches = 0;
};
std::mutex g_prof_mu;
std::map<std::string, ProfEntry> g_prof;
std::atomic<bool> g_prof_on{false};
}  // namespace

ProfScope::ProfScope(const char *family, hipStream_t s) : family_(family), s_(s) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return;
  if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) {
    a_ = b_ = nullptr;
    return;
  }
  (void)hipEventRecord(a_, s_);
}

ProfScope::~ProfScope() {
  if (!a_) return;
  (void)hipEventRecord(b_, s_);
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof[family_].pending.emplace_back(a_, b_);
}

}  // namespace kg

using namespace kg;

extern "C" {

const char *kg_last_error(void) { return g_last_error.c_str(); }
int kg_abi_version(void) { return KG_ABI_VERSION; }

int kg_ctx_create(int device, uint64_t hbm_budget_bytes, kg_ctx **out) {
  return guarded([&] {
    if (!out) fail(KG_ERR_INTERNAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    KG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(KG_ERR_INTERNAL, "no such device");
    KG_HIP(hipSetDevice(device));
    auto *c = new kg_ctx();
    c->device = device;
    size_t free_b = 0, total_b = 0;
    KG_HIP(hipMemGetInfo(&free_b, &total_b));
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(free_b * 0.9);
    c->total_hbm = total_b;
    hipDeviceProp_t prop;
    KG_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    c->stream = make_stream(device);
    *out = c;
  });
}

void kg_ctx_destroy(kg_ctx *ctx) {
  if (!ctx) return;
  (void)hipDeviceSynchronize();
  cache_release_all(ctx);
  destroy_stream(ctx->stream);
  delete ctx;
}

uint64_t kg_ctx_hbm_in_use(const kg_ctx *ctx) { return ctx ? ctx->in_use.load() : 0; }

int kg_ctx_mem_stats(kg_ctx *ctx, kg_mem_stats *out) {
  return guarded([&] {
    if (!ctx || !out) fail(KG_ERR_INTERNAL, "no context or output");
    std::lock_guard<std::mutex> g(ctx->cache_mu);
    out->in_use = ctx->in_use.load();
    out->cached = ctx->cached;
    out->hip_mallocs = ctx->n_mallocs.load();
    out->hip_frees = ctx->n_frees;
    out->cache_hits = ctx->n_hits;
    out->releases = ctx->n_releases;
  });
}

int kg_dev_alloc(kg_ctx *ctx, uint64_t bytes, void **out) {
  return guarded([&] { *out = dev_alloc(ctx, bytes); });
}

int kg_dev_free(kg_ctx *ctx, void *p) {
  return guarded([&] {
    // Plain allocations are not size-tracked by pointer; ask HIP for it.
    if (!p) return;
    size_t sz = 0;
    hipDeviceptr_t base = nullptr;
    if (hipMemGetAddressRange(&base, &sz, (hipDeviceptr_t)p) != hipSuccess) sz = 0;
    dev_free(ctx, p, sz);
  });
}

int kg_memcpy_h2d(kg_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  return guarded([&] {
    KG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_memcpy_d2h(kg_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  return guarded([&] {
    KG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_ctx_synchronize(kg_ctx *ctx) {
  return guarded([&] {
    KG_HIP(hipSetDevice(ctx->device));
    KG_HIP(hipDeviceSynchronize());
  });
}

int64_t kg_batch_num_rows(const kg_batch *b) { return b ? b->nrows : 0; }
int32_t kg_batch_num_columns(const kg_batch *b) { return b ? (int32_t)b->cols.size() : 0; }

int kg_batch_column(const kg_batch *b, int32_t i, kg_column *out) {
  return guarded([&] {
    if (!b || i < 0 || i >= (int32_t)b->cols.size()) fail(KG_ERR_INTERNAL, "column index");
    const auto &c = b->cols[i];
    out->type = c.type;
    out->precision = c.precision;
    out->scale = c.scale;
    out->flags = KG_COL_DEVICE;
    out->length = b->nrows;
    out->offset = 0;
    out->values = c.values.get();
    out->validity = c.validity.bytes() ? c.validity.get<uint8_t>() : nullptr;
  });
}

int kg_batch_column_to_host(const kg_batch *b, int32_t i, void *values, uint8_t *valid_bytes) {
  return guarded([&] {
    if (!b || i < 0 || i >= (int32_t)b->cols.size()) fail(KG_ERR_INTERNAL, "column index");
    const auto &c = b->cols[i];
    int w = type_width(c.type);
    hipStream_t s = b->ctx->stream;
    int64_t n = b->nrows;
    if (n == 0) return;
    if (w == 0) {
      std::vector<uint8_t> bits((size_t)ceil_div(n, 8));
      KG_HIP(hipMemcpyAsync(bits.data(), c.values.get(), bits.size(), hipMemcpyDeviceToHost, s));
      KG_HIP(hipStreamSynchronize(s));
      for (int64_t r = 0; r < n; r++)
        static_cast<uint8_t *>(values)[r] = (bits[r >> 3] >> (r & 7)) & 1;
    } else if (values) {
      KG_HIP(hipMemcpyAsync(values, c.values.get(), (size_t)n * w, hipMemcpyDeviceToHost, s));
    }
    if (valid_bytes) {
      if (c.validity.bytes()) {
        std::vector<uint8_t> bits((size_t)ceil_div(n, 8));
        KG_HIP(hipMemcpyAsync(bits.data(), c.validity.get(), bits.size(), hipMemcpyDeviceToHost, s));
        KG_HIP(hipStreamSynchronize(s));
        for (int64_t r = 0; r < n; r++) valid_bytes[r] = (bits[r >> 3] >> (r & 7)) & 1;
      } else {
        memset(valid_bytes, 1, (size_t)n);
      }
    }
    KG_HIP(hipStreamSynchronize(s));
  });
}

void kg_batch_release(kg_batch *b) { delete b; }

void kg_prof_enable(int on) { g_prof_on.store(on != 0); }

void kg_prof_reset(void) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto &kv : g_prof)
    for (auto &p : kv.second.pending) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  g_prof.clear();
}

int kg_prof_get(const char *family, double *total_ms, int64_t *launches) {
  return guarded([&] {
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto it = g_prof.find(family);
    if (it == g_prof.end()) {
      *total_ms = 0.0;
      *launches = 0;
      return;
    }
    ProfEntry &e = it->second;
    for (auto &p : e.pending) {
      KG_HIP(hipEventSynchronize(p.second));
      float ms = 0.f;
      KG_HIP(hipEventElapsedTime(&ms, p.first, p.second));
      e.ms += ms;
      e.launches += 1;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    e.pending.clear();
    *total_ms = e.ms;
    *launches = e.launches;
  });
}

}  // extern "C"
