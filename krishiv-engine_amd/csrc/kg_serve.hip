// kg_serve.hip -- batched, continuously served join-aggregate requests
// (include/kgpu.h, kg_server_*): SELECT key, SUM(pay), COUNT(*) FROM build
// JOIN probe ON key = fk GROUP BY key per request.
//
// A batch is every request admitted while a stream slot is free.  Its work
// is four phases, each one persistent-kernel launch over a device work list
// of (request, row range) tiles, so requests of 1K and 100M rows share a
// launch and balance across the 256 CUs:
//   init    the batch's own tables (16-byte slots {key, rows}) and records
//   build   build keys inserted into their request's table (CAS); rows counted
//   probe   probe rows looked up in their request's table (own or resident);
//           a hit adds rows to COUNT and pay*rows to SUM (64-bit, wrapping:
//           SUM(bigint) without ANSI overflow checks)
//   output  occupied records compacted into the request's output region
// Tiles are sized per batch so that each phase has about two per workgroup:
// a batch of a few small requests still spreads over every CU.
// The four launches read one device descriptor at a fixed address, so they
// are captured once per stream slot as a hipGraph and replayed per batch
// after the descriptor is uploaded.  Two slots (streams at default priority)
// let the next batch form and launch while one runs.
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <map>
#include <set>
#include <thread>

#include "kg_aggops.hpp"

namespace kg {

constexpr int kSrvBlock = 256;
constexpr int64_t kSrvMinTile = 512;    // rows (or slots) per tile, at least
constexpr int64_t kSrvTile = 1 << 16;   // and at most
constexpr int kSrvSlots = 2;            // batches in flight
constexpr int kSrvU = 4;                // probe rows in flight per lane

struct alignas(16) SSlot {
  uint64_t key;
  uint64_t rows;
};

struct SReq {
  const int64_t *bkeys;
  int64_t nb;
  const int64_t *pkeys;
  const uint8_t *pvalid;
  const int64_t *ppay;
  int64_t np;
  SSlot *slots;  // cap + 1 (slot cap: the key equal to EMPTY)
  uint64_t cap;
  int32_t log2cap;
  int32_t own;
  uint64_t *acc;  // (cap + 1) x {COUNT, SUM}; a group exists where COUNT != 0
  int64_t *out_key, *out_sum, *out_cnt;
};

struct STile {
  int32_t req;
  int32_t pad;
  int64_t lo, hi;
};

struct SDesc {
  int32_t nreq;
  int32_t ntiles[4];
  unsigned int next[4];
  SReq *reqs;
  STile *tiles[4];
  unsigned long long *groups;  // per request
};

__device__ __forceinline__ uint64_t sslot_start(uint64_t key, int log2cap) {
  return dev::table_hash(key, 0x3C6EF372FE94F82BULL) >> (64 - log2cap);
}

// Claims tiles of phase `ph` until none is left (every workgroup exits).
template <class F>
__device__ __forceinline__ void for_tiles(SDesc *D, int ph, F &&f) {
  __shared__ int s_t;
  while (true) {
    if (threadIdx.x == 0) s_t = (int)atomicAdd(&D->next[ph], 1u);
    __syncthreads();
    int t = s_t;
    __syncthreads();
    if (t >= D->ntiles[ph]) break;
    const STile tile = D->tiles[ph][t];
    f(D->reqs[tile.req], tile.req, tile.lo, tile.hi);
  }
}

__global__ void __launch_bounds__(kSrvBlock) k_srv_init(SDesc *D) {
  for_tiles(D, 0, [&](const SReq &r, int, int64_t lo, int64_t hi) {
    for (int64_t s = lo + threadIdx.x; s < hi; s += blockDim.x) {
      if (r.own) r.slots[s] = SSlot{kEmpty, 0};
      r.acc[2 * s] = 0;
      r.acc[2 * s + 1] = 0;
    }
  });
}

__device__ __forceinline__ void sbuild_row(SSlot *slots, uint64_t cap, int log2cap, uint64_t key) {
  uint64_t s;
  if (key == kEmpty) {
    s = cap;
  } else {
    s = sslot_start(key, log2cap);
    uint64_t mask = cap - 1;
    while (true) {
      uint64_t k = __hip_atomic_load(&slots[s].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k == key) break;
      if (k == kEmpty) {
        unsigned long long prev = atomicCAS((unsigned long long *)&slots[s].key, kEmpty, key);
        if (prev == kEmpty || prev == key) break;
      }
      s = (s + 1) & mask;
    }
  }
  atomicAdd((unsigned long long *)&slots[s].rows, 1ull);
}

__global__ void __launch_bounds__(kSrvBlock) k_srv_build(SDesc *D) {
  for_tiles(D, 1, [&](const SReq &r, int, int64_t lo, int64_t hi) {
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
      sbuild_row(r.slots, r.cap, r.log2cap, (uint64_t)r.bkeys[i]);
  });
}

// kSrvU rows a lane: their key and payload loads, then their home-slot loads
// (16 bytes: key and rows together), are issued before any is used.  The
// key equal to EMPTY lives in slot `cap`, whose key field stays EMPTY.
__global__ void __launch_bounds__(kSrvBlock) k_srv_probe(SDesc *D) {
  for_tiles(D, 2, [&](const SReq &r, int, int64_t lo, int64_t hi) {
    const uint64_t mask = r.cap - 1;
    for (int64_t i0 = lo; i0 < hi; i0 += (int64_t)kSrvBlock * kSrvU) {
      uint64_t key[kSrvU], pay[kSrvU], at[kSrvU];
      bool act[kSrvU];
#pragma unroll
      for (int u = 0; u < kSrvU; u++) {
        const int64_t i = i0 + (int64_t)u * kSrvBlock + threadIdx.x;
        act[u] = i < hi && (!r.pvalid || ((r.pvalid[i >> 3] >> (i & 7)) & 1));
        key[u] = act[u] ? (uint64_t)r.pkeys[i] : 0;
        pay[u] = act[u] ? (uint64_t)r.ppay[i] : 0;
      }
      uint64_t k0[kSrvU], m0[kSrvU];
#pragma unroll
      for (int u = 0; u < kSrvU; u++) {
        at[u] = key[u] == kEmpty ? r.cap : sslot_start(key[u], r.log2cap);
        const SSlot t = act[u] ? r.slots[at[u]] : SSlot{kEmpty, 0};
        k0[u] = t.key;
        m0[u] = t.rows;
      }
#pragma unroll
      for (int u = 0; u < kSrvU; u++) {
        uint64_t k = k0[u], m = m0[u], s = at[u];
        while (act[u] && k != key[u] && k != kEmpty) {
          s = (s + 1) & mask;
          const SSlot t = r.slots[s];
          k = t.key;
          m = t.rows;
        }
        if (act[u] && k == key[u] && m) {
          add_u64_nr<false>(r.acc + 2 * s, m);
          add_u64_nr<false>(r.acc + 2 * s + 1, pay[u] * m);
        }
      }
    }
  });
}

// Output positions are taken once per wave (a ballot and one atomic on the
// request's counter), not once per group.
__global__ void __launch_bounds__(kSrvBlock) k_srv_output(SDesc *D) {
  for_tiles(D, 3, [&](const SReq &r, int q, int64_t lo, int64_t hi) {
    for (int64_t s0 = lo; s0 < hi; s0 += blockDim.x) {
      const int64_t s = s0 + threadIdx.x;
      const bool has = s < hi && r.acc[2 * s] != 0;
      const uint64_t m = __ballot(has);
      if (!m) continue;
      const int lane = __lane_id();
      const int lead = __ffsll((unsigned long long)m) - 1;
      unsigned long long base = 0;
      if (lane == lead) base = atomicAdd(D->groups + q, (unsigned long long)__popcll(m));
      base = __shfl(base, lead);
      if (!has) continue;
      const uint64_t o = base + __popcll(m & ((1ull << lane) - 1));
      r.out_key[o] = s == (int64_t)r.cap ? (int64_t)kEmpty : (int64_t)r.slots[s].key;
      r.out_sum[o] = (int64_t)r.acc[2 * s + 1];
      r.out_cnt[o] = (int64_t)r.acc[2 * s];
    }
  });
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

static int slog2(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

}  // namespace kg

using namespace kg;

struct kg_server {
  kg_ctx *ctx = nullptr;
  int32_t max_batch = 256;
  bool use_graph = true;
  uint64_t arena_bytes = 0;
  struct Resident {
    DevBuf slots;
    uint64_t cap = 0;
    int log2cap = 0;
    kg_hash_join *join = nullptr;  // the same keys for large requests
  };
  std::vector<Resident> tables;
  struct Pending {
    kg_serve_request r;
    uint64_t cap = 0;
    uint64_t bytes = 0;
  };
  std::deque<Pending> queue;
  // Probes of kLargeProbe rows or more run on `nworkers` worker threads
  // (each request on its own stream), so batches keep launching meanwhile;
  // with two, one large request's host waits are filled by another's
  // kernels, at the price of each sharing the device with the other.
  struct Large {
    kg_serve_request r;
    kg_hash_join *table = nullptr;  // the resident table's join, if any
  };
  struct LargeDone {
    int64_t id;
    kg_batch *b;
    int code;
    std::string msg;
  };
  std::mutex large_mu;
  std::condition_variable large_cv;
  std::deque<Large> large;
  std::set<kg_hash_join *> large_busy;  // resident joins a worker is using
  std::deque<LargeDone> large_done;
  int64_t large_outstanding = 0;  // queued + running (main thread's count)
  // finished requests whose ids did not fit a progress call's id buffer:
  // reported first by the next call
  std::deque<int64_t> unreported;
  bool large_stop = false;
  std::vector<std::thread> large_workers;
  int nworkers = 1;
  struct Slot {
    hipStream_t s = nullptr;
    DevBuf arena;          // tables, records, outputs of the batch
    DevBuf desc_dev;       // SDesc + SReq[max_batch] + tiles + group counters
    uint8_t *desc_host = nullptr;  // pinned mirror of desc_dev
    uint64_t desc_bytes = 0;
    uint64_t max_tiles = 0;
    hipGraphExec_t graph = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Pending> reqs;
    std::vector<SReq> sreqs;
    // each request's result columns (key, SUM, COUNT; cap + 1 rows), written
    // by the output phase and handed to the caller's batch as they are
    std::vector<DevBuf> outs;
  };
  Slot slots[kSrvSlots];
  std::map<int64_t, kg_batch *> results;
  int64_t inflight = 0;
  uint64_t launches = 0;
};

namespace kg {

// Requests with probes this long run alone through the radix-partitioned
// join aggregate (kg_joinagg_*): their time is the HBM traffic of their rows,
// which the batch's tile kernels (one global atomic per matching row) would
// only slow down.  Smaller requests are packed into batches.
constexpr int64_t kLargeProbe = 4 << 20;
constexpr int kLargeWorkers = 1;  // default; KG_SERVE_LARGE_WORKERS (1 to 4) overrides
constexpr int64_t kBatchProbeRows = 8 << 20;

static void ck(int rc) {
  if (rc != KG_OK) throw Error{rc, std::string(kg_last_error())};
}

static kg_column dev_i64(const void *p, int64_t n, const uint8_t *valid = nullptr) {
  kg_column c{};
  c.type = KG_T_INT64;
  c.flags = KG_COL_DEVICE;
  c.length = n;
  c.values = p;
  c.validity = valid;
  return c;
}

static kg_hash_join *build_join(kg_ctx *ctx, const int64_t *keys, int64_t n) {
  kg_join_spec spec{};
  spec.join_type = KG_JOIN_INNER;
  spec.num_keys = 1;
  kg_hash_join *j = nullptr;
  ck(kg_join_create(ctx, &spec, (uint64_t)(n > 0 ? n : 1), &j));
  std::unique_ptr<kg_hash_join, void (*)(kg_hash_join *)> guard(j, kg_join_destroy);
  if (n > 0) {
    kg_column k = dev_i64(keys, n);
    ck(kg_join_build_sink(j, 1, &k, n));
  }
  ck(kg_join_build_finalize(j));
  return guard.release();
}

// One large request, start to finish (the partitioned fused aggregate).
static kg_batch *run_large(kg_server *S, const kg_serve_request &r, kg_hash_join *table) {
  const bool own = r.build_keys != nullptr;
  kg_hash_join *j = own ? build_join(S->ctx, r.build_keys, r.build_rows) : table;
  std::unique_ptr<kg_hash_join, void (*)(kg_hash_join *)> guard(own ? j : nullptr,
                                                                 kg_join_destroy);
  kg_column it = dev_i64(nullptr, 0);
  kg_agg_spec aggs[2] = {{KG_AGG_SUM, 0}, {KG_AGG_COUNT_STAR, 0}};
  ck(kg_joinagg_create(j, 1, &it, 2, aggs, 0));
  kg_column cols[2] = {dev_i64(r.probe_keys, r.probe_rows, r.probe_valid),
                       dev_i64(r.probe_pay, r.probe_rows)};
  int32_t ic[1] = {1};
  ck(kg_joinagg_probe(j, 2, cols, r.probe_rows, ic));
  kg_batch *b = nullptr;
  ck(kg_joinagg_result(j, &b));
  return b;
}

// A worker takes the oldest queued large request whose resident table (the
// join object its aggregate state lives in) no other worker is using.
static void large_loop(kg_server *S) {
  (void)hipSetDevice(S->ctx->device);
  while (true) {
    kg_server::Large L;
    {
      std::unique_lock<std::mutex> g(S->large_mu);
      auto pick = [&] {
        for (auto it = S->large.begin(); it != S->large.end(); ++it)
          if (!it->table || !S->large_busy.count(it->table)) return it;
        return S->large.end();
      };
      S->large_cv.wait(g, [&] { return S->large_stop || pick() != S->large.end(); });
      if (S->large_stop) return;
      auto it = pick();
      L = *it;
      S->large.erase(it);
      if (L.table) S->large_busy.insert(L.table);
    }
    kg_server::LargeDone d{L.r.id, nullptr, KG_OK, {}};
    try {
      d.b = run_large(S, L.r, L.table);
    } catch (const Error &e) {
      d.code = e.code;
      d.msg = e.msg;
    } catch (const std::exception &e) {
      d.code = KG_ERR_INTERNAL;
      d.msg = e.what();
    }
    std::lock_guard<std::mutex> g(S->large_mu);
    if (L.table) S->large_busy.erase(L.table);
    S->large_done.push_back(std::move(d));
    S->large_cv.notify_all();
  }
}

// An id to the caller's buffer, or to the next call's when it is full.
static void report(kg_server *S, int64_t id, int64_t *done_ids, int32_t max, int32_t *ndone) {
  if (*ndone < max) done_ids[(*ndone)++] = id;
  else S->unreported.push_back(id);
}

// Moves finished large requests into the results (waiting for one when
// `block`).  A failed one fails the call that meets it first, unless this
// call has already reported ids: it then stays queued for the next call, so
// no reported id is lost behind an error.
static void collect_large(kg_server *S, bool block, int64_t *done_ids, int32_t max,
                          int32_t *ndone) {
  std::unique_lock<std::mutex> g(S->large_mu);
  if (block) S->large_cv.wait(g, [&] { return !S->large_done.empty(); });
  while (!S->large_done.empty() && *ndone < max) {
    if (S->large_done.front().code != KG_OK && *ndone > 0) return;
    kg_server::LargeDone d = std::move(S->large_done.front());
    S->large_done.pop_front();
    S->large_outstanding--;
    if (d.code != KG_OK) {
      g.unlock();
      fail(d.code, "request " + std::to_string(d.id) + ": " + d.msg);
    }
    S->results[d.id] = d.b;
    done_ids[(*ndone)++] = d.id;
  }
}

static uint64_t req_cap(kg_server *S, const kg_serve_request &r) {
  if (!r.build_keys) {
    if (r.table < 0 || r.table >= (int32_t)S->tables.size()) fail(KG_ERR_INTERNAL, "no such table");
    return S->tables[r.table].cap;
  }
  uint64_t n = (uint64_t)(r.build_rows < 256 ? 256 : r.build_rows);
  return 1ull << slog2(n * 2);
}

// Arena bytes of a request: own table slots and records (its result
// columns are allocated apart, and handed over with its result).
static uint64_t req_bytes(uint64_t cap, bool own) {
  uint64_t per = (own ? 16 : 0) + 16;
  return ((cap + 1) * per + 255) / 256 * 256;
}

static void layout_desc(kg_server *S, kg_server::Slot &sl) {
  sl.max_tiles = 1 << 16;
  sl.desc_bytes = sizeof(SDesc) + sizeof(SReq) * S->max_batch + 4 * sizeof(STile) * sl.max_tiles +
                  8 * S->max_batch + 256;
  sl.desc_dev.reset(S->ctx, sl.desc_bytes);
  KG_HIP(hipHostMalloc((void **)&sl.desc_host, sl.desc_bytes, hipHostMallocDefault));
}

static SDesc *desc_ptr(kg_server::Slot &sl) { return sl.desc_dev.get<SDesc>(); }

// Workgroups of each persistent phase launch: four a CU (16 waves).
static unsigned srv_grid(kg_server *S) {
  return (unsigned)((S->ctx->num_cus > 0 ? S->ctx->num_cus : 256) * 4);
}

static void launch_phases(kg_server *S, kg_server::Slot &sl) {
  const unsigned grid = srv_grid(S);
  SDesc *D = desc_ptr(sl);
  k_srv_init<<<grid, kSrvBlock, 0, sl.s>>>(D);
  k_srv_build<<<grid, kSrvBlock, 0, sl.s>>>(D);
  k_srv_probe<<<grid, kSrvBlock, 0, sl.s>>>(D);
  k_srv_output<<<grid, kSrvBlock, 0, sl.s>>>(D);
}

static void capture_graph(kg_server *S, kg_server::Slot &sl) {
  hipGraph_t g = nullptr;
  KG_HIP(hipStreamBeginCapture(sl.s, hipStreamCaptureModeThreadLocal));
  launch_phases(S, sl);
  KG_HIP(hipStreamEndCapture(sl.s, &g));
  KG_HIP(hipGraphInstantiate(&sl.graph, g, nullptr, nullptr, 0));
  KG_HIP(hipGraphDestroy(g));
}

// Forms and launches one batch on a free slot from the head of the queue.
static bool launch_batch(kg_server *S, kg_server::Slot &sl) {
  if (S->queue.empty()) return false;
  StreamScope kg_scope(sl.s);
  sl.reqs.clear();
  sl.sreqs.clear();
  uint64_t used = 0;
  int64_t rows = 0;
  while (!S->queue.empty() && (int32_t)sl.reqs.size() < S->max_batch) {
    auto &p = S->queue.front();
    if (used + p.bytes > S->arena_bytes) {
      if (sl.reqs.empty()) fail(KG_ERR_MEMORY, "a request larger than the server's arena");
      break;
    }
    // a batch completes as a whole: bound its probe rows so that a small
    // request does not wait long for the largest one beside it
    if (!sl.reqs.empty() && rows + p.r.probe_rows > kBatchProbeRows) break;
    rows += p.r.probe_rows;
    used += p.bytes;
    sl.reqs.push_back(p);
    S->queue.pop_front();
  }
  // the result columns; refused for memory, the requests go back to the
  // front of the queue
  try {
    sl.outs.clear();
    sl.outs.resize(3 * sl.reqs.size());
    for (size_t q = 0; q < sl.reqs.size(); q++)
      for (int c = 0; c < 3; c++) sl.outs[3 * q + c].reset(S->ctx, (sl.reqs[q].cap + 1) * 8 + 8);
  } catch (...) {
    for (size_t q = sl.reqs.size(); q-- > 0;) S->queue.push_front(sl.reqs[q]);
    sl.reqs.clear();
    sl.outs.clear();
    throw;
  }
  // host descriptor
  uint8_t *h = sl.desc_host;
  SDesc *HD = reinterpret_cast<SDesc *>(h);
  SReq *HR = reinterpret_cast<SReq *>(h + sizeof(SDesc));
  STile *HT = reinterpret_cast<STile *>(h + sizeof(SDesc) + sizeof(SReq) * S->max_batch);
  unsigned long long *HG = reinterpret_cast<unsigned long long *>(
      h + sizeof(SDesc) + sizeof(SReq) * S->max_batch + 4 * sizeof(STile) * sl.max_tiles);
  uint8_t *dbase = sl.desc_dev.get<uint8_t>();
  *HD = SDesc{};
  HD->nreq = (int32_t)sl.reqs.size();
  HD->reqs = reinterpret_cast<SReq *>(dbase + sizeof(SDesc));
  for (int ph = 0; ph < 4; ph++)
    HD->tiles[ph] = reinterpret_cast<STile *>(dbase + sizeof(SDesc) + sizeof(SReq) * S->max_batch +
                                              sizeof(STile) * sl.max_tiles * ph);
  HD->groups = reinterpret_cast<unsigned long long *>(
      dbase + sizeof(SDesc) + sizeof(SReq) * S->max_batch + 4 * sizeof(STile) * sl.max_tiles);
  uint8_t *arena = sl.arena.get<uint8_t>();
  uint64_t at = 0;
  // per phase, tiles of about total / (2 x grid) rows (or slots)
  int64_t tot[4] = {0, 0, 0, 0};
  for (const auto &p : sl.reqs) {
    tot[0] += (int64_t)p.cap + 1;
    tot[1] += p.r.build_keys ? p.r.build_rows : 0;
    tot[2] += p.r.probe_rows;
    tot[3] += (int64_t)p.cap + 1;
  }
  int64_t tile[4];
  for (int ph = 0; ph < 4; ph++) {
    const int64_t want = ceil_div(ceil_div(tot[ph], 2 * (int64_t)srv_grid(S)), 256) * 256;
    tile[ph] = std::min(kSrvTile, std::max(kSrvMinTile, want));
  }
  int ntiles[4] = {0, 0, 0, 0};
  auto add_tiles = [&](int ph, int q, int64_t n) {
    for (int64_t lo = 0; lo < n; lo += tile[ph]) {
      if ((uint64_t)ntiles[ph] >= sl.max_tiles) fail(KG_ERR_UNSUPPORTED, "too many tiles in a batch");
      STile &t = HT[sl.max_tiles * ph + ntiles[ph]++];
      t.req = q;
      t.lo = lo;
      t.hi = lo + tile[ph] < n ? lo + tile[ph] : n;
    }
  };
  for (int q = 0; q < (int)sl.reqs.size(); q++) {
    const kg_serve_request &r = sl.reqs[q].r;
    SReq s{};
    bool own = r.build_keys != nullptr;
    s.cap = sl.reqs[q].cap;
    s.log2cap = slog2(s.cap);
    s.own = own;
    if (own) {
      s.slots = reinterpret_cast<SSlot *>(arena + at);
      at += (s.cap + 1) * 16;
    } else {
      s.slots = S->tables[r.table].slots.get<SSlot>();
    }
    s.acc = reinterpret_cast<uint64_t *>(arena + at);
    at += (s.cap + 1) * 16;
    s.out_key = sl.outs[3 * q].get<int64_t>();
    s.out_sum = sl.outs[3 * q + 1].get<int64_t>();
    s.out_cnt = sl.outs[3 * q + 2].get<int64_t>();
    at = (at + 255) / 256 * 256;
    s.bkeys = r.build_keys;
    s.nb = own ? r.build_rows : 0;
    s.pkeys = r.probe_keys;
    s.pvalid = r.probe_valid;
    s.ppay = r.probe_pay;
    s.np = r.probe_rows;
    HR[q] = s;
    sl.sreqs.push_back(s);
    add_tiles(0, q, (int64_t)s.cap + 1);
    if (own) add_tiles(1, q, s.nb);
    add_tiles(2, q, s.np);
    add_tiles(3, q, (int64_t)s.cap + 1);
    HG[q] = 0;
  }
  for (int ph = 0; ph < 4; ph++) HD->ntiles[ph] = ntiles[ph];
  // upload the descriptor (every tile list up to its used length)
  auto up = [&](uint64_t off, uint64_t bytes) {
    if (bytes) KG_HIP(hipMemcpyAsync(dbase + off, h + off, bytes, hipMemcpyHostToDevice, sl.s));
  };
  up(0, sizeof(SDesc) + sizeof(SReq) * sl.reqs.size());
  for (int ph = 0; ph < 4; ph++)
    up(sizeof(SDesc) + sizeof(SReq) * S->max_batch + sizeof(STile) * sl.max_tiles * ph,
       sizeof(STile) * ntiles[ph]);
  up(sizeof(SDesc) + sizeof(SReq) * S->max_batch + 4 * sizeof(STile) * sl.max_tiles,
     8 * sl.reqs.size());
  if (S->use_graph) {
    if (!sl.graph) capture_graph(S, sl);
    KG_HIP(hipGraphLaunch(sl.graph, sl.s));
  } else {
    launch_phases(S, sl);
    KG_HIP(hipGetLastError());
  }
  KG_HIP(hipEventRecord(sl.done, sl.s));
  sl.busy = true;
  S->inflight += (int64_t)sl.reqs.size();
  S->launches++;
  return true;
}

static void retire(kg_server *S, kg_server::Slot &sl, int64_t *done_ids, int32_t max,
                   int32_t *ndone) {
  StreamScope kg_scope(sl.s);
  // the group counts, into the pinned mirror of the descriptor (the batch
  // that used it has finished); each request's result columns were written
  // in place by the output phase
  const size_t nq = sl.reqs.size();
  const size_t goff = sizeof(SDesc) + sizeof(SReq) * S->max_batch + 4 * sizeof(STile) * sl.max_tiles;
  unsigned long long *groups = reinterpret_cast<unsigned long long *>(sl.desc_host + goff);
  KG_HIP(hipMemcpyAsync(groups, sl.desc_dev.get<uint8_t>() + goff, 8 * nq, hipMemcpyDeviceToHost,
                        sl.s));
  KG_HIP(hipStreamSynchronize(sl.s));
  for (size_t q = 0; q < nq; q++) {
    auto b = std::make_unique<kg_batch>();
    b->ctx = S->ctx;
    b->nrows = (int64_t)groups[q];
    b->cols.resize(3);
    for (int c = 0; c < 3; c++) {
      b->cols[c].type = KG_T_INT64;
      b->cols[c].values = std::move(sl.outs[3 * q + c]);
    }
    S->results[sl.reqs[q].r.id] = b.release();
    report(S, sl.reqs[q].r.id, done_ids, max, ndone);
  }
  sl.outs.clear();
  S->inflight -= (int64_t)sl.reqs.size();
  sl.busy = false;
  sl.reqs.clear();
}

}  // namespace kg

extern "C" {

int kg_server_create(kg_ctx *ctx, uint64_t arena_bytes, int32_t max_batch, int32_t use_graph,
                     kg_server **out) {
  return guarded([&] {
    *out = nullptr;
    if (max_batch < 1 || max_batch > 4096) fail(KG_ERR_UNSUPPORTED, "1 to 4096 requests per batch");
    auto S = std::make_unique<kg_server>();
    S->ctx = ctx;
    S->max_batch = max_batch;
    S->use_graph = use_graph != 0;
    S->arena_bytes = arena_bytes / kSrvSlots;
    S->nworkers = kLargeWorkers;
    if (const char *w = getenv("KG_SERVE_LARGE_WORKERS")) {
      const int v = atoi(w);
      if (v < 1 || v > 4) fail(KG_ERR_UNSUPPORTED, "KG_SERVE_LARGE_WORKERS must be 1 to 4");
      S->nworkers = v;
    }
    for (auto &sl : S->slots) {
      sl.s = make_stream(ctx->device);
      sl.arena.reset(ctx, S->arena_bytes);
      layout_desc(S.get(), sl);
      KG_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    *out = S.release();
  });
}

int kg_server_add_table(kg_server *S, const int64_t *keys, int64_t n, int32_t *table_id) {
  return guarded([&] {
    if (n < 0) fail(KG_ERR_INTERNAL, "negative row count");
    kg_server::Resident t;
    uint64_t want = (uint64_t)(n < 256 ? 256 : n);
    t.cap = 1ull << slog2(want * 2);
    t.log2cap = slog2(t.cap);
    t.slots.reset(S->ctx, (t.cap + 1) * 16);
    DevBuf scratch(S->ctx, (t.cap + 1) * 16);
    // a one-request descriptor of its own (the stream slots' descriptors
    // may still hold a batch's group counts): init + build phases
    std::vector<STile> tiles[2];
    for (int64_t lo = 0; lo < (int64_t)t.cap + 1; lo += kSrvTile)
      tiles[0].push_back(STile{0, 0, lo, lo + kSrvTile < (int64_t)t.cap + 1 ? lo + kSrvTile
                                                                          : (int64_t)t.cap + 1});
    for (int64_t lo = 0; lo < n; lo += kSrvTile)
      tiles[1].push_back(STile{0, 0, lo, lo + kSrvTile < n ? lo + kSrvTile : n});
    uint64_t bytes = sizeof(SDesc) + sizeof(SReq) +
                     sizeof(STile) * (tiles[0].size() + tiles[1].size());
    std::vector<uint8_t> h(bytes);
    DevBuf d(S->ctx, bytes);
    uint8_t *dbase = d.get<uint8_t>();
    SDesc D{};
    D.nreq = 1;
    D.reqs = reinterpret_cast<SReq *>(dbase + sizeof(SDesc));
    D.tiles[0] = reinterpret_cast<STile *>(dbase + sizeof(SDesc) + sizeof(SReq));
    D.tiles[1] = D.tiles[0] + tiles[0].size();
    D.tiles[2] = D.tiles[3] = D.tiles[1];
    D.ntiles[0] = (int32_t)tiles[0].size();
    D.ntiles[1] = (int32_t)tiles[1].size();
    SReq r{};
    r.bkeys = keys;
    r.nb = n;
    r.slots = t.slots.get<SSlot>();
    r.cap = t.cap;
    r.log2cap = t.log2cap;
    r.own = 1;
    r.acc = scratch.get<uint64_t>();
    memcpy(h.data(), &D, sizeof(SDesc));
    memcpy(h.data() + sizeof(SDesc), &r, sizeof(SReq));
    memcpy(h.data() + sizeof(SDesc) + sizeof(SReq), tiles[0].data(), sizeof(STile) * tiles[0].size());
    if (!tiles[1].empty())
      memcpy(h.data() + sizeof(SDesc) + sizeof(SReq) + sizeof(STile) * tiles[0].size(),
             tiles[1].data(), sizeof(STile) * tiles[1].size());
    hipStream_t st = S->slots[0].s;
    KG_HIP(hipMemcpyAsync(dbase, h.data(), bytes, hipMemcpyHostToDevice, st));
    const unsigned grid = srv_grid(S);
    k_srv_init<<<grid, kSrvBlock, 0, st>>>(d.get<SDesc>());
    k_srv_build<<<grid, kSrvBlock, 0, st>>>(d.get<SDesc>());
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(st));
    t.join = build_join(S->ctx, keys, n);
    *table_id = (int32_t)S->tables.size();
    S->tables.push_back(std::move(t));
  });
}

int kg_server_submit(kg_server *S, const kg_serve_request *r) {
  return guarded([&] {
    if (!r->probe_keys || !r->probe_pay) fail(KG_ERR_INTERNAL, "request without probe columns");
    if (r->build_keys == nullptr && (r->table < 0 || r->table >= (int32_t)S->tables.size()))
      fail(KG_ERR_INTERNAL, "request names no resident table");
    if (r->probe_rows >= kLargeProbe) {
      kg_server::Large L{*r, r->build_keys ? nullptr : S->tables[r->table].join};
      std::lock_guard<std::mutex> g(S->large_mu);
      while (S->large_workers.size() < (size_t)S->nworkers)
        S->large_workers.emplace_back(large_loop, S);
      S->large.push_back(L);
      S->large_outstanding++;
      S->large_cv.notify_all();
      return;
    }
    kg_server::Pending p;
    p.r = *r;
    p.cap = req_cap(S, *r);
    p.bytes = req_bytes(p.cap, r->build_keys != nullptr);
    if (p.bytes > S->arena_bytes)
      fail(KG_ERR_MEMORY, "request larger than the server's arena share");
    S->queue.push_back(p);
  });
}

int kg_server_progress(kg_server *S, int32_t wait, int64_t *done_ids, int32_t max,
                       int32_t *ndone) {
  return guarded([&] {
    *ndone = 0;
    while (!S->unreported.empty() && *ndone < max) {
      done_ids[(*ndone)++] = S->unreported.front();
      S->unreported.pop_front();
    }
    bool waited = false;
    for (int pass = 0; pass < 2; pass++) {
      for (auto &sl : S->slots) {
        if (!sl.busy) continue;
        hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) {
          if (!(wait && !waited)) continue;
          KG_HIP(hipEventSynchronize(sl.done));
          waited = true;
        } else if (q != hipSuccess) {
          KG_HIP(q);
        }
        retire(S, sl, done_ids, max, ndone);
      }
      for (auto &sl : S->slots)
        if (!sl.busy && !S->queue.empty()) launch_batch(S, sl);
      collect_large(S, false, done_ids, max, ndone);
      if (!wait || waited || *ndone > 0) break;
      if (S->inflight == 0) {
        // nothing batched in flight: wait for a large request, if any
        if (S->large_outstanding > 0 && *ndone < max) collect_large(S, true, done_ids, max, ndone);
        break;
      }
    }
  });
}

int kg_server_take_result(kg_server *S, int64_t id, kg_batch **out) {
  return guarded([&] {
    auto it = S->results.find(id);
    if (it == S->results.end()) fail(KG_ERR_INTERNAL, "no finished request with that id");
    *out = it->second;
    S->results.erase(it);
  });
}

int64_t kg_server_pending(const kg_server *S) {
  return S ? (int64_t)S->queue.size() + S->inflight + S->large_outstanding +
                 (int64_t)S->unreported.size()
           : 0;
}

void kg_server_destroy(kg_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->large_mu);
    S->large_stop = true;
    S->large.clear();
    S->large_cv.notify_all();
  }
  for (auto &t : S->large_workers) t.join();
  for (auto &d : S->large_done) kg_batch_release(d.b);
  for (auto &sl : S->slots) {
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    if (sl.graph) (void)hipGraphExecDestroy(sl.graph);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.desc_host) (void)hipHostFree(sl.desc_host);
  }
  for (auto &kv : S->results) kg_batch_release(kv.second);
  for (auto &t : S->tables) kg_join_destroy(t.join);
  hipStream_t ss[kSrvSlots];
  for (int i = 0; i < kSrvSlots; i++) ss[i] = S->slots[i].s;
  delete S;
  for (auto s : ss) destroy_stream(s);
}

}  // extern "C"
