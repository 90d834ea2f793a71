// kg_grace.hip -- the grace hash join (join/grace.rs:99-1031) fused with the
// per-partition aggregate of the plan in include/kgpu.h (kg_grace_*).
//
//  sink     H_kg -> partition_of -> stable radix scatter (kg_repartition);
//           each sunk batch stays partitioned in HBM, or, once the context's
//           HBM budget would be passed, in pinned host memory (the spill).
//  finish   per partition: the build rows aggregated per key (the device
//           hash aggregation: row count, non-NULL payload count, payload
//           sum), a read-only lookup table of those aggregates, then the
//           probe rows streamed against it.  A probe row with key k adds
//           nb_k pairs and pay * Sb_k to its partition (wave-reduced, one
//           atomic per wave), the factorized form of the pair walk
//           (join/probe.rs:266-337) for a SUM of payload products.
#include <hip/hip_runtime.h>

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

struct GSlot {
  uint64_t key;
  uint64_t nb;   // build rows with the key
  uint64_t nbv;  // of which with a non-NULL payload
  uint64_t sb;   // their payload sum (wrapping)
};

constexpr uint64_t kGEmpty = 0x8000000000000000ULL;

__device__ __forceinline__ uint64_t gslot_start(uint64_t key, int log2cap) {
  return dev::table_hash(key, 0x51ED27A1D3C4B9E5ULL) >> (64 - log2cap);
}

// Lookup table of the build partition's per-key aggregates (keys are
// distinct: no CAS race on a key).  Slot cap holds the key equal to EMPTY.
__global__ void k_grace_table(const int64_t *keys, const uint8_t *key_valid, const int64_t *nb,
                              const int64_t *nbv, const int64_t *sb, int64_t n, GSlot *slots,
                              uint64_t mask, int log2cap, uint64_t cap) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (key_valid && !((key_valid[i >> 3] >> (i & 7)) & 1)) continue;  // the NULL key group
    uint64_t k = (uint64_t)keys[i];
    uint64_t s;
    if (k == kGEmpty) {
      s = cap;
    } else {
      s = gslot_start(k, log2cap);
      while (true) {
        unsigned long long prev = atomicCAS((unsigned long long *)&slots[s].key, kGEmpty, k);
        if (prev == kGEmpty) break;
        s = (s + 1) & mask;
      }
    }
    slots[s].key = k;
    slots[s].nb = (uint64_t)nb[i];
    slots[s].nbv = (uint64_t)nbv[i];
    slots[s].sb = (uint64_t)sb[i];
  }
}

__global__ void k_grace_init(GSlot *slots, uint64_t n) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n;
       s += (uint64_t)gridDim.x * blockDim.x)
    slots[s] = GSlot{kGEmpty, 0, 0, 0};
}

// Probe rows against the table; out[0..2] of the partition: pairs, pairs
// with both payloads non-NULL, payload-product sum (all wrapping u64).
__global__ void k_grace_probe(DCol key, DCol pay, int64_t n, const GSlot *slots, uint64_t mask,
                              int log2cap, uint64_t cap, unsigned long long *out) {
  uint64_t cnt = 0, vcnt = 0, sum = 0;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x; b < n; b += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = b + threadIdx.x;
    if (i < n && dev::valid(key, i)) {
      uint64_t k = (uint64_t)dev::load_i64(key, i);
      const GSlot *hit = nullptr;
      if (k == kGEmpty) {
        if (slots[cap].nb) hit = &slots[cap];
      } else {
        uint64_t s = gslot_start(k, log2cap);
        while (true) {
          uint64_t sk = slots[s].key;
          if (sk == k) {
            hit = &slots[s];
            break;
          }
          if (sk == kGEmpty) break;
          s = (s + 1) & mask;
        }
      }
      if (hit) {
        cnt += hit->nb;
        if (dev::valid(pay, i)) {
          vcnt += hit->nbv;
          sum += (uint64_t)dev::load_i64(pay, i) * hit->sb;
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off);
    vcnt += __shfl_xor(vcnt, off);
    sum += __shfl_xor(sum, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (cnt) atomicAdd(out, (unsigned long long)cnt);
    if (vcnt) atomicAdd(out + 1, (unsigned long long)vcnt);
    if (sum) atomicAdd(out + 2, (unsigned long long)sum);
  }
}

// Range of the non-NULL payloads: [min, max] via atomics.
__global__ void k_minmax_i64(DCol c, int64_t n, long long *mm) {
  long long lo = INT64_MAX, hi = INT64_MIN;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!dev::valid(c, i)) continue;
    long long v = (long long)dev::load_i64(c, i);
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int off = 32; off > 0; off >>= 1) {
    long long a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(mm, lo);
    atomicMax(mm + 1, hi);
  }
}

static void ck(int rc) {
  if (rc != KG_OK) throw Error{rc, std::string(kg_last_error())};
}

static int glog2(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

}  // namespace kg

using namespace kg;

struct kg_grace {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  uint32_t parts = 64;
  uint64_t table_budget = 0;
  // one partitioned piece per sunk batch: device (a kg_batch) or host copy
  struct Piece {
    kg_batch *dev = nullptr;
    // the spilled copy, in pinned host memory (hipHostMalloc)
    uint8_t *hk = nullptr, *hv = nullptr, *hp = nullptr, *hpv = nullptr;
    bool key_nullable = false, pay_nullable = false;
    int64_t rows = 0;
    std::vector<int64_t> off;
  };
  std::vector<Piece> build, probe;
  DevBuf mm;  // [build min, build max, probe min, probe max]
  uint64_t spilled = 0;
};

namespace kg {

static void free_piece(kg_grace::Piece &p) {
  if (p.dev) kg_batch_release(p.dev);
  p.dev = nullptr;
  for (uint8_t *q : {p.hk, p.hv, p.hp, p.hpv})
    if (q) (void)hipHostFree(q);
  p.hk = p.hv = p.hp = p.hpv = nullptr;
}

// Moves a partitioned piece from HBM to pinned host memory.
static void spill_piece(kg_grace *g, kg_grace::Piece &piece) {
  if (!piece.dev) return;
  kg_column dk{}, dp{};
  ck(kg_batch_column(piece.dev, 0, &dk));
  ck(kg_batch_column(piece.dev, 1, &dp));
  int64_t n = piece.rows;
  size_t vb = (size_t)ceil_div(n, 32) * 4;
  KG_HIP(hipHostMalloc((void **)&piece.hk, (size_t)n * 8, hipHostMallocDefault));
  KG_HIP(hipHostMalloc((void **)&piece.hp, (size_t)n * 8, hipHostMallocDefault));
  KG_HIP(hipMemcpyAsync(piece.hk, dk.values, (size_t)n * 8, hipMemcpyDeviceToHost, g->s));
  KG_HIP(hipMemcpyAsync(piece.hp, dp.values, (size_t)n * 8, hipMemcpyDeviceToHost, g->s));
  if (dk.validity) {
    KG_HIP(hipHostMalloc((void **)&piece.hv, vb, hipHostMallocDefault));
    KG_HIP(hipMemcpyAsync(piece.hv, dk.validity, vb, hipMemcpyDeviceToHost, g->s));
  }
  if (dp.validity) {
    KG_HIP(hipHostMalloc((void **)&piece.hpv, vb, hipHostMallocDefault));
    KG_HIP(hipMemcpyAsync(piece.hpv, dp.validity, vb, hipMemcpyDeviceToHost, g->s));
  }
  KG_HIP(hipStreamSynchronize(g->s));
  kg_batch_release(piece.dev);
  piece.dev = nullptr;
  g->spilled += (uint64_t)n * 16 + (dk.validity ? vb : 0) + (dp.validity ? vb : 0);
}

// HBM a sub-batch of `rows` needs while it is partitioned: the staged key and
// payload, their partitioned copy, partition ids and the block histograms.
static uint64_t sink_bytes(int64_t rows) { return (uint64_t)rows * 40 + (4u << 20); }

static bool fits(kg_grace *g, uint64_t need) {
  return !g->ctx->budget || g->ctx->in_use.load() + need + g->table_budget <= g->ctx->budget;
}

// Past the HBM budget (with one partition's tables kept free), partitioned
// pieces move to pinned host memory, oldest first; a batch is partitioned in
// sub-batches that fit beside what stays in HBM.
static void sink_rows(kg_grace *g, std::vector<kg_grace::Piece> &store, const kg_column *key,
                      const kg_column *pay, int64_t n, int side) {
  if (n <= 0) return;
  if (!is_int_like(key->type) || type_width(key->type) != 8 || !is_int_like(pay->type) ||
      type_width(pay->type) != 8)
    fail(KG_ERR_UNSUPPORTED, "grace join of 64-bit integer keys and payloads");
  check_column(*key, n, "grace key");
  check_column(*pay, n, "grace payload");
  for (int64_t done = 0; done < n;) {
    int64_t m = n - done;
    if (g->ctx->budget) {
      for (auto *v : {&g->build, &g->probe})
        for (auto &pc : *v)
          if (!fits(g, sink_bytes(m < (1 << 16) ? m : (1 << 16)))) spill_piece(g, pc);
      uint64_t used = g->ctx->in_use.load() + g->table_budget + (4u << 20);
      int64_t room = used < g->ctx->budget ? (int64_t)((g->ctx->budget - used) / 40) : 0;
      if (room < 1024 && room < m)
        fail(KG_ERR_MEMORY, "HBM budget too small to partition a grace batch");
      if (m > room) m = room;
    }
    kg_column k = *key, v = *pay;
    k.offset += done;
    v.offset += done;
    k.length -= done;
    v.length -= done;
    // payload range for the factorization check
    {
      Staged st(g->ctx, g->s);
      DCol pc = st.add(v, m);
      k_minmax_i64<<<stream_grid(m, 256), 256, 0, g->s>>>(pc, m,
                                                          g->mm.get<long long>() + 2 * side);
      KG_HIP(hipGetLastError());
      KG_HIP(hipStreamSynchronize(g->s));
    }
    kg_column cols[2] = {k, v};
    int32_t kc = 0;
    kg_grace::Piece piece;
    piece.off.assign(g->parts + 1, 0);
    ck(kg_repartition(g->ctx, 2, cols, m, 1, &kc, g->parts, piece.off.data(), &piece.dev));
    piece.rows = m;
    kg_column dk{}, dp{};
    ck(kg_batch_column(piece.dev, 0, &dk));
    ck(kg_batch_column(piece.dev, 1, &dp));
    piece.key_nullable = dk.validity != nullptr;
    piece.pay_nullable = dp.validity != nullptr;
    if (!fits(g, 0)) spill_piece(g, piece);
    store.push_back(std::move(piece));
    done += m;
  }
}

// Partition p's rows of a piece as key and payload columns (device or host).
static void slice(const kg_grace::Piece &pc, uint32_t p, kg_column &k, kg_column &v, int64_t &n) {
  int64_t a = pc.off[p], b = pc.off[p + 1];
  n = b - a;
  if (pc.dev) {
    ck(kg_batch_column(pc.dev, 0, &k));
    ck(kg_batch_column(pc.dev, 1, &v));
  } else {
    k = kg_column{KG_T_INT64, 0, 0, 0, pc.rows, 0, pc.hk, pc.hv};
    v = kg_column{KG_T_INT64, 0, 0, 0, pc.rows, 0, pc.hp, pc.hpv};
  }
  k.offset = a;
  v.offset = a;
  k.length = v.length = b;
}

}  // namespace kg

extern "C" {

int kg_grace_create(kg_ctx *ctx, uint32_t partitions, uint64_t table_budget_bytes, kg_grace **out) {
  return guarded([&] {
    *out = nullptr;
    if (partitions < 1 || partitions > 4096) fail(KG_ERR_UNSUPPORTED, "1 to 4096 partitions");
    auto g = std::make_unique<kg_grace>();
    g->ctx = ctx;
    g->parts = partitions;
    g->table_budget = table_budget_bytes ? table_budget_bytes : (8ull << 30);
    g->s = make_stream(ctx->device);
    StreamScope kg_scope(g->s);
    g->mm.reset(ctx, 32);
    long long init[4] = {INT64_MAX, INT64_MIN, INT64_MAX, INT64_MIN};
    KG_HIP(hipMemcpy(g->mm.get(), init, 32, hipMemcpyHostToDevice));
    *out = g.release();
  });
}

int kg_grace_build_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows) {
  return guarded([&] { StreamScope kg_scope(g->s); sink_rows(g, g->build, key, pay, nrows, 0); });
}

int kg_grace_probe_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows) {
  return guarded([&] { StreamScope kg_scope(g->s); sink_rows(g, g->probe, key, pay, nrows, 1); });
}

int kg_grace_finish(kg_grace *g, int64_t *counts, int64_t *sums, uint8_t *sum_valid,
                    uint64_t *spilled_bytes) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    long long mm[4];
    KG_HIP(hipMemcpy(mm, g->mm.get(), 32, hipMemcpyDeviceToHost));
    bool any_b = mm[0] <= mm[1], any_p = mm[2] <= mm[3];
    if (any_b && any_p) {
      // every pair's product in [0, 2^31): the per-pair % 2^31 is the identity
      bool nonneg = mm[0] >= 0 && mm[2] >= 0;
      unsigned __int128 prod = (unsigned __int128)(uint64_t)(mm[1] > 0 ? mm[1] : 0) *
                               (unsigned __int128)(uint64_t)(mm[3] > 0 ? mm[3] : 0);
      if (!nonneg || prod >= ((unsigned __int128)1 << 31))
        fail(KG_ERR_UNSUPPORTED,
             "payload products reach 2^31 or are negative: the per-pair modulus does not factor");
    }
    DevBuf out(g->ctx, (uint64_t)g->parts * 3 * 8);
    KG_HIP(hipMemsetAsync(out.get(), 0, out.bytes(), g->s));
    const kg_column key_t{KG_T_INT64, 0, 0, 0, 0, 0, nullptr, nullptr};
    const kg_column in_t{KG_T_INT64, 0, 0, 0, 0, 0, nullptr, nullptr};
    kg_agg_spec aggs[3] = {{KG_AGG_COUNT_STAR, 0}, {KG_AGG_COUNT, 0}, {KG_AGG_SUM, 0}};
    for (uint32_t p = 0; p < g->parts; p++) {
      int64_t nb = 0, np = 0;
      for (auto &pc : g->build) nb += pc.off[p + 1] - pc.off[p];
      for (auto &pc : g->probe) np += pc.off[p + 1] - pc.off[p];
      if (nb == 0 || np == 0) continue;
      // 1. the build partition aggregated per key
      kg_hash_agg *agg = nullptr;
      ck(kg_agg_create(g->ctx, &key_t, 1, &in_t, 3, aggs, 0, (uint64_t)nb, &agg));
      std::unique_ptr<kg_hash_agg, void (*)(kg_hash_agg *)> agg_guard(agg, kg_agg_destroy);
      for (auto &pc : g->build) {
        kg_column k, v;
        int64_t n;
        slice(pc, p, k, v, n);
        if (n) ck(kg_agg_sink(agg, &k, &v, n));
      }
      ck(kg_agg_finalize(agg));
      kg_batch *res = nullptr;
      ck(kg_agg_result(agg, &res));
      std::unique_ptr<kg_batch, void (*)(kg_batch *)> res_guard(res, kg_batch_release);
      int64_t groups = kg_batch_num_rows(res);
      kg_column rk, rn, rnv, rs;
      ck(kg_batch_column(res, 0, &rk));
      ck(kg_batch_column(res, 1, &rn));
      ck(kg_batch_column(res, 2, &rnv));
      ck(kg_batch_column(res, 3, &rs));
      // 2. the lookup table
      uint64_t cap = 1ull << glog2((uint64_t)(groups < 256 ? 256 : groups) * 2);
      if ((cap + 1) * sizeof(GSlot) > g->table_budget)
        fail(KG_ERR_MEMORY, "a grace partition's table exceeds the table budget");
      DevBuf slots(g->ctx, (cap + 1) * sizeof(GSlot));
      k_grace_init<<<stream_grid(cap + 1, 256), 256, 0, g->s>>>(slots.get<GSlot>(), cap + 1);
      {
        ProfScope prof("grace_build", g->s);
        k_grace_table<<<stream_grid(groups, 256), 256, 0, g->s>>>(
            static_cast<const int64_t *>(rk.values), rk.validity,
            static_cast<const int64_t *>(rn.values), static_cast<const int64_t *>(rnv.values),
            static_cast<const int64_t *>(rs.values), groups, slots.get<GSlot>(), cap - 1,
            glog2(cap), cap);
        KG_HIP(hipGetLastError());
      }
      // 3. the probe partition streamed against it
      for (auto &pc : g->probe) {
        kg_column k, v;
        int64_t n;
        slice(pc, p, k, v, n);
        if (!n) continue;
        // rows [offset, offset + n): device slices in place, host slices
        // copied to HBM
        Staged st(g->ctx, g->s);
        DCol dk = st.add(k, n);
        DCol dv = st.add(v, n);
        ProfScope prof("grace_probe", g->s);
        k_grace_probe<<<stream_grid(n, 256), 256, 0, g->s>>>(
            dk, dv, n, slots.get<GSlot>(), cap - 1, glog2(cap), cap,
            out.get<unsigned long long>() + 3 * p);
        KG_HIP(hipGetLastError());
        if (st.staged_host()) KG_HIP(hipStreamSynchronize(g->s));
      }
      KG_HIP(hipStreamSynchronize(g->s));
    }
    std::vector<uint64_t> h((size_t)g->parts * 3);
    KG_HIP(hipMemcpy(h.data(), out.get(), h.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t p = 0; p < g->parts; p++) {
      counts[p] = (int64_t)h[3 * p];
      sums[p] = (int64_t)h[3 * p + 2];
      if (sum_valid) sum_valid[p] = h[3 * p + 1] != 0;
    }
    if (spilled_bytes) *spilled_bytes = g->spilled;
  });
}

void kg_grace_destroy(kg_grace *g) {
  if (!g) return;
  if (g->s) (void)hipStreamSynchronize(g->s);
  for (auto &p : g->build) free_piece(p);
  for (auto &p : g->probe) free_piece(p);
  hipStream_t s = g->s;
  delete g;
  destroy_stream(s);
}

}  // extern "C"
