// kg_frame.hip -- exchange frames written in HBM: after ExchangeSink's
// split, each reduce partition's rows leave as one frame
// (crates/krishiv-shuffle/src/frame.rs:1-26): a 44-byte little-endian
// header (magic KXFR, version 1, kind, compression 0, exchange id, stage
// attempt, map partition, task attempt, sequence, reduce partition, body
// length, CRC32C of the body) and an Arrow IPC stream of one batch as the
// body (frame.rs:205-221: the schema message, the record batch message,
// the end-of-stream marker).
//
// The host builds the IPC metadata (small flatbuffers, one record batch
// message a partition, from the per-partition null counts the device
// counts) and a list of copy items; the device copies each partition's
// slice of every column into its body (validity and boolean bitmaps
// shifted to bit 0, every buffer zero-padded to 8 bytes), then computes the
// CRC32C of every body in 64 KiB chunks (slicing-by-8 tables in LDS, 256
// lanes of 256 bytes a chunk, combined by multiplying CRCs by powers of x
// modulo the polynomial) and writes it into the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

constexpr uint32_t kCrcPoly = 0x82F63B78u;  // Castagnoli, reflected
constexpr int kCrcBlock = 256;
constexpr int64_t kCrcSeg = 256;                    // bytes a lane
constexpr int64_t kCrcChunk = kCrcSeg * kCrcBlock;  // bytes a workgroup
constexpr int64_t kCopyChunk = 64 * 1024;           // bytes a copy item at most

// a * b modulo the CRC polynomial, both in the reflected representation
// (bit 31 is x^0): the product of two remainders.
__host__ __device__ inline uint32_t crc_mulmod(uint32_t a, uint32_t b) {
  uint32_t p = 0;
  for (int k = 0; k < 32; k++) {
    if (a & (0x80000000u >> k)) p ^= b;
    b = (b & 1) ? (b >> 1) ^ kCrcPoly : b >> 1;
  }
  return p;
}

// x^(8 n) modulo the polynomial: what n zero bytes do to a CRC register.
__host__ __device__ inline uint32_t crc_x8n(uint64_t n) {
  uint32_t r = 0x80000000u;     // x^0
  uint32_t sq = 0x00800000u;    // x^8
  while (n) {
    if (n & 1) r = crc_mulmod(r, sq);
    sq = crc_mulmod(sq, sq);
    n >>= 1;
  }
  return r;
}

// The register after feeding `bytes` from register value c (no init, no
// final xor): the raw CRC.  words: the segment as little-endian 32-bit words.
__device__ __forceinline__ uint32_t crc_words(uint32_t c, const uint32_t *w, int nw,
                                              const uint32_t (*T)[256]) {
  int k = 0;
  for (; k + 1 < nw; k += 2) {
    const uint32_t a = w[k] ^ c, b = w[k + 1];
    c = T[7][a & 0xFF] ^ T[6][(a >> 8) & 0xFF] ^ T[5][(a >> 16) & 0xFF] ^ T[4][a >> 24] ^
        T[3][b & 0xFF] ^ T[2][(b >> 8) & 0xFF] ^ T[1][(b >> 16) & 0xFF] ^ T[0][b >> 24];
  }
  if (k < nw) {
    const uint32_t a = w[k] ^ c;
    c = T[3][a & 0xFF] ^ T[2][(a >> 8) & 0xFF] ^ T[1][(a >> 16) & 0xFF] ^ T[0][a >> 24];
  }
  return c;
}

__device__ __forceinline__ uint32_t crc_bytes(uint32_t c, const uint8_t *p, int n,
                                              const uint32_t (*T)[256]) {
  for (int i = 0; i < n; i++) c = T[0][(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}

struct CrcChunk {
  uint64_t at;   // first byte (in the output)
  uint32_t len;  // bytes (<= kCrcChunk)
  uint32_t frame;
};

// Raw CRC of each chunk: a lane's 256-byte segment, then the workgroup's
// segments combined in order (a tree: each level shifts the left CRC over
// the right part's bytes).
__global__ void __launch_bounds__(kCrcBlock) k_crc_chunks(const uint8_t *data,
                                                          const CrcChunk *chunks,
                                                          uint32_t *chunk_crc) {
  __shared__ uint32_t T[8][256];
  __shared__ uint32_t s_crc[kCrcBlock];
  __shared__ uint32_t s_len[kCrcBlock];
  {
    uint32_t c = threadIdx.x;
    for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ kCrcPoly : c >> 1;
    T[0][threadIdx.x] = c;
  }
  __syncthreads();
  for (int t = 1; t < 8; t++) {
    const uint32_t prev = T[t - 1][threadIdx.x];
    T[t][threadIdx.x] = (prev >> 8) ^ T[0][prev & 0xFF];
    __syncthreads();
  }
  const CrcChunk ch = chunks[blockIdx.x];
  const int64_t lo = (int64_t)threadIdx.x * kCrcSeg;
  const int64_t rem = (int64_t)ch.len - lo;
  const int len = rem <= 0 ? 0 : (int)(rem < kCrcSeg ? rem : kCrcSeg);
  uint32_t c = 0;
  if (len > 0) {
    const uint8_t *p = data + ch.at + lo;
    // to a 4-byte boundary, then words, then the tail
    int head = (int)((4 - ((uintptr_t)p & 3)) & 3);
    if (head > len) head = len;
    c = crc_bytes(c, p, head, T);
    const int nw = (len - head) / 4;
    c = crc_words(c, reinterpret_cast<const uint32_t *>(p + head), nw, T);
    c = crc_bytes(c, p + head + 4 * nw, len - head - 4 * nw, T);
  }
  s_crc[threadIdx.x] = c;
  s_len[threadIdx.x] = (uint32_t)len;
  __syncthreads();
  for (int stride = 1; stride < kCrcBlock; stride <<= 1) {
    uint32_t nc = 0, nl = 0;
    const bool left = (threadIdx.x % (2 * stride)) == 0;
    if (left) {
      const uint32_t rl = s_len[threadIdx.x + stride];
      nc = rl ? crc_mulmod(s_crc[threadIdx.x], crc_x8n(rl)) ^ s_crc[threadIdx.x + stride]
              : s_crc[threadIdx.x];
      nl = s_len[threadIdx.x] + rl;
    }
    __syncthreads();
    if (left) {
      s_crc[threadIdx.x] = nc;
      s_len[threadIdx.x] = nl;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) chunk_crc[blockIdx.x] = s_crc[0];
}

struct CrcFrame {
  uint64_t header_at;  // the frame's first byte
  uint64_t body_len;
  uint32_t first, count;  // its chunks
};

// Each frame's chunks combined, the initial and final all-ones applied
// (crc32c(M) = ~(raw(M) ^ x^(8 |M|) * 0xFFFFFFFF)), written at header
// bytes 40..44; crc_out (optional) gets it too.
__global__ void k_crc_frames(const CrcFrame *frames, int64_t nframes, const CrcChunk *chunks,
                             const uint32_t *chunk_crc, uint8_t *out, uint32_t *crc_out) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nframes;
       f += (int64_t)gridDim.x * blockDim.x) {
    const CrcFrame F = frames[f];
    uint32_t raw = 0;
    const uint32_t full = crc_x8n(kCrcChunk);
    for (uint32_t k = 0; k < F.count; k++) {
      const CrcChunk &ch = chunks[F.first + k];
      raw = crc_mulmod(raw, ch.len == kCrcChunk ? full : crc_x8n(ch.len)) ^ chunk_crc[F.first + k];
    }
    const uint32_t crc = ~(raw ^ crc_mulmod(0xFFFFFFFFu, crc_x8n(F.body_len)));
    if (out) {
      uint8_t *h = out + F.header_at + 40;
      h[0] = (uint8_t)crc;
      h[1] = (uint8_t)(crc >> 8);
      h[2] = (uint8_t)(crc >> 16);
      h[3] = (uint8_t)(crc >> 24);
    }
    if (crc_out) crc_out[f] = crc;
  }
}

// ---------------------------------------------------------------- copies
struct CopyItem {
  const uint8_t *src;
  uint64_t dst;      // output byte
  uint64_t src_bit;  // bit copies: the first source bit
  uint32_t nbytes;   // bytes written (data, then zeros)
  uint32_t ndata;    // byte copies: source bytes; bit copies: source bits
  uint32_t bits;     // 1: a bitmap shifted to bit 0
  uint32_t src_bytes;  // bit copies: readable source bytes from src_bit / 8
};

// One workgroup an item: bytes (words where both ends allow) or a shifted
// bitmap; the padding after the data is written as zeros.
__global__ void __launch_bounds__(256) k_frame_copy(const CopyItem *items, uint8_t *out) {
  const CopyItem it = items[blockIdx.x];
  uint8_t *d = out + it.dst;
  if (it.bits) {
    const uint8_t *s = it.src + (it.src_bit >> 3);
    const int sh = (int)(it.src_bit & 7);
    const uint32_t nb = (it.ndata + 7) / 8;
    for (uint32_t j = threadIdx.x; j < it.nbytes; j += blockDim.x) {
      uint8_t v = 0;
      if (j < nb) {
        uint32_t w = s[j];
        if (sh && j + 1 < it.src_bytes) w |= (uint32_t)s[j + 1] << 8;
        v = (uint8_t)(w >> sh);
        if (j == nb - 1 && (it.ndata & 7)) v &= (uint8_t)((1u << (it.ndata & 7)) - 1);
      }
      d[j] = v;
    }
    return;
  }
  if ((((uintptr_t)it.src | (uintptr_t)d) & 3) == 0) {
    const uint32_t nw = it.ndata / 4;
    const uint32_t *s4 = reinterpret_cast<const uint32_t *>(it.src);
    uint32_t *d4 = reinterpret_cast<uint32_t *>(d);
    for (uint32_t j = threadIdx.x; j < nw; j += blockDim.x) d4[j] = s4[j];
    for (uint32_t j = 4 * nw + threadIdx.x; j < it.nbytes; j += blockDim.x)
      d[j] = j < it.ndata ? it.src[j] : 0;
    return;
  }
  for (uint32_t j = threadIdx.x; j < it.nbytes; j += blockDim.x)
    d[j] = j < it.ndata ? it.src[j] : 0;
}

// Null counts of each (partition, column) of columns with validity: rows
// of partition p lie in [off[p], off[p + 1]).
__global__ void k_frame_nulls(const uint8_t *valid, int64_t voff, const int64_t *off,
                              uint32_t parts, int64_t n, int col, int ncols,
                              unsigned long long *nulls) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i + voff;
    if ((valid[j >> 3] >> (j & 7)) & 1) continue;
    uint32_t lo = 0, hi = parts;  // the partition whose range holds i
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) / 2;
      if (off[m] <= i) lo = m;
      else hi = m;
    }
    atomicAdd(nulls + (uint64_t)lo * ncols + col, 1ULL);
  }
}

// ---------------------------------------------------------------- IPC metadata
// A flatbuffer written front to back: every uoffset points forward, each
// table's vtable sits just before it (soffset positive); alignment is
// relative to the buffer's start.
class Fb {
 public:
  std::vector<uint8_t> b;
  void pad(size_t a) {
    while (b.size() % a) b.push_back(0);
  }
  template <class T>
  size_t put(T v) {
    const size_t at = b.size();
    b.resize(at + sizeof v);
    std::memcpy(&b[at], &v, sizeof v);
    return at;
  }
  template <class T>
  void set(size_t at, T v) {
    std::memcpy(&b[at], &v, sizeof v);
  }
  void link(size_t slot, size_t target) { set<uint32_t>(slot, (uint32_t)(target - slot)); }
  struct F {
    int id;
    int size;  // 1, 2, 4, 8; an offset is a 4-byte field linked later
    uint64_t v;
  };
  // The table's position; slots[i] = the position of fs[i].
  size_t table(std::vector<F> fs, std::vector<size_t> *slots) {
    int nid = 0;
    for (auto &f : fs) nid = std::max(nid, f.id + 1);
    pad(2);
    const size_t vt = b.size();
    b.resize(vt + 4 + 2 * (size_t)nid, 0);
    pad(4);
    const size_t tp = b.size();
    put<int32_t>((int32_t)(tp - vt));
    std::vector<int> order(fs.size());
    for (size_t i = 0; i < fs.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return fs[x].size > fs[y].size; });
    if (slots) slots->assign(fs.size(), 0);
    for (int i : order) {
      pad((size_t)fs[i].size);
      const size_t at = b.size();
      switch (fs[i].size) {
        case 1: put<uint8_t>((uint8_t)fs[i].v); break;
        case 2: put<uint16_t>((uint16_t)fs[i].v); break;
        case 4: put<uint32_t>((uint32_t)fs[i].v); break;
        default: put<uint64_t>(fs[i].v);
      }
      set<uint16_t>(vt + 4 + 2 * (size_t)fs[i].id, (uint16_t)(at - tp));
      if (slots) (*slots)[i] = at;
    }
    set<uint16_t>(vt, (uint16_t)(4 + 2 * nid));
    set<uint16_t>(vt + 2, (uint16_t)(b.size() - tp));
    return tp;
  }
  size_t string(const std::string &s) {
    pad(4);
    const size_t at = put<uint32_t>((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    b.push_back(0);
    return at;
  }
  // a vector of 16-byte structs of two int64 each
  size_t pairs(const std::vector<std::pair<int64_t, int64_t>> &v) {
    while (b.size() % 8 != 4) b.push_back(0);
    const size_t at = put<uint32_t>((uint32_t)v.size());
    for (auto &p : v) {
      put<int64_t>(p.first);
      put<int64_t>(p.second);
    }
    return at;
  }
  // a vector of n offsets (linked later); returns its position, slots
  size_t offsets(size_t n, std::vector<size_t> *slots) {
    pad(4);
    const size_t at = put<uint32_t>((uint32_t)n);
    slots->clear();
    for (size_t i = 0; i < n; i++) slots->push_back(put<uint32_t>(0));
    return at;
  }
};

// Arrow's Schema.fbs / Message.fbs codes
enum : uint8_t { MH_SCHEMA = 1, MH_RECORD_BATCH = 3 };
enum : uint8_t { AT_INT = 2, AT_FLOAT = 3, AT_BOOL = 6, AT_DECIMAL = 7, AT_DATE = 8, AT_TIMESTAMP = 10 };
constexpr uint16_t kMetadataV5 = 4;

static void field_type(Fb &fb, const kg_column &c, uint8_t *type_type, size_t *type_pos) {
  switch (c.type) {
    case KG_T_BOOL: *type_type = AT_BOOL; *type_pos = fb.table({}, nullptr); return;
    case KG_T_INT8: case KG_T_INT16: case KG_T_INT32: case KG_T_INT64:
    case KG_T_UINT8: case KG_T_UINT16: case KG_T_UINT32: case KG_T_UINT64: {
      const bool sgn = c.type <= KG_T_INT64;
      *type_type = AT_INT;
      *type_pos = fb.table({{0, 4, (uint64_t)(8 * type_width(c.type))}, {1, 1, sgn ? 1u : 0u}},
                           nullptr);
      return;
    }
    case KG_T_FLOAT32: case KG_T_FLOAT64:
      *type_type = AT_FLOAT;
      *type_pos = fb.table({{0, 2, c.type == KG_T_FLOAT32 ? 1u : 2u}}, nullptr);
      return;
    case KG_T_DATE32:
      *type_type = AT_DATE;
      *type_pos = fb.table({{0, 2, 0}}, nullptr);  // DateUnit DAY
      return;
    case KG_T_TIMESTAMP:
      *type_type = AT_TIMESTAMP;
      *type_pos = fb.table({{0, 2, 2}}, nullptr);  // TimeUnit MICROSECOND, no zone
      return;
    case KG_T_DECIMAL64: case KG_T_DECIMAL128:
      *type_type = AT_DECIMAL;
      *type_pos = fb.table({{0, 4, (uint64_t)(uint32_t)c.precision},
                            {1, 4, (uint64_t)(uint32_t)c.scale},
                            {2, 4, c.type == KG_T_DECIMAL64 ? 64u : 128u}},
                           nullptr);
      return;
    default: fail(KG_ERR_UNSUPPORTED, "a column type with no Arrow IPC encoding here");
  }
}

// An encapsulated IPC message: continuation, metadata length, the
// flatbuffer padded to 8 bytes (the body follows).
static std::vector<uint8_t> encapsulate(const std::vector<uint8_t> &fb) {
  const size_t meta = (fb.size() + 7) / 8 * 8;
  std::vector<uint8_t> out(8 + meta, 0);
  const uint32_t cont = 0xFFFFFFFFu;
  const int32_t len = (int32_t)meta;
  std::memcpy(out.data(), &cont, 4);
  std::memcpy(out.data() + 4, &len, 4);
  std::memcpy(out.data() + 8, fb.data(), fb.size());
  return out;
}

static std::vector<uint8_t> schema_message(int ncols, const kg_column *cols,
                                           const char *const *names) {
  Fb fb;
  const size_t root = fb.put<uint32_t>(0);
  std::vector<size_t> ms;
  const size_t msg = fb.table({{0, 2, kMetadataV5}, {1, 1, MH_SCHEMA}, {2, 4, 0}}, &ms);
  fb.link(root, msg);
  std::vector<size_t> ss;
  const size_t schema = fb.table({{1, 4, 0}}, &ss);  // endianness Little (the default)
  fb.link(ms[2], schema);
  std::vector<size_t> fslots;
  const size_t fields = fb.offsets((size_t)ncols, &fslots);
  fb.link(ss[0], fields);
  for (int c = 0; c < ncols; c++) {
    std::vector<size_t> fs;
    uint8_t tt = 0;
    const size_t field = fb.table({{0, 4, 0}, {1, 1, 1}, {2, 1, 0}, {3, 4, 0}, {5, 4, 0}}, &fs);
    fb.link(fslots[c], field);
    const std::string nm = names && names[c] ? names[c] : "c" + std::to_string(c);
    fb.link(fs[0], fb.string(nm));
    size_t tpos = 0;
    field_type(fb, cols[c], &tt, &tpos);
    fb.b[fs[2]] = tt;
    fb.link(fs[3], tpos);
    std::vector<size_t> none;
    fb.link(fs[4], fb.offsets(0, &none));  // no children
  }
  return encapsulate(fb.b);
}

static std::vector<uint8_t> batch_message(int64_t rows,
                                          const std::vector<std::pair<int64_t, int64_t>> &nodes,
                                          const std::vector<std::pair<int64_t, int64_t>> &bufs,
                                          int64_t body_len) {
  Fb fb;
  const size_t root = fb.put<uint32_t>(0);
  std::vector<size_t> ms;
  const size_t msg = fb.table(
      {{0, 2, kMetadataV5}, {1, 1, MH_RECORD_BATCH}, {2, 4, 0}, {3, 8, (uint64_t)body_len}}, &ms);
  fb.link(root, msg);
  std::vector<size_t> rs;
  const size_t rb = fb.table({{0, 8, (uint64_t)rows}, {1, 4, 0}, {2, 4, 0}}, &rs);
  fb.link(ms[2], rb);
  fb.link(rs[1], fb.pairs(nodes));
  fb.link(rs[2], fb.pairs(bufs));
  return encapsulate(fb.b);
}

static uint64_t pad8(uint64_t v) { return (v + 7) / 8 * 8; }

}  // namespace kg

using namespace kg;

extern "C" {

int kg_frame_end(const kg_frame_header *h, uint8_t *out) {
  return guarded([&] {
    // an End frame: an empty body, CRC32C of nothing is 0 (frame.rs:132-150)
    std::memset(out, 0, KG_FRAME_HEADER_BYTES);
    std::memcpy(out, "KXFR", 4);
    out[4] = 1;
    out[6] = (uint8_t)KG_FRAME_END;
    std::memcpy(out + 8, &h->id, 8);
    std::memcpy(out + 16, &h->stage_attempt, 4);
    std::memcpy(out + 20, &h->map, 4);
    std::memcpy(out + 24, &h->attempt, 4);
    std::memcpy(out + 28, &h->sequence, 4);
    std::memcpy(out + 32, &h->reduce, 4);
  });
}

int kg_crc32c(kg_ctx *ctx, const uint8_t *data, uint64_t len, uint32_t *out) {
  return guarded([&] {
    *out = 0;
    if (len == 0) return;
    hipStream_t s = make_stream(ctx->device);
    struct Done {
      hipStream_t s;
      ~Done() { destroy_stream(s); }
    } done{s};
    StreamScope kg_scope(s);
    const uint64_t nchunks = (len + kCrcChunk - 1) / kCrcChunk;
    if (nchunks > UINT32_MAX) fail(KG_ERR_UNSUPPORTED, "input too large");
    std::vector<CrcChunk> ch(nchunks);
    for (uint64_t k = 0; k < nchunks; k++)
      ch[k] = CrcChunk{k * kCrcChunk, (uint32_t)std::min<uint64_t>(kCrcChunk, len - k * kCrcChunk), 0};
    CrcFrame F{0, len, 0, (uint32_t)nchunks};
    DevBuf dch(ctx, nchunks * sizeof(CrcChunk)), dcrc(ctx, nchunks * 4), dF(ctx, sizeof F),
        dout(ctx, 4);
    KG_HIP(hipMemcpyAsync(dch.get(), ch.data(), nchunks * sizeof(CrcChunk), hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(dF.get(), &F, sizeof F, hipMemcpyHostToDevice, s));
    {
      ProfScope prof("crc32c", s);
      k_crc_chunks<<<(unsigned)nchunks, kCrcBlock, 0, s>>>(data, dch.get<CrcChunk>(),
                                                          dcrc.get<uint32_t>());
      k_crc_frames<<<1, 64, 0, s>>>(dF.get<CrcFrame>(), 1, dch.get<CrcChunk>(),
                                    dcrc.get<uint32_t>(), nullptr, dout.get<uint32_t>());
      KG_HIP(hipGetLastError());
    }
    KG_HIP(hipMemcpyAsync(out, dout.get(), 4, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
  });
}

int kg_exchange_frames(kg_ctx *ctx, int32_t ncols, const kg_column *cols,
                       const char *const *names, const int64_t *part_offsets, uint32_t partitions,
                       const kg_frame_header *base, uint8_t *out, uint64_t capacity,
                       uint64_t *frame_offsets, uint64_t *total) {
  return guarded([&] {
    *total = 0;
    if (ncols <= 0) fail(KG_ERR_INTERNAL, "no columns");
    if (partitions == 0) fail(KG_ERR_INTERNAL, "no partitions");
    const int64_t nrows = part_offsets[partitions];
    if (part_offsets[0] != 0) fail(KG_ERR_INTERNAL, "partition offsets start at 0");
    for (uint32_t p = 0; p < partitions; p++)
      if (part_offsets[p + 1] < part_offsets[p]) fail(KG_ERR_INTERNAL, "partition offsets decrease");
    for (int c = 0; c < ncols; c++) {
      check_column(cols[c], nrows, "frame column");
      if (type_width(cols[c].type) < 0) fail(KG_ERR_UNSUPPORTED, "frame column type");
    }
    hipStream_t s = make_stream(ctx->device);
    struct Done {
      hipStream_t s;
      ~Done() { destroy_stream(s); }
    } done{s};
    Staged st(ctx, s);
    std::vector<DCol> dc(ncols);
    for (int c = 0; c < ncols; c++) dc[c] = st.add(cols[c], nrows);
    // null counts per (partition, column)
    std::vector<uint64_t> nulls((size_t)partitions * ncols, 0);
    {
      bool any = false;
      for (int c = 0; c < ncols; c++) any |= dc[c].validity != nullptr;
      if (any && nrows > 0) {
        DevBuf doff(ctx, (partitions + 1) * 8), dn(ctx, nulls.size() * 8);
        KG_HIP(hipMemcpyAsync(doff.get(), part_offsets, (partitions + 1) * 8, hipMemcpyHostToDevice, s));
        KG_HIP(hipMemsetAsync(dn.get(), 0, nulls.size() * 8, s));
        for (int c = 0; c < ncols; c++)
          if (dc[c].validity)
            k_frame_nulls<<<stream_grid(nrows, 256), 256, 0, s>>>(
                dc[c].validity, dc[c].offset, doff.get<int64_t>(), partitions, nrows, c, ncols,
                dn.get<unsigned long long>());
        KG_HIP(hipGetLastError());
        KG_HIP(hipMemcpyAsync(nulls.data(), dn.get(), nulls.size() * 8, hipMemcpyDeviceToHost, s));
        KG_HIP(hipStreamSynchronize(s));
      }
    }
    // the layout of every frame, its metadata, its copy items
    const std::vector<uint8_t> schema = schema_message(ncols, cols, names);
    const uint8_t eos[8] = {0xFF, 0xFF, 0xFF, 0xFF, 0, 0, 0, 0};
    std::vector<uint8_t> blob;  // headers and IPC metadata, uploaded once
    struct Pending {
      uint64_t blob_at, dst, len;
    };
    std::vector<Pending> blobs;
    std::vector<CopyItem> items;
    std::vector<CrcChunk> chunks;
    std::vector<CrcFrame> frames;
    uint64_t at = 0;
    for (uint32_t p = 0; p < partitions; p++) {
      frame_offsets[p] = at;
      const int64_t r0 = part_offsets[p], rows = part_offsets[p + 1] - r0;
      if (rows == 0) continue;
      // the body's buffers: validity then values, each padded to 8 bytes
      std::vector<std::pair<int64_t, int64_t>> nodes, bufs;
      uint64_t body = 0;
      struct Buf {
        int col;
        bool validity;
        uint64_t off, len;
      };
      std::vector<Buf> layout;
      for (int c = 0; c < ncols; c++) {
        const uint64_t nn = nulls[(size_t)p * ncols + c];
        nodes.push_back({rows, (int64_t)nn});
        const uint64_t vlen = nn ? (uint64_t)(rows + 7) / 8 : 0;
        bufs.push_back({(int64_t)body, (int64_t)vlen});
        if (vlen) layout.push_back({c, true, body, vlen});
        body += pad8(vlen);
        const int w = type_width(cols[c].type);
        const uint64_t dlen = w == 0 ? (uint64_t)(rows + 7) / 8 : (uint64_t)rows * w;
        bufs.push_back({(int64_t)body, (int64_t)dlen});
        layout.push_back({c, false, body, dlen});
        body += pad8(dlen);
      }
      const std::vector<uint8_t> rb = batch_message(rows, nodes, bufs, (int64_t)body);
      const uint64_t body_at = at + KG_FRAME_HEADER_BYTES;
      const uint64_t data_at = body_at + schema.size() + rb.size();
      const uint64_t body_len = schema.size() + rb.size() + body + 8;
      if (body_len > UINT32_MAX) fail(KG_ERR_UNSUPPORTED, "a frame body past 4 GiB");
      // header (CRC patched on the device), schema, batch metadata
      uint8_t h[KG_FRAME_HEADER_BYTES] = {};
      std::memcpy(h, "KXFR", 4);
      h[4] = 1;
      h[6] = (uint8_t)KG_FRAME_DATA;
      std::memcpy(h + 8, &base->id, 8);
      std::memcpy(h + 16, &base->stage_attempt, 4);
      std::memcpy(h + 20, &base->map, 4);
      std::memcpy(h + 24, &base->attempt, 4);
      std::memcpy(h + 28, &base->sequence, 4);
      std::memcpy(h + 32, &p, 4);
      const uint32_t bl = (uint32_t)body_len;
      std::memcpy(h + 36, &bl, 4);
      const uint64_t b0 = blob.size();
      blob.insert(blob.end(), h, h + KG_FRAME_HEADER_BYTES);
      blob.insert(blob.end(), schema.begin(), schema.end());
      blob.insert(blob.end(), rb.begin(), rb.end());
      blobs.push_back({b0, at, (uint64_t)KG_FRAME_HEADER_BYTES + schema.size() + rb.size()});
      const uint64_t e0 = blob.size();
      blob.insert(blob.end(), eos, eos + 8);
      blobs.push_back({e0, data_at + body, 8});
      for (const Buf &b : layout) {
        const DCol &d = dc[b.col];
        const int w = type_width(cols[b.col].type);
        if (b.validity || w == 0) {
          const uint8_t *src = b.validity ? d.validity : static_cast<const uint8_t *>(d.values);
          const uint64_t first = (uint64_t)(d.offset + r0);
          const uint64_t readable = (uint64_t)(d.offset + nrows + 7) / 8 - first / 8;
          CopyItem it{src, data_at + b.off, first, (uint32_t)pad8(b.len), (uint32_t)rows, 1,
                      (uint32_t)std::min<uint64_t>(readable, UINT32_MAX)};
          if (pad8(b.len) > UINT32_MAX) fail(KG_ERR_UNSUPPORTED, "a bitmap past 4 GiB");
          items.push_back(it);
        } else {
          const uint8_t *src = static_cast<const uint8_t *>(d.values) + (uint64_t)(d.offset + r0) * w;
          const uint64_t padded = pad8(b.len);
          for (uint64_t o = 0; o < padded; o += kCopyChunk) {
            const uint64_t nb = std::min<uint64_t>(kCopyChunk, padded - o);
            const uint64_t nd = b.len > o ? std::min<uint64_t>(nb, b.len - o) : 0;
            items.push_back(CopyItem{src + o, data_at + b.off + o, 0, (uint32_t)nb, (uint32_t)nd, 0, 0});
          }
        }
      }
      // the body's CRC chunks
      CrcFrame F{at, body_len, (uint32_t)chunks.size(), 0};
      for (uint64_t o = 0; o < body_len; o += kCrcChunk) {
        chunks.push_back(CrcChunk{body_at + o, (uint32_t)std::min<uint64_t>(kCrcChunk, body_len - o),
                                  (uint32_t)frames.size()});
        F.count++;
      }
      frames.push_back(F);
      at += KG_FRAME_HEADER_BYTES + body_len;
    }
    frame_offsets[partitions] = at;
    *total = at;
    if (!out) return;  // sizes only
    if (capacity < at) fail(KG_ERR_INTERNAL, "frame buffer too small");
    if (frames.empty()) return;
    // the metadata copies are items too (from the uploaded blob)
    DevBuf dblob(ctx, blob.size());
    KG_HIP(hipMemcpyAsync(dblob.get(), blob.data(), blob.size(), hipMemcpyHostToDevice, s));
    for (const Pending &b : blobs)
      items.push_back(CopyItem{dblob.get<uint8_t>() + b.blob_at, b.dst, 0, (uint32_t)b.len,
                               (uint32_t)b.len, 0, 0});
    DevBuf ditems(ctx, items.size() * sizeof(CopyItem)), dch(ctx, chunks.size() * sizeof(CrcChunk)),
        dcrc(ctx, chunks.size() * 4), dfr(ctx, frames.size() * sizeof(CrcFrame));
    KG_HIP(hipMemcpyAsync(ditems.get(), items.data(), items.size() * sizeof(CopyItem),
                          hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(dch.get(), chunks.data(), chunks.size() * sizeof(CrcChunk),
                          hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(dfr.get(), frames.data(), frames.size() * sizeof(CrcFrame),
                          hipMemcpyHostToDevice, s));
    {
      ProfScope prof("frame_copy", s);
      k_frame_copy<<<(unsigned)items.size(), 256, 0, s>>>(ditems.get<CopyItem>(), out);
      KG_HIP(hipGetLastError());
    }
    {
      ProfScope prof("frame_crc", s);
      k_crc_chunks<<<(unsigned)chunks.size(), kCrcBlock, 0, s>>>(out, dch.get<CrcChunk>(),
                                                                dcrc.get<uint32_t>());
      k_crc_frames<<<stream_grid((int64_t)frames.size(), 64, 1), 64, 0, s>>>(
          dfr.get<CrcFrame>(), (int64_t)frames.size(), dch.get<CrcChunk>(), dcrc.get<uint32_t>(),
          out, nullptr);
      KG_HIP(hipGetLastError());
    }
    KG_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
