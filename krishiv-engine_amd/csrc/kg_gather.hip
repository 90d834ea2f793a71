// kg_gather.hip -- host batches gathered into HBM through pinned staging
// (the aggregation's and the fused join aggregate's small host batches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "kg_gather.hpp"

namespace kg {

namespace {

constexpr int64_t kGatherRows = 1 << 22;
constexpr uint64_t kGatherBytes = 256ull << 20;

// n bits from src (starting at bit sbit) to dst (starting at bit dbit), LSB
// first, as Arrow bitmaps; src == nullptr writes ones.
void copy_bits(uint8_t *dst, int64_t dbit, const uint8_t *src, int64_t sbit, int64_t n) {
  auto get = [&](int64_t i) { return src ? (src[(sbit + i) >> 3] >> ((sbit + i) & 7)) & 1 : 1; };
  auto put = [&](int64_t i, int b) {
    uint8_t &d = dst[(dbit + i) >> 3];
    const int sh = (int)((dbit + i) & 7);
    d = (uint8_t)((d & ~(1u << sh)) | ((unsigned)b << sh));
  };
  int64_t i = 0;
  for (; i < n && ((dbit + i) & 7); i++) put(i, get(i));
  const int64_t full = (n - i) >> 3;
  uint8_t *d = dst + ((dbit + i) >> 3);
  if (!src) {
    memset(d, 0xFF, (size_t)full);
  } else if (((sbit + i) & 7) == 0) {
    memcpy(d, src + ((sbit + i) >> 3), (size_t)full);
  } else {
    const int sh = (int)((sbit + i) & 7);
    const uint8_t *p = src + ((sbit + i) >> 3);
    for (int64_t k = 0; k < full; k++) d[k] = (uint8_t)((p[k] >> sh) | (p[k + 1] << (8 - sh)));
  }
  i += full * 8;
  for (; i < n; i++) put(i, get(i));
}

}  // namespace

int64_t gather_max_rows(const int *width, int ncols) {
  uint64_t bits = 0;  // bits a row: values and validity
  for (int c = 0; c < ncols; c++) {
    if (width[c] < 0) return 0;
    bits += (width[c] ? 8u * (uint64_t)width[c] : 1u) + 1;
  }
  int64_t cap = kGatherRows;
  while (cap > 4096 && (uint64_t)cap * bits / 8 > kGatherBytes) cap >>= 1;
  return cap;
}

int64_t gather_rows_for(int64_t nrows, int64_t max_rows) {
  int64_t want = 65536;
  while (want < nrows * 64 && want < max_rows) want <<= 1;
  return std::min(want, max_rows);
}

HostGather::HostGather(kg_ctx *ctx, int ncols, const int *width, int64_t cap)
    : ctx_(ctx), ncols_(ncols), cap_(cap) {
  if (ncols < 1 || ncols > kGatherCols) fail(KG_ERR_INTERNAL, "gather columns");
  const int64_t srows = cap / 8;
  size_t off = 0;
  for (int c = 0; c < ncols; c++) {
    width_[c] = width[c];
    const uint64_t w = (uint64_t)width[c];
    vals_[c].reset(ctx, w ? cap * w : cap / 8);
    valid_[c].reset(ctx, cap / 8);
    voff_[c] = off;
    off += ((w ? srows * w : srows / 8) + 255) & ~(size_t)255;
    boff_[c] = off;
    off += ((size_t)srows / 8 + 255) & ~(size_t)255;
  }
  for (Slot &sl : slot_) {
    KG_HIP(hipHostMalloc((void **)&sl.host, off, hipHostMallocDefault));
    KG_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  }
}

HostGather::~HostGather() {
  for (Slot &sl : slot_) {
    if (sl.done) (void)hipEventSynchronize(sl.done);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.host) (void)hipHostFree(sl.host);
  }
}

bool HostGather::same_types(const kg_column *const *cols) const {
  if (rows_ == 0) return true;
  for (int c = 0; c < ncols_; c++)
    if (cols[c]->type != type_[c].type || cols[c]->precision != type_[c].precision ||
        cols[c]->scale != type_[c].scale)
      return false;
  return true;
}

// The open slot's rows to HBM; the other slot opens once its last DMA has
// read the staging.
void HostGather::submit_slot(hipStream_t s) {
  Slot &sl = open();
  if (sl.rows == 0) return;
  for (int c = 0; c < ncols_; c++) {
    const int w = width_[c];
    const uint64_t vb = w ? (uint64_t)sl.rows * w : (uint64_t)ceil_div(sl.rows, 8);
    uint8_t *dv = vals_[c].get<uint8_t>() + (w ? (uint64_t)sl.base * w : (uint64_t)sl.base / 8);
    KG_HIP(hipMemcpyAsync(dv, sl.host + voff_[c], vb, hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(valid_[c].get<uint8_t>() + sl.base / 8, sl.host + boff_[c],
                          (size_t)ceil_div(sl.rows, 8), hipMemcpyHostToDevice, s));
  }
  KG_HIP(hipEventRecord(sl.done, s));
  const int64_t next = sl.base + sl.rows;
  cur_ ^= 1;
  Slot &nx = open();
  KG_HIP(hipEventSynchronize(nx.done));
  nx.base = next;
  nx.rows = 0;
}

void HostGather::append(hipStream_t s, const kg_column *const *cols, int64_t n) {
  if (rows_ == 0)
    for (int c = 0; c < ncols_; c++) {
      type_[c] = kg_column{};
      type_[c].type = cols[c]->type;
      type_[c].precision = cols[c]->precision;
      type_[c].scale = cols[c]->scale;
    }
  const int64_t srows = slot_rows();
  for (int64_t done = 0; done < n;) {
    Slot &sl = open();
    const int64_t take = std::min<int64_t>(n - done, srows - sl.rows);
    for (int c = 0; c < ncols_; c++) {
      const kg_column &src = *cols[c];
      const int w = width_[c];
      const int64_t first = src.offset + done;
      if (w)
        memcpy(sl.host + voff_[c] + (size_t)sl.rows * w,
               static_cast<const uint8_t *>(src.values) + (size_t)first * w, (size_t)take * w);
      else
        copy_bits(sl.host + voff_[c], sl.rows, static_cast<const uint8_t *>(src.values), first,
                  take);
      copy_bits(sl.host + boff_[c], sl.rows, src.validity, first, take);
      if (src.validity) any_valid_[c] = true;
    }
    sl.rows += take;
    rows_ += take;
    done += take;
    if (sl.rows == srows) submit_slot(s);
  }
}

void HostGather::submit(hipStream_t s, kg_column *out) {
  cur_before_submit_ = cur_;
  submit_slot(s);
  for (int c = 0; c < ncols_; c++) {
    out[c] = type_[c];
    out[c].flags = KG_COL_DEVICE;
    out[c].length = rows_;
    out[c].offset = 0;
    out[c].values = vals_[c].get();
    out[c].validity = any_valid_[c] ? valid_[c].get<uint8_t>() : nullptr;
  }
}

void HostGather::clear() {
  rows_ = 0;
  for (int c = 0; c < ncols_; c++) any_valid_[c] = false;
  // the open slot (emptied by the submit) starts the buffers again; the
  // stream orders its DMA after the work that read them
  open().base = 0;
  open().rows = 0;
}

void HostGather::keep() {
  // the slot open before the submit stays open, so later rows keep their
  // place in it; its copy in flight finishes first
  cur_ = cur_before_submit_;
  (void)hipEventSynchronize(open().done);
}

}  // namespace kg
