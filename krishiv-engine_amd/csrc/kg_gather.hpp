// kg_gather.hpp -- small host batches gathered into HBM before an operator
// takes them.
//
// A batch from host memory alone costs a synchronous copy and a launch over
// a few workgroups.  A gather copies each host batch with the CPU into one of
// two pinned staging slots (so the caller's buffers are free when the call
// returns); a full slot goes by DMA on the operator's stream into HBM
// buffers of cap rows; the operator takes the gathered rows as one device
// batch when the next batch does not fit, a device batch arrives, or at the
// end.  Validity bitmaps (and BOOL values) are re-aligned bit by bit while
// they are copied.
#pragma once

#include <hip/hip_runtime.h>

#include "kg_common.hpp"

namespace kg {

constexpr int kGatherCols = 33;  // the most columns a gather holds

// The most rows a gather of columns of these widths (bytes a value; 0 for
// bit-packed BOOL) holds: 4M rows, fewer when the buffers would pass 256 MB.
int64_t gather_max_rows(const int *width, int ncols);

// The rows of a gather for batches of nrows: 64 batches, 64K at least,
// max_rows at most (powers of two).
int64_t gather_rows_for(int64_t nrows, int64_t max_rows);

class HostGather {
 public:
  HostGather(kg_ctx *ctx, int ncols, const int *width, int64_t cap);
  ~HostGather();
  HostGather(const HostGather &) = delete;
  HostGather &operator=(const HostGather &) = delete;

  int64_t rows() const { return rows_; }
  int64_t cap() const { return cap_; }
  int64_t slot_rows() const { return cap_ / 8; }
  int ncols() const { return ncols_; }
  int width(int c) const { return width_[c]; }
  // Whether these columns have the gathered rows' types (type, precision,
  // scale); true when nothing is gathered.
  bool same_types(const kg_column *const *cols) const;
  // The type of gathered column c (valid while rows() > 0).
  int type(int c) const { return type_[c].type; }
  // Rows [0, n) of each host column (at its offset) into the staging; full
  // slots go to HBM on s.  The caller checked rows() + n <= cap().
  void append(hipStream_t s, const kg_column *const *cols, int64_t n);
  // Every gathered row to HBM on s; out[c] views column c's gathered rows
  // (device columns, valid once s reaches them).
  void submit(hipStream_t s, kg_column *out);
  // After submit: the rows were taken (clear), or taking them failed (keep:
  // they stay gathered and the next submit copies them again).
  void clear();
  void keep();

 private:
  struct Slot {
    uint8_t *host = nullptr;     // pinned
    hipEvent_t done = nullptr;   // its last DMA
    int64_t base = 0, rows = 0;  // first gathered row, rows staged
  };
  Slot &open() { return slot_[cur_]; }
  void submit_slot(hipStream_t s);

  kg_ctx *ctx_;
  int ncols_;
  int64_t cap_;
  int width_[kGatherCols]{};
  size_t voff_[kGatherCols]{}, boff_[kGatherCols]{};  // within a staging slot
  bool any_valid_[kGatherCols]{};  // some gathered batch had a validity bitmap
  kg_column type_[kGatherCols]{};
  DevBuf vals_[kGatherCols], valid_[kGatherCols];  // cap rows each
  int64_t rows_ = 0;               // rows gathered (the open slot's included)
  Slot slot_[2];
  int cur_ = 0, cur_before_submit_ = 0;
};

}  // namespace kg
