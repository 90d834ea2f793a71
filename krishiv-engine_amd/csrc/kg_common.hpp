// kg_common.hpp -- shared runtime pieces of libkgpu: errors at the C ABI,
// the context (device, HBM budget), device buffers, columns as kernels see
// them, and launch profiling.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <mutex>
#include <string>
#include <vector>

#include "kgpu.h"

namespace kg {

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};

const char *class_of(int code);
void set_last_error(int code, const std::string &msg);

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

#define KG_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      ::kg::fail(KG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Wraps a C entry point body: any kg::Error (or other exception) becomes a
// status code and the thread's last error.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return KG_OK;
  } catch (const Error &e) {
    set_last_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error(KG_ERR_MEMORY, "host allocation failed");
    return KG_ERR_MEMORY;
  } catch (const std::exception &e) {
    set_last_error(KG_ERR_INTERNAL, e.what());
    return KG_ERR_INTERNAL;
  } catch (...) {
    set_last_error(KG_ERR_INTERNAL, "unknown failure");
    return KG_ERR_INTERNAL;
  }
}

}  // namespace kg

// ---------------------------------------------------------------- context
namespace kg {
// An event recorded on a stream; destroyed with its last holder.
class Fence {
 public:
  explicit Fence(hipStream_t s);
  ~Fence();
  bool done() const;
  Fence(const Fence &) = delete;
  Fence &operator=(const Fence &) = delete;

 private:
  hipEvent_t e_ = nullptr;
};

// Shape parameters of the hash join and its fused aggregate, per size class
// of the build side (kg_tune_*): the table's load factor, the staged
// partition scatter's rows per lane, and the partition merge's workgroup.
struct TuneCfg {
  int32_t load_pct = 50;     // table capacity >= rows * 100 / load_pct (a power of two);
                             // untuned classes take a per-class value (tune_for)
  int32_t scatter_rows = 0;  // 4, 8 or 12; 0: by the row width
  int32_t merge_lanes = 1024;  // 1024 (one workgroup a CU, the whole LDS) or 512 (two)
};
}  // namespace kg

struct kg_ctx {
  int device = 0;
  uint64_t budget = 0;
  std::atomic<uint64_t> in_use{0};
  hipStream_t stream = nullptr;  // for context-level calls
  int num_cus = 0;
  uint64_t total_hbm = 0;
  // freed device blocks kept for reuse, by rounded size
  std::mutex cache_mu;
  std::multimap<uint64_t, void *> cache;
  uint64_t cached = 0;
  // allocator counters (kg_ctx_mem_stats); the plain ones under cache_mu
  std::atomic<uint64_t> n_mallocs{0};
  uint64_t n_frees = 0, n_hits = 0, n_releases = 0;
  // blocks freed with stream order: reusable once every event completes
  struct Pending {
    uint64_t rb;
    void *p;
    std::vector<std::shared_ptr<kg::Fence>> waits;
  };
  std::vector<Pending> pending;
  // events after the last enqueued reads of caller-owned device columns
  std::vector<std::shared_ptr<kg::Fence>> use_fences;
  // tuned shapes by build-side size class (ceil log2 of the rows)
  std::mutex tune_mu;
  std::map<int, kg::TuneCfg> tune;
};

namespace kg {

void *dev_alloc(kg_ctx *ctx, uint64_t bytes);
// The tuned shape for a build side of `rows` (defaults where none is stored).
TuneCfg tune_for(kg_ctx *ctx, int64_t rows);
void tune_store(kg_ctx *ctx, int64_t rows, const TuneCfg &c);
// Returns a size class to its untuned default.
void tune_clear(kg_ctx *ctx, int64_t rows);
// Whether a size class has a tuned (or set) shape.
bool tune_has(kg_ctx *ctx, int64_t rows);
// Frees a block.  With a stream, the block is reused only after the work
// enqueued on that stream so far, and every read of caller columns enqueued
// so far (use fences), has completed; without one, after a device
// synchronize.
void dev_free(kg_ctx *ctx, void *p, uint64_t bytes, hipStream_t stream = nullptr);
// Records that work reading caller-owned device columns has been enqueued
// on stream s (blocks freed later wait for it).
void note_column_use(kg_ctx *ctx, hipStream_t s);

// The stream whose work uses the device buffers allocated in the current
// scope of this thread: an operator entry point sets its own stream, so the
// buffers it allocates are freed in that stream's order.
hipStream_t alloc_stream();
class StreamScope {
 public:
  explicit StreamScope(hipStream_t s);
  ~StreamScope();
  StreamScope(const StreamScope &) = delete;
  StreamScope &operator=(const StreamScope &) = delete;

 private:
  hipStream_t prev_;
};

// RAII device buffer charged to a context's HBM budget.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(kg_ctx *ctx, uint64_t bytes) { reset(ctx, bytes); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept { swap(o); }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      swap(o);
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void reset(kg_ctx *ctx, uint64_t bytes) {
    release();
    ctx_ = ctx;
    bytes_ = bytes;
    s_ = alloc_stream();
    p_ = bytes ? dev_alloc(ctx, bytes) : nullptr;
  }
  // Keeps the buffer when it is already large enough.
  void ensure(kg_ctx *ctx, uint64_t bytes) {
    if (bytes > bytes_) reset(ctx, bytes + bytes / 4);
  }
  void release() {
    if (p_) dev_free(ctx_, p_, bytes_, s_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <class T = void>
  T *get() const {
    return static_cast<T *>(p_);
  }
  uint64_t bytes() const { return bytes_; }
  void *detach() {
    void *p = p_;
    p_ = nullptr;
    bytes_ = 0;
    return p;
  }

 private:
  void swap(DevBuf &o) {
    std::swap(ctx_, o.ctx_);
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    std::swap(s_, o.s_);
  }
  kg_ctx *ctx_ = nullptr;
  void *p_ = nullptr;
  uint64_t bytes_ = 0;
  hipStream_t s_ = nullptr;  // the stream whose order frees it (none: device-wide)
};

// ---------------------------------------------------------------- columns
// A column as kernels read it (device pointers only).
struct DCol {
  int32_t type;
  int32_t precision;
  int32_t scale;
  int32_t width;  // bytes per value (0 for bit-packed BOOL)
  int64_t offset;
  const void *values;
  const uint8_t *validity;
};

int type_width(int t);  // -1 for an unknown type
bool is_integral(int t);
bool is_float(int t);
bool is_decimal(int t);
bool is_int_like(int t);  // integers, dates, timestamps, decimal64, bool

// Host-side staging of kg_column inputs: host columns are copied to HBM
// (synchronously, so the caller may free them on return); device columns
// are used in place.
class Staged {
 public:
  Staged(kg_ctx *ctx, hipStream_t s) : ctx_(ctx), s_(s), scope_(s) {}
  // the caller's device columns are read by work enqueued on s_ by now
  ~Staged() {
    if (device_) note_column_use(ctx_, s_);
  }
  DCol add(const kg_column &c, int64_t nrows);
  bool staged_host() const { return !bufs_.empty(); }
  // The staged copies, for a caller that reads them after this call.
  std::vector<DevBuf> take_buffers() { return std::move(bufs_); }

 private:
  kg_ctx *ctx_;
  hipStream_t s_;
  StreamScope scope_;
  bool device_ = false;
  std::vector<DevBuf> bufs_;
};

void check_column(const kg_column &c, int64_t nrows, const char *what);

// hipFuncSetAttribute(MaxDynamicSharedMemorySize) for the current device,
// made only when `bytes` passes what was set before for this kernel.
hipError_t set_dynamic_lds(const void *fn, size_t bytes);

// Small device-to-host readbacks through this thread's pinned buffer.  A
// copy into pageable memory waits for itself (the runtime stages it), a copy
// into pinned memory queues on the stream: a call that reads several values
// back waits once.  Space taken is returned when the PinnedReads goes out of
// scope, so the copies into it must have finished by then (the caller
// synchronizes its stream before it reads the values).
class PinnedReads {
 public:
  PinnedReads();
  ~PinnedReads();
  PinnedReads(const PinnedReads &) = delete;
  PinnedReads &operator=(const PinnedReads &) = delete;
  template <class T>
  T *take(size_t n) {
    return static_cast<T *>(take_bytes(n * sizeof(T)));
  }

 private:
  void *take_bytes(size_t bytes);
  size_t mark_;
};

// ---------------------------------------------------------------- profiling
// Records a HIP event pair around a launch when profiling is on.
class ProfScope {
 public:
  ProfScope(const char *family, hipStream_t s);
  ~ProfScope();

 private:
  const char *family_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Streams: every operator owns one stream at the default priority.
hipStream_t make_stream(int device);
void destroy_stream(hipStream_t s);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// RuntimeBloom::expected_fpp of a split-block filter of nblocks 32-byte
// blocks after `items` hashes (bloom.rs:270-308).
double sbbf_expected_fpp(uint64_t nblocks, uint64_t items);

// Grid size for a streaming kernel over n items: enough workgroups to fill
// the 256 CUs several times over, capped (grid-stride loops do the rest).
inline unsigned stream_grid(int64_t n, int block, int64_t per_thread = 4) {
  int64_t g = ceil_div(n, (int64_t)block * per_thread);
  if (g < 1) g = 1;
  if (g > 256 * 32) g = 256 * 32;
  return (unsigned)g;
}

}  // namespace kg

// ---------------------------------------------------------------- batches
struct kg_batch {
  kg_ctx *ctx = nullptr;
  int64_t nrows = 0;
  struct Col {
    int32_t type = 0, precision = 0, scale = 0;
    kg::DevBuf values;
    kg::DevBuf validity;  // LSB bitmap, or empty when no NULLs
  };
  std::vector<Col> cols;
};
