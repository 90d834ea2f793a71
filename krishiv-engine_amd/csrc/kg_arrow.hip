// kg_arrow.hip -- the Arrow C Data Interface (and C Device Data Interface)
// at the library's boundary: an Arrow array viewed as a kg_column without a
// copy, and a device batch exported as an Arrow struct array, either
// zero-copy in HBM (ArrowDeviceArray, ARROW_DEVICE_ROCM) or copied to host
// memory.  The reference's operators exchange arrow_array RecordBatches
// (crates/krishiv-types/src/batch.rs); a binding passes them across the C
// ABI this way.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kg_common.hpp"

namespace kg {

// Format string of an Arrow C schema -> (type, precision, scale).
static void type_of_format(const char *f, kg_column *out) {
  if (!f) fail(KG_ERR_INTERNAL, "an Arrow schema without a format");
  out->precision = out->scale = 0;
  const std::string s(f);
  if (s == "b") out->type = KG_T_BOOL;
  else if (s == "c") out->type = KG_T_INT8;
  else if (s == "s") out->type = KG_T_INT16;
  else if (s == "i") out->type = KG_T_INT32;
  else if (s == "l") out->type = KG_T_INT64;
  else if (s == "C") out->type = KG_T_UINT8;
  else if (s == "S") out->type = KG_T_UINT16;
  else if (s == "I") out->type = KG_T_UINT32;
  else if (s == "L") out->type = KG_T_UINT64;
  else if (s == "f") out->type = KG_T_FLOAT32;
  else if (s == "g") out->type = KG_T_FLOAT64;
  else if (s == "tdD") out->type = KG_T_DATE32;
  else if (s == "tdm" || s.rfind("ts", 0) == 0 || s.rfind("tD", 0) == 0)
    out->type = KG_T_TIMESTAMP;  // 64-bit temporals: date64, timestamps, durations
  else if (s.rfind("d:", 0) == 0) {
    int p = 0, sc = 0, bits = 128;
    const int n = std::sscanf(f + 2, "%d,%d,%d", &p, &sc, &bits);
    if (n < 2) fail(KG_ERR_INTERNAL, "a malformed decimal format");
    if (bits == 128) out->type = KG_T_DECIMAL128;
    else if (bits == 64) out->type = KG_T_DECIMAL64;
    else fail(KG_ERR_UNSUPPORTED, "a decimal of " + std::to_string(bits) + " bits");
    out->precision = p;
    out->scale = sc;
  } else {
    fail(KG_ERR_UNSUPPORTED, "the Arrow format '" + s + "' at the C ABI");
  }
}

static std::string format_of(const kg_batch::Col &c) {
  switch (c.type) {
    case KG_T_BOOL: return "b";
    case KG_T_INT8: return "c";
    case KG_T_INT16: return "s";
    case KG_T_INT32: return "i";
    case KG_T_INT64: return "l";
    case KG_T_UINT8: return "C";
    case KG_T_UINT16: return "S";
    case KG_T_UINT32: return "I";
    case KG_T_UINT64: return "L";
    case KG_T_FLOAT32: return "f";
    case KG_T_FLOAT64: return "g";
    case KG_T_DATE32: return "tdD";
    case KG_T_TIMESTAMP: return "tsu:";
    case KG_T_DECIMAL64:
      return "d:" + std::to_string(c.precision) + "," + std::to_string(c.scale) + ",64";
    case KG_T_DECIMAL128:
      return "d:" + std::to_string(c.precision) + "," + std::to_string(c.scale);
    default: fail(KG_ERR_UNSUPPORTED, "a column type with no Arrow format here");
  }
}

static void view(const ArrowArray *a, const ArrowSchema *s, bool device, kg_column *out) {
  if (!a || !s || !a->release || !s->release) fail(KG_ERR_INTERNAL, "a released Arrow array");
  if (s->dictionary || a->dictionary || a->n_children != 0)
    fail(KG_ERR_UNSUPPORTED, "a dictionary or nested Arrow array at the C ABI");
  kg_column c{};
  type_of_format(s->format, &c);
  if (a->n_buffers != 2) fail(KG_ERR_INTERNAL, "a primitive Arrow array has two buffers");
  c.flags = device ? KG_COL_DEVICE : 0;
  c.length = a->length;
  c.offset = a->offset;
  c.values = a->buffers[1];
  c.validity = a->null_count == 0 ? nullptr : static_cast<const uint8_t *>(a->buffers[0]);
  if (c.length > 0 && !c.values) fail(KG_ERR_INTERNAL, "an Arrow array without values");
  *out = c;
}

// ---- exported arrays: one allocation holds every struct and pointer array
struct Export {
  kg_batch *batch = nullptr;  // device export: the batch moved in
  std::vector<std::vector<uint8_t>> host;  // host export: the copied buffers
  std::vector<ArrowArray> children;
  std::vector<ArrowArray *> child_ptrs;
  std::vector<std::vector<const void *>> bufs;
  const void *parent_bufs[1] = {nullptr};
};

struct SchemaExport {
  std::vector<std::string> formats, names;
  std::vector<ArrowSchema> children;
  std::vector<ArrowSchema *> child_ptrs;
};

static void release_child_array(ArrowArray *a) { a->release = nullptr; }
static void release_child_schema(ArrowSchema *s) { s->release = nullptr; }

static void release_array(ArrowArray *a) {
  if (!a || !a->release) return;
  auto *e = static_cast<Export *>(a->private_data);
  for (auto &c : e->children)
    if (c.release) c.release(&c);
  if (e->batch) kg_batch_release(e->batch);
  delete e;
  a->release = nullptr;
}

static void release_schema(ArrowSchema *s) {
  if (!s || !s->release) return;
  auto *e = static_cast<SchemaExport *>(s->private_data);
  for (auto &c : e->children)
    if (c.release) c.release(&c);
  delete e;
  s->release = nullptr;
}

static void export_schema(const kg_batch *b, const char *const *names, ArrowSchema *out) {
  auto e = std::make_unique<SchemaExport>();
  const size_t n = b->cols.size();
  e->formats.reserve(n);
  e->names.reserve(n);
  e->children.resize(n);
  e->child_ptrs.resize(n);
  for (size_t c = 0; c < n; c++) {
    e->formats.push_back(format_of(b->cols[c]));
    e->names.push_back(names && names[c] ? names[c] : "c" + std::to_string(c));
  }
  for (size_t c = 0; c < n; c++) {
    ArrowSchema &s = e->children[c];
    s = ArrowSchema{};
    s.format = e->formats[c].c_str();
    s.name = e->names[c].c_str();
    s.flags = ARROW_FLAG_NULLABLE;
    s.release = release_child_schema;
    e->child_ptrs[c] = &s;
  }
  *out = ArrowSchema{};
  out->format = "+s";
  out->name = "";
  out->n_children = (int64_t)n;
  out->children = e->child_ptrs.data();
  out->release = release_schema;
  out->private_data = e.release();
}

static int64_t null_count_of(const kg_batch *b, size_t c, hipStream_t s) {
  const auto &col = b->cols[c];
  if (!col.validity.bytes() || b->nrows == 0) return 0;
  std::vector<uint8_t> bits((size_t)ceil_div(b->nrows, 8));
  KG_HIP(hipMemcpyAsync(bits.data(), col.validity.get(), bits.size(), hipMemcpyDeviceToHost, s));
  KG_HIP(hipStreamSynchronize(s));
  int64_t valid = 0;
  for (int64_t i = 0; i < b->nrows; i++) valid += (bits[i >> 3] >> (i & 7)) & 1;
  return b->nrows - valid;
}

// The struct array of a batch; buffer(c, k) gives child c's buffer k.
template <class Buf>
static void export_array(const kg_batch *b, Export *e, ArrowArray *out,
                         const std::vector<int64_t> &nulls, Buf buffer) {
  const size_t n = b->cols.size();
  e->children.resize(n);
  e->child_ptrs.resize(n);
  e->bufs.resize(n);
  for (size_t c = 0; c < n; c++) {
    e->bufs[c] = {nulls[c] ? buffer(c, 0) : nullptr, buffer(c, 1)};
    ArrowArray &a = e->children[c];
    a = ArrowArray{};
    a.length = b->nrows;
    a.null_count = nulls[c];
    a.n_buffers = 2;
    a.buffers = e->bufs[c].data();
    a.release = release_child_array;
    e->child_ptrs[c] = &a;
  }
  *out = ArrowArray{};
  out->length = b->nrows;
  out->n_buffers = 1;  // a struct array's validity (none)
  out->buffers = e->parent_bufs;
  out->n_children = (int64_t)n;
  out->children = e->child_ptrs.data();
  out->release = release_array;
  out->private_data = e;
}

}  // namespace kg

using namespace kg;

extern "C" {

int kg_column_from_arrow(const struct ArrowArray *array, const struct ArrowSchema *schema,
                         int32_t on_device, kg_column *out) {
  return guarded([&] { view(array, schema, on_device != 0, out); });
}

int kg_column_from_arrow_device(const struct ArrowDeviceArray *array,
                                const struct ArrowSchema *schema, kg_column *out) {
  return guarded([&] {
    bool device = false;
    switch (array->device_type) {
      case ARROW_DEVICE_ROCM: device = true; break;
      case ARROW_DEVICE_CPU:
      case ARROW_DEVICE_ROCM_HOST: break;
      default: fail(KG_ERR_UNSUPPORTED, "an Arrow device array of another device type");
    }
    if (device && array->sync_event)
      KG_HIP(hipEventSynchronize(*static_cast<hipEvent_t *>(array->sync_event)));
    view(&array->array, schema, device, out);
  });
}

int kg_batch_export_device(kg_batch *b, const char *const *names, struct ArrowDeviceArray *out,
                           struct ArrowSchema *schema_out) {
  return guarded([&] {
    hipStream_t s = b->ctx->stream;
    KG_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> nulls(b->cols.size());
    for (size_t c = 0; c < b->cols.size(); c++) nulls[c] = null_count_of(b, c, s);
    export_schema(b, names, schema_out);
    auto e = std::make_unique<Export>();
    e->batch = b;
    *out = ArrowDeviceArray{};
    export_array(b, e.get(), &out->array, nulls, [&](size_t c, int k) -> const void * {
      return k == 0 ? b->cols[c].validity.get() : b->cols[c].values.get();
    });
    e.release();
    out->device_id = b->ctx->device;
    out->device_type = ARROW_DEVICE_ROCM;
    out->sync_event = nullptr;
  });
}

int kg_batch_export_host(const kg_batch *b, const char *const *names, struct ArrowArray *out,
                         struct ArrowSchema *schema_out) {
  return guarded([&] {
    hipStream_t s = b->ctx->stream;
    auto e = std::make_unique<Export>();
    const size_t n = b->cols.size();
    std::vector<int64_t> nulls(n);
    e->host.resize(2 * n);
    for (size_t c = 0; c < n; c++) {
      const auto &col = b->cols[c];
      const int w = type_width(col.type);
      const uint64_t vb = w == 0 ? (uint64_t)ceil_div(b->nrows, 8) : (uint64_t)b->nrows * w;
      e->host[2 * c + 1].resize(vb + 8);
      if (vb) KG_HIP(hipMemcpyAsync(e->host[2 * c + 1].data(), col.values.get(), vb,
                                    hipMemcpyDeviceToHost, s));
      if (col.validity.bytes() && b->nrows) {
        e->host[2 * c].resize((size_t)ceil_div(b->nrows, 8) + 8);
        KG_HIP(hipMemcpyAsync(e->host[2 * c].data(), col.validity.get(),
                              (size_t)ceil_div(b->nrows, 8), hipMemcpyDeviceToHost, s));
      }
    }
    KG_HIP(hipStreamSynchronize(s));
    for (size_t c = 0; c < n; c++) {
      int64_t valid = b->nrows;
      if (!e->host[2 * c].empty()) {
        valid = 0;
        for (int64_t i = 0; i < b->nrows; i++) valid += (e->host[2 * c][i >> 3] >> (i & 7)) & 1;
      }
      nulls[c] = b->nrows - valid;
    }
    export_schema(b, names, schema_out);
    Export *ep = e.release();
    export_array(b, ep, out, nulls, [&](size_t c, int k) -> const void * {
      auto &v = ep->host[2 * c + k];
      return v.empty() ? nullptr : v.data();
    });
  });
}

}  // extern "C"
