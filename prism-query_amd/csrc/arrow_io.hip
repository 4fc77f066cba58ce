// Arrow C Data Interface import/export and host<->HBM column transfer.
#include "arrow_io.hpp"

#include <cstdlib>

namespace prism {

const char* dtype_format(DType t) {
  switch (t) {
    case DType::Int32: return "i";
    case DType::Int64: return "l";
    case DType::Float64: return "g";
    case DType::Utf8: return "u";
    case DType::Bool: return "b";
    case DType::Date32: return "tdD";
    case DType::UInt32: return "I";
    case DType::Decimal128: return "d:38,0";
  }
  return "n";
}

std::string dtype_name(DType t) {
  switch (t) {
    case DType::Int32: return "Int32";
    case DType::Int64: return "Int64";
    case DType::Float64: return "Float64";
    case DType::Utf8: return "Utf8";
    case DType::Bool: return "Boolean";
    case DType::Date32: return "Date32";
    case DType::UInt32: return "UInt32";
    case DType::Decimal128: return "Decimal128(38, 0)";
  }
  return "?";
}

int dtype_width(DType t) {
  switch (t) {
    case DType::Int32: case DType::Date32: case DType::UInt32: return 4;
    case DType::Int64: case DType::Float64: return 8;
    case DType::Decimal128: return 16;
    default: return 0;
  }
}

// `outputs`: also the types only operators produce (Decimal128(38,0) sums),
// for writers of operator results.
static DType parse_format(const char* f, bool outputs = false) {
  std::string s = f ? f : "";
  if (outputs && s == "d:38,0") return DType::Decimal128;
  if (s == "i") return DType::Int32;
  if (s == "l") return DType::Int64;
  if (s == "g") return DType::Float64;
  if (s == "u") return DType::Utf8;
  if (s == "b") return DType::Bool;
  if (s == "tdD") return DType::Date32;
  if (s == "I") return DType::UInt32;
  fail(PRISM_ERR_ARROW, "Arrow error: unsupported Arrow type (format '" + s + "')");
}

static HostColumn import_child(const ArrowArray* ca, const ArrowSchema* cs, int64_t length, int64_t parent_off,
                               bool outputs) {
  HostColumn c;
  c.type = parse_format(cs->format, outputs);
  c.name = cs->name ? cs->name : "";
  c.nullable = (cs->flags & ARROW_FLAG_NULLABLE) != 0;
  c.length = length;
  c.offset = ca->offset + parent_off;
  c.null_count = ca->null_count;
  const int64_t need = c.type == DType::Utf8 ? 3 : 2;
  if (ca->n_buffers < need) fail(PRISM_ERR_ARROW, "Arrow error: column '" + c.name + "' has too few buffers");
  c.validity = static_cast<const uint8_t*>(ca->buffers[0]);
  if (c.null_count == 0) c.validity = nullptr;
  if (c.type == DType::Utf8) {
    c.offsets = static_cast<const int32_t*>(ca->buffers[1]);
    c.data = ca->buffers[2];
    if (length && !c.offsets) fail(PRISM_ERR_ARROW, "Arrow error: Utf8 column without offsets");
  } else {
    c.data = ca->buffers[1];
    if (length && !c.data) fail(PRISM_ERR_ARROW, "Arrow error: column '" + c.name + "' without data buffer");
  }
  return c;
}

HostBatch import_batch(const ArrowArray* a, const ArrowSchema* s, bool outputs) {
  if (!a || !s) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL ArrowArray/ArrowSchema");
  if (!s->format || std::string(s->format) != "+s")
    fail(PRISM_ERR_ARROW, "Arrow error: expected a RecordBatch exported as a struct array");
  if (a->n_children != s->n_children) fail(PRISM_ERR_ARROW, "Arrow error: array/schema child count mismatch");
  HostBatch b;
  b.n_rows = a->length;
  for (int64_t i = 0; i < a->n_children; ++i) b.cols.push_back(import_child(a->children[i], s->children[i], a->length, a->offset, outputs));
  return b;
}

std::vector<HostColumn> import_schema(const ArrowSchema* s) {
  if (!s || !s->format || std::string(s->format) != "+s")
    fail(PRISM_ERR_ARROW, "Arrow error: expected a struct schema");
  std::vector<HostColumn> cols;
  for (int64_t i = 0; i < s->n_children; ++i) {
    HostColumn c;
    c.type = parse_format(s->children[i]->format);
    c.name = s->children[i]->name ? s->children[i]->name : "";
    c.nullable = (s->children[i]->flags & ARROW_FLAG_NULLABLE) != 0;
    cols.push_back(c);
  }
  return cols;
}

prism_column DevColumn::desc() const {
  prism_column c;
  memset(&c, 0, sizeof(c));
  c.type = static_cast<int32_t>(type == DType::UInt32 ? DType::Int32 : type);
  c.offset = dev_offset;
  c.data = data.ptr;
  c.offsets = offsets.as<int32_t>();
  c.validity = validity.as<uint8_t>();
  return c;
}

// Host -> HBM through two pinned staging halves, so the memcpy into one half
// overlaps the DMA out of the other.
void staged_h2d(prism_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  constexpr size_t kHalf = 8u << 20;
  std::lock_guard<std::mutex> lock(ctx->staging_mu);
  char* pin = static_cast<char*>(pinned_staging(ctx, 2 * kHalf));
  static thread_local hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev)
    if (!e) PRISM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  bool used[2] = {false, false};
  size_t done = 0;
  int k = 0;
  while (done < bytes) {
    size_t n = bytes - done < kHalf ? bytes - done : kHalf;
    if (used[k]) PRISM_HIP(hipEventSynchronize(ev[k]));
    memcpy(pin + k * kHalf, static_cast<const char*>(src) + done, n);
    PRISM_HIP(hipMemcpyAsync(static_cast<char*>(dst) + done, pin + k * kHalf, n, hipMemcpyHostToDevice, ctx->stream));
    PRISM_HIP(hipEventRecord(ev[k], ctx->stream));
    used[k] = true;
    done += n;
    k ^= 1;
  }
  for (int i = 0; i < 2; ++i)
    if (used[i]) PRISM_HIP(hipEventSynchronize(ev[i]));
}

DevColumn upload(prism_ctx* ctx, const HostColumn& c) {
  Ctx* cx = as_ctx(ctx);
  DevColumn d;
  d.type = c.type;
  d.name = c.name;
  d.nullable = c.nullable;
  d.length = c.length;
  const int64_t start = c.offset & ~7LL;
  d.dev_offset = c.offset - start;
  const int64_t n = d.dev_offset + c.length;
  const int w = dtype_width(c.type);
  if (c.type == DType::Utf8) {
    std::vector<int32_t> o(static_cast<size_t>(n + 1));
    const int32_t base = c.offsets ? c.offsets[start] : 0;
    for (int64_t i = 0; i <= n; ++i) {
      const int64_t v = c.offsets ? static_cast<int64_t>(c.offsets[start + i]) - base : 0;
      if (i && (v < o[i - 1] || v > INT32_MAX))
        fail(PRISM_ERR_ARROW, "Arrow error: Utf8 offsets decrease at row " + std::to_string(start + i - 1));
      o[i] = static_cast<int32_t>(v);
    }
    d.offsets.alloc(cx, o.size() * 4);
    staged_h2d(ctx, d.offsets.ptr, o.data(), o.size() * 4);
    const size_t nb = static_cast<size_t>(o[n]);
    d.data.alloc(cx, nb ? nb : 1);
    if (nb) staged_h2d(ctx, d.data.ptr, static_cast<const uint8_t*>(c.data) + base, nb);
  } else if (c.type == DType::Bool) {
    const size_t nb = static_cast<size_t>((n + 7) / 8);
    d.data.alloc(cx, nb ? nb : 1);
    if (nb) staged_h2d(ctx, d.data.ptr, static_cast<const uint8_t*>(c.data) + start / 8, nb);
  } else {
    const size_t nb = static_cast<size_t>(n) * w;
    d.data.alloc(cx, nb ? nb : 8);
    if (nb) staged_h2d(ctx, d.data.ptr, static_cast<const uint8_t*>(c.data) + start * w, nb);
  }
  if (c.validity) {
    const size_t nb = static_cast<size_t>((n + 7) / 8);
    d.validity.alloc(cx, nb ? nb : 1);
    if (nb) staged_h2d(ctx, d.validity.ptr, c.validity + start / 8, nb);
    d.null_count = c.null_count;
  }
  return d;
}

OutColumn download(prism_ctx* ctx, const DevColumn& c) {
  if (c.dev_offset != 0) fail(PRISM_ERR_INTERNAL, "Internal error: download of an offset device column");
  OutColumn o;
  o.type = c.type;
  o.name = c.name;
  o.nullable = c.nullable;
  o.length = c.length;
  const int64_t n = c.length;
  hipStream_t st = ctx->stream;
  if (c.type == DType::Utf8) {
    o.offsets.resize(static_cast<size_t>(n + 1));
    PRISM_HIP(hipMemcpyAsync(o.offsets.data(), c.offsets.ptr, (n + 1) * 4, hipMemcpyDeviceToHost, st));
    PRISM_HIP(hipStreamSynchronize(st));
    o.data.resize(static_cast<size_t>(o.offsets[n]));
    if (!o.data.empty())
      PRISM_HIP(hipMemcpyAsync(o.data.data(), c.data.ptr, o.data.size(), hipMemcpyDeviceToHost, st));
  } else if (c.type == DType::Bool) {
    o.data.resize(static_cast<size_t>((n + 7) / 8));
    if (!o.data.empty()) PRISM_HIP(hipMemcpyAsync(o.data.data(), c.data.ptr, o.data.size(), hipMemcpyDeviceToHost, st));
  } else {
    o.data.resize(static_cast<size_t>(n) * dtype_width(c.type));
    if (!o.data.empty()) PRISM_HIP(hipMemcpyAsync(o.data.data(), c.data.ptr, o.data.size(), hipMemcpyDeviceToHost, st));
  }
  if (c.validity.ptr) {
    o.validity.resize(static_cast<size_t>((n + 7) / 8));
    if (!o.validity.empty())
      PRISM_HIP(hipMemcpyAsync(o.validity.data(), c.validity.ptr, o.validity.size(), hipMemcpyDeviceToHost, st));
  }
  PRISM_HIP(hipStreamSynchronize(st));
  if (!o.validity.empty()) {
    int64_t valid = 0;
    for (int64_t i = 0; i < n; ++i) valid += (o.validity[i >> 3] >> (i & 7)) & 1;
    o.null_count = n - valid;
    if (o.null_count == 0) o.validity.clear();
  }
  return o;
}

// ---------------------------------------------------------------------------
// Export.
// ---------------------------------------------------------------------------
namespace {

struct ColHolder {
  OutColumn col;
  const void* buffers[3];
};

struct StructHolder {
  std::vector<ArrowArray*> children;
  const void* buffers[1] = {nullptr};
};

struct SchemaHolder {
  std::string format, name;
  std::vector<ArrowSchema*> children;
};

const uint8_t kEmpty[16] = {0};

void release_col(ArrowArray* a) {
  delete static_cast<ColHolder*>(a->private_data);
  a->release = nullptr;
}

void release_struct(ArrowArray* a) {
  auto* h = static_cast<StructHolder*>(a->private_data);
  for (ArrowArray* c : h->children) {
    if (c->release) c->release(c);
    delete c;
  }
  delete h;
  a->release = nullptr;
}

void release_schema(ArrowSchema* s) {
  auto* h = static_cast<SchemaHolder*>(s->private_data);
  for (ArrowSchema* c : h->children) {
    if (c->release) c->release(c);
    delete c;
  }
  delete h;
  s->release = nullptr;
}

void fill_schema(ArrowSchema* s, const std::string& format, const std::string& name, bool nullable) {
  auto* h = new SchemaHolder();
  h->format = format;
  h->name = name;
  s->format = h->format.c_str();
  s->name = h->name.c_str();
  s->metadata = nullptr;
  s->flags = nullable ? ARROW_FLAG_NULLABLE : 0;
  s->n_children = 0;
  s->children = nullptr;
  s->dictionary = nullptr;
  s->release = release_schema;
  s->private_data = h;
}

void fill_array(ArrowArray* a, OutColumn&& col) {
  auto* h = new ColHolder();
  h->col = std::move(col);
  OutColumn& c = h->col;
  a->length = c.length;
  a->null_count = c.validity.empty() ? 0 : c.null_count;
  a->offset = 0;
  a->n_children = 0;
  a->children = nullptr;
  a->dictionary = nullptr;
  h->buffers[0] = c.validity.empty() ? nullptr : c.validity.data();
  if (c.type == DType::Utf8) {
    if (c.offsets.empty()) c.offsets.push_back(0);
    h->buffers[1] = c.offsets.data();
    h->buffers[2] = c.data.empty() ? static_cast<const void*>(kEmpty) : c.data.data();
    a->n_buffers = 3;
  } else {
    h->buffers[1] = c.data.empty() ? static_cast<const void*>(kEmpty) : c.data.data();
    a->n_buffers = 2;
  }
  a->buffers = h->buffers;
  a->release = release_col;
  a->private_data = h;
}

std::string format_of(const OutColumn& c) {
  if (c.type == DType::Decimal128)
    return "d:" + std::to_string(c.precision) + "," + std::to_string(c.scale);
  return dtype_format(c.type);
}

}  // namespace

void export_array(OutColumn&& col, ArrowArray* out, ArrowSchema* out_schema) {
  if (!out || !out_schema) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL output array/schema");
  fill_schema(out_schema, format_of(col), col.name, col.nullable);
  fill_array(out, std::move(col));
}

void export_batch(std::vector<OutColumn>&& cols, int64_t n_rows, ArrowArray* out, ArrowSchema* out_schema) {
  if (!out || !out_schema) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL output array/schema");
  const size_t k = cols.size();
  // schema
  fill_schema(out_schema, "+s", "", false);
  auto* sh = static_cast<SchemaHolder*>(out_schema->private_data);
  for (size_t i = 0; i < k; ++i) {
    auto* cs = new ArrowSchema();
    fill_schema(cs, format_of(cols[i]), cols[i].name, cols[i].nullable);
    sh->children.push_back(cs);
  }
  out_schema->n_children = static_cast<int64_t>(k);
  out_schema->children = sh->children.data();
  // array
  auto* h = new StructHolder();
  for (size_t i = 0; i < k; ++i) {
    auto* ca = new ArrowArray();
    fill_array(ca, std::move(cols[i]));
    h->children.push_back(ca);
  }
  out->length = n_rows;
  out->null_count = 0;
  out->offset = 0;
  out->n_buffers = 1;
  out->buffers = h->buffers;
  out->n_children = static_cast<int64_t>(k);
  out->children = h->children.data();
  out->dictionary = nullptr;
  out->release = release_struct;
  out->private_data = h;
}

}  // namespace prism
