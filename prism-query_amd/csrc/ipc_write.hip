// Arrow IPC stream writer for prism-jni's reply.  prism-jni serialises the
// plan's output batches with arrow's StreamWriter (prism-jni/src/lib.rs:111-121):
// the schema message, one record batch message per batch, then the
// end-of-stream marker, and nothing at all when the plan produced no batch.
// Here the same stream is written without arrow, from the Arrow C Data batches
// the operators export:
//   [0xFFFFFFFF] [int32 metadata size] [flatbuffer Message, padded to 8] [body]
// with the body's buffers 8-byte aligned and padded (validity bitmaps only for
// columns with nulls, sliced inputs rebased to offset 0, Utf8 offsets rebased
// to start at 0).  The flatbuffers are built back to front, as the format's
// builders do, by the small builder below.
#include <cstring>
#include <string>
#include <vector>

#include "arrow_io.hpp"

namespace prism {
namespace ipcw {

// Back-to-front flatbuffer builder.  An object is named by its distance from
// the end of the buffer ("off"), which does not change as the buffer grows at
// the front.
class Builder {
 public:
  uint32_t size() const { return static_cast<uint32_t>(b_.size()); }

  // Pad so that, after `n` more bytes, the size is a multiple of `align`.
  void prep(uint32_t align, uint32_t n) {
    if (align > min_align_) min_align_ = align;
    const uint32_t pad = (align - ((size() + n) % align)) % align;
    b_.insert(b_.begin(), pad, 0);
  }
  void push(const void* p, uint32_t n) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    b_.insert(b_.begin(), q, q + n);
  }
  template <class T>
  uint32_t scalar(T v) {
    prep(sizeof(T), sizeof(T));
    push(&v, sizeof(T));
    return size();
  }
  // A uoffset to an object built earlier.
  uint32_t uoffset(uint32_t target) {
    prep(4, 4);
    const uint32_t v = size() + 4 - target;
    push(&v, 4);
    return size();
  }
  uint32_t string(const std::string& s) {
    prep(4, static_cast<uint32_t>(s.size()) + 1 + 4);
    const uint8_t z = 0;
    push(&z, 1);
    push(s.data(), static_cast<uint32_t>(s.size()));
    const uint32_t n = static_cast<uint32_t>(s.size());
    push(&n, 4);
    return size();
  }
  uint32_t offset_vector(const std::vector<uint32_t>& objs) {
    prep(4, 4 * static_cast<uint32_t>(objs.size()) + 4);
    for (size_t i = objs.size(); i-- > 0;) uoffset(objs[i]);
    const uint32_t n = static_cast<uint32_t>(objs.size());
    push(&n, 4);
    return size();
  }
  // Vector of 16-byte structs of two int64 (FieldNode, Buffer), 8-aligned.
  uint32_t pair_vector(const std::vector<std::pair<int64_t, int64_t>>& v) {
    prep(8, 16 * static_cast<uint32_t>(v.size()));
    for (size_t i = v.size(); i-- > 0;) {
      push(&v[i].second, 8);
      push(&v[i].first, 8);
    }
    prep(4, 4);  // the elements start 8-aligned, so this adds nothing
    const uint32_t n = static_cast<uint32_t>(v.size());
    push(&n, 4);
    return size();
  }

  void start_table() {
    fields_.clear();
    table_start_ = size();
  }
  template <class T>
  void add_scalar(int id, T v) {
    fields_.push_back({id, scalar(v)});
  }
  void add_offset(int id, uint32_t target) { fields_.push_back({id, uoffset(target)}); }
  uint32_t end_table() {
    prep(4, 4);
    const int32_t placeholder = 0;
    push(&placeholder, 4);
    const uint32_t table = size();
    int max_id = -1;
    for (auto& f : fields_) max_id = f.first > max_id ? f.first : max_id;
    std::vector<uint16_t> vt(static_cast<size_t>(max_id + 1), 0);
    for (auto& f : fields_) vt[static_cast<size_t>(f.first)] = static_cast<uint16_t>(table - f.second);
    const uint16_t vt_bytes = static_cast<uint16_t>(4 + 2 * vt.size());
    const uint16_t obj_bytes = static_cast<uint16_t>(table - table_start_);
    for (size_t i = vt.size(); i-- > 0;) push(&vt[i], 2);
    push(&obj_bytes, 2);
    push(&vt_bytes, 2);
    const uint32_t vtable = size();
    const int32_t so = static_cast<int32_t>(vtable - table);  // table start - vtable start
    std::memcpy(b_.data() + (size() - table), &so, 4);
    return table;
  }
  std::vector<uint8_t> finish(uint32_t root) {
    prep(min_align_ > 8 ? min_align_ : 8, 4);
    uoffset(root);
    return b_;
  }

 private:
  std::vector<uint8_t> b_;
  std::vector<std::pair<int, uint32_t>> fields_;
  uint32_t table_start_ = 0, min_align_ = 1;
};

// Message.fbs / Schema.fbs tags and field indices written here.
enum : uint8_t { kHdrSchema = 1, kHdrRecordBatch = 3 };
enum : uint8_t { kTInt = 2, kTFloat = 3, kTUtf8 = 5, kTBool = 6, kTDecimal = 7, kTDate = 8 };
constexpr int16_t kMetadataV5 = 4;

static uint32_t type_table(Builder& b, DType t, uint8_t& tag) {
  b.start_table();
  switch (t) {
    case DType::Int32: case DType::Int64: case DType::UInt32:
      tag = kTInt;
      b.add_scalar<int32_t>(0, t == DType::Int64 ? 64 : 32);  // bitWidth
      b.add_scalar<uint8_t>(1, t != DType::UInt32);           // is_signed
      break;
    case DType::Float64:
      tag = kTFloat;
      b.add_scalar<int16_t>(0, 2);  // DOUBLE
      break;
    case DType::Utf8: tag = kTUtf8; break;
    case DType::Bool: tag = kTBool; break;
    case DType::Date32:
      tag = kTDate;
      b.add_scalar<int16_t>(0, 0);  // DAY (the field's default is MILLISECOND)
      break;
    case DType::Decimal128:
      tag = kTDecimal;
      b.add_scalar<int32_t>(0, 38);   // precision
      b.add_scalar<int32_t>(1, 0);    // scale
      b.add_scalar<int32_t>(2, 128);  // bitWidth
      break;
  }
  return b.end_table();
}

static std::vector<uint8_t> schema_message(const std::vector<HostColumn>& cols) {
  Builder b;
  std::vector<uint32_t> fields;
  for (const HostColumn& c : cols) {
    const uint32_t name = b.string(c.name);
    const uint32_t children = b.offset_vector({});
    uint8_t tag = 0;
    const uint32_t type = type_table(b, c.type, tag);
    b.start_table();
    b.add_offset(0, name);
    b.add_scalar<uint8_t>(1, c.nullable ? 1 : 0);
    b.add_scalar<uint8_t>(2, tag);
    b.add_offset(3, type);
    b.add_offset(5, children);
    fields.push_back(b.end_table());
  }
  const uint32_t fv = b.offset_vector(fields);
  b.start_table();
  b.add_scalar<int16_t>(0, 0);  // little-endian
  b.add_offset(1, fv);
  const uint32_t schema = b.end_table();
  b.start_table();
  b.add_scalar<int16_t>(0, kMetadataV5);
  b.add_scalar<uint8_t>(1, kHdrSchema);
  b.add_offset(2, schema);
  b.add_scalar<int64_t>(3, 0);
  return b.finish(b.end_table());
}

// Bits [from, from + n) of a bitmap as a bitmap starting at bit 0.
static void append_bits(std::vector<uint8_t>& out, const uint8_t* bits, int64_t from, int64_t n) {
  const size_t at = out.size();
  out.resize(at + static_cast<size_t>((n + 7) / 8), 0);
  for (int64_t i = 0; i < n; ++i)
    if ((bits[(from + i) >> 3] >> ((from + i) & 7)) & 1) out[at + static_cast<size_t>(i >> 3)] |= 1u << (i & 7);
}

struct Body {
  std::vector<uint8_t> bytes;
  std::vector<std::pair<int64_t, int64_t>> nodes, buffers;
  // Start a buffer: pad the body to 8 bytes.
  void open() { bytes.resize((bytes.size() + 7) & ~static_cast<size_t>(7), 0); }
  void close(size_t start) { buffers.push_back({static_cast<int64_t>(start), static_cast<int64_t>(bytes.size() - start)}); }
  void raw(const void* p, size_t n) {
    open();
    const size_t s = bytes.size();
    bytes.insert(bytes.end(), static_cast<const uint8_t*>(p), static_cast<const uint8_t*>(p) + n);
    close(s);
  }
};

static int64_t count_nulls(const HostColumn& c) {
  if (!c.validity) return 0;
  if (c.null_count >= 0) return c.null_count;
  int64_t k = 0;
  for (int64_t i = 0; i < c.length; ++i) k += !((c.validity[(c.offset + i) >> 3] >> ((c.offset + i) & 7)) & 1);
  return k;
}

static std::vector<uint8_t> batch_message(const HostBatch& hb, std::vector<uint8_t>& body_out) {
  Body B;
  for (const HostColumn& c : hb.cols) {
    const int64_t n = c.length, nulls = count_nulls(c);
    B.nodes.push_back({n, nulls});
    B.open();
    size_t s = B.bytes.size();
    if (nulls) append_bits(B.bytes, c.validity, c.offset, n);
    B.close(s);
    switch (c.type) {
      case DType::Bool:
        B.open();
        s = B.bytes.size();
        append_bits(B.bytes, static_cast<const uint8_t*>(c.data), c.offset, n);
        B.close(s);
        break;
      case DType::Utf8: {
        std::vector<int32_t> o(static_cast<size_t>(n) + 1);
        const int32_t base = c.offsets[c.offset];
        for (int64_t i = 0; i <= n; ++i) o[static_cast<size_t>(i)] = c.offsets[c.offset + i] - base;
        B.raw(o.data(), o.size() * 4);
        B.raw(static_cast<const uint8_t*>(c.data) + base, static_cast<size_t>(o.back()));
        break;
      }
      default: {
        const size_t w = static_cast<size_t>(dtype_width(c.type));
        B.raw(static_cast<const uint8_t*>(c.data) + static_cast<size_t>(c.offset) * w, static_cast<size_t>(n) * w);
      }
    }
  }
  B.open();
  Builder b;
  const uint32_t bufs = b.pair_vector(B.buffers);
  const uint32_t nodes = b.pair_vector(B.nodes);
  b.start_table();
  b.add_scalar<int64_t>(0, hb.n_rows);
  b.add_offset(1, nodes);
  b.add_offset(2, bufs);
  const uint32_t rb = b.end_table();
  b.start_table();
  b.add_scalar<int16_t>(0, kMetadataV5);
  b.add_scalar<uint8_t>(1, kHdrRecordBatch);
  b.add_offset(2, rb);
  b.add_scalar<int64_t>(3, static_cast<int64_t>(B.bytes.size()));
  body_out = std::move(B.bytes);
  return b.finish(b.end_table());
}

static void put_message(std::vector<uint8_t>& out, const std::vector<uint8_t>& meta, const std::vector<uint8_t>& body) {
  const uint32_t cont = 0xFFFFFFFFu;
  const uint32_t len = static_cast<uint32_t>((meta.size() + 7) & ~static_cast<size_t>(7));
  out.insert(out.end(), reinterpret_cast<const uint8_t*>(&cont), reinterpret_cast<const uint8_t*>(&cont) + 4);
  out.insert(out.end(), reinterpret_cast<const uint8_t*>(&len), reinterpret_cast<const uint8_t*>(&len) + 4);
  out.insert(out.end(), meta.begin(), meta.end());
  out.resize(out.size() + (len - meta.size()), 0);
  out.insert(out.end(), body.begin(), body.end());
}

}  // namespace ipcw
}  // namespace prism

using namespace prism;

extern "C" {

prism_status prism_ipc_write(const struct ArrowArray* batches, uint32_t n_batches, const struct ArrowSchema* schema,
                             uint8_t** out_bytes, uint64_t* out_len) {
  return guarded([&] {
    if (!out_bytes || !out_len || (n_batches && (!batches || !schema)))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL batches/schema/output");
    *out_bytes = nullptr;
    *out_len = 0;
    if (n_batches == 0) return;  // the reference writes no stream for an empty result
    std::vector<HostBatch> hb;
    for (uint32_t i = 0; i < n_batches; ++i) hb.push_back(import_batch(&batches[i], schema, true));
    std::vector<uint8_t> out;
    ipcw::put_message(out, ipcw::schema_message(hb[0].cols), {});
    for (const HostBatch& b : hb) {
      std::vector<uint8_t> body;
      const std::vector<uint8_t> meta = ipcw::batch_message(b, body);
      ipcw::put_message(out, meta, body);
    }
    const uint32_t eos[2] = {0xFFFFFFFFu, 0u};
    out.insert(out.end(), reinterpret_cast<const uint8_t*>(eos), reinterpret_cast<const uint8_t*>(eos) + 8);
    uint8_t* p = static_cast<uint8_t*>(malloc(out.size()));
    if (!p) fail(PRISM_ERR_RESOURCE, "Resource error: no host memory for the IPC stream");
    std::memcpy(p, out.data(), out.size());
    *out_bytes = p;
    *out_len = out.size();
  });
}

void prism_ipc_buffer_free(uint8_t* bytes) { free(bytes); }

}  // extern "C"
