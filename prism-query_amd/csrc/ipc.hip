// Arrow IPC stream ingest.  prism-jni receives every plan's input as an Arrow
// IPC stream in a JNI byte array and decodes it with arrow's StreamReader
// (prism-jni/src/lib.rs:96-121) before the operators copy it again.  Here the
// stream is decoded without arrow: a stream is a sequence of encapsulated
// messages
//   [0xFFFFFFFF] [int32 metadata size] [flatbuffer Message, padded to 8] [body]
// ending with a zero metadata size (the 0xFFFFFFFF continuation is optional,
// as in the pre-1.0 format).  Only the metadata this path needs is read: the
// Schema's fields (name, nullability, type) and each RecordBatch's length,
// field nodes and buffers.  Dictionaries, body compression, big-endian
// streams and nested types are rejected with an Arrow error.
//
// Two outputs: Arrow C Data struct arrays whose buffers point into the
// caller's bytes (no copy), and device columns whose buffers are copied once
// from the bytes into HBM through pinned staging.
#include <cstring>
#include <string>
#include <vector>

#include "arrow_io.hpp"

namespace prism {
namespace ipc {

[[noreturn]] static void bad(const std::string& m) { fail(PRISM_ERR_ARROW, "Arrow error: IPC stream " + m); }

// Little-endian flatbuffer reader over one message's metadata.
struct Fb {
  const uint8_t* base;
  size_t size;
  void need(size_t at, size_t n) const {
    if (at > size || n > size - at) bad("metadata is truncated");
  }
  template <class T>
  T get(size_t at) const {
    need(at, sizeof(T));
    T v;
    memcpy(&v, base + at, sizeof(T));
    return v;
  }
};

// A flatbuffer table: a signed offset to its vtable, then the fields.
struct Table {
  const Fb* fb = nullptr;
  size_t pos = 0, vt = 0;
  uint16_t vt_size = 0;
  Table() = default;
  Table(const Fb* f, size_t p) : fb(f), pos(p) {
    const int64_t v = static_cast<int64_t>(p) - fb->get<int32_t>(p);
    if (v < 0) bad("has a bad vtable offset");
    vt = static_cast<size_t>(v);
    vt_size = fb->get<uint16_t>(vt);
  }
  size_t field(int i) const {  // position of field i, 0 when absent
    const size_t o = 4 + 2 * static_cast<size_t>(i);
    if (o + 2 > vt_size) return 0;
    const uint16_t off = fb->get<uint16_t>(vt + o);
    return off ? pos + off : 0;
  }
  template <class T>
  T scalar(int i, T dflt) const {
    const size_t f = field(i);
    return f ? fb->get<T>(f) : dflt;
  }
  size_t ref(int i) const {  // target of an offset field, 0 when absent
    const size_t f = field(i);
    return f ? f + fb->get<uint32_t>(f) : 0;
  }
  bool has(int i) const { return field(i) != 0; }
  Table table(int i) const {
    const size_t r = ref(i);
    if (!r) bad("misses a required table");
    return Table(fb, r);
  }
  // vector: position of element 0 and element count
  std::pair<size_t, uint32_t> vec(int i) const {
    const size_t r = ref(i);
    if (!r) return {0, 0};
    return {r + 4, fb->get<uint32_t>(r)};
  }
  std::string str(int i) const {
    const size_t r = ref(i);
    if (!r) return std::string();
    const uint32_t n = fb->get<uint32_t>(r);
    fb->need(r + 4, n);
    return std::string(reinterpret_cast<const char*>(fb->base + r + 4), n);
  }
};

// Message.fbs / Schema.fbs field indices and union tags used here.
enum : uint8_t { kHdrSchema = 1, kHdrDictionary = 2, kHdrRecordBatch = 3 };
enum : uint8_t { kTInt = 2, kTFloat = 3, kTUtf8 = 5, kTBool = 6, kTDate = 8 };

struct FieldInfo {
  std::string name;
  bool nullable = true;
  DType type = DType::Int64;
};

struct Batch {
  int64_t n_rows = 0;
  std::vector<HostColumn> cols;  // pointers into the caller's bytes
};

struct Stream {
  std::vector<FieldInfo> fields;
  std::vector<Batch> batches;
};

static DType field_type(const Table& f) {
  const uint8_t tt = f.scalar<uint8_t>(2, 0);
  if (f.has(4)) bad("uses dictionary encoding, which is not supported");
  if (f.vec(5).second) bad("has nested fields, which are not supported");
  switch (tt) {
    case kTInt: {
      const Table t = f.table(3);
      const int32_t bits = t.scalar<int32_t>(0, 0);
      const bool sgn = t.scalar<uint8_t>(1, 0) != 0;
      if (sgn && bits == 32) return DType::Int32;
      if (sgn && bits == 64) return DType::Int64;
      bad("has an integer type of " + std::to_string(bits) + (sgn ? " signed" : " unsigned") +
          " bits, which is not supported");
    }
    case kTFloat:
      if (f.table(3).scalar<int16_t>(0, 0) != 2) bad("has a float type other than Float64");
      return DType::Float64;
    case kTUtf8: return DType::Utf8;
    case kTBool: return DType::Bool;
    case kTDate:
      if (f.table(3).scalar<int16_t>(0, 1) != 0) bad("has a Date64 field, which is not supported");
      return DType::Date32;
    default: bad("has a field type (tag " + std::to_string(tt) + ") that is not supported");
  }
}

static void read_schema(const Table& s, Stream& out) {
  if (s.scalar<int16_t>(0, 0) != 0) bad("is big-endian");
  const auto fv = s.vec(1);
  for (uint32_t i = 0; i < fv.second; ++i) {
    const size_t e = fv.first + 4 * static_cast<size_t>(i);
    const Table f(s.fb, e + s.fb->get<uint32_t>(e));
    FieldInfo fi;
    fi.name = f.str(0);
    fi.nullable = f.scalar<uint8_t>(1, 0) != 0;
    fi.type = field_type(f);
    out.fields.push_back(fi);
  }
}

static void read_batch(const Table& rb, const uint8_t* body, uint64_t body_len, Stream& out) {
  if (rb.has(3)) bad("uses body compression, which is not supported");
  Batch b;
  b.n_rows = rb.scalar<int64_t>(0, 0);
  if (b.n_rows < 0) bad("has a negative batch length");
  const auto nodes = rb.vec(1), bufs = rb.vec(2);
  if (nodes.second != out.fields.size()) bad("batch does not match its schema (field nodes)");
  const Fb* fb = rb.fb;
  uint32_t bi = 0;
  auto buffer = [&](const void** p, int64_t* len) {
    if (bi >= bufs.second) bad("batch has too few buffers");
    const size_t at = bufs.first + 16 * static_cast<size_t>(bi++);
    const int64_t off = fb->get<int64_t>(at), n = fb->get<int64_t>(at + 8);
    if (off < 0 || n < 0 || static_cast<uint64_t>(off) > body_len || static_cast<uint64_t>(n) > body_len - off)
      bad("buffer lies outside the message body");
    *p = body + off;
    *len = n;
  };
  for (size_t k = 0; k < out.fields.size(); ++k) {
    const size_t at = nodes.first + 16 * k;
    const int64_t len = fb->get<int64_t>(at), nulls = fb->get<int64_t>(at + 8);
    if (len != b.n_rows || nulls < 0 || nulls > len) bad("field node does not match the batch length");
    HostColumn c;
    c.type = out.fields[k].type;
    c.name = out.fields[k].name;
    c.nullable = out.fields[k].nullable;
    c.length = len;
    c.offset = 0;
    c.null_count = nulls;
    const void* vp;
    int64_t vlen;
    buffer(&vp, &vlen);
    // Lengths are untrusted: every size test divides the buffer length
    // instead of multiplying the row count, so nothing can wrap.
    auto bitmap_short = [len](int64_t blen) { return blen < len / 8 + (len % 8 != 0); };
    if (nulls > 0) {
      if (bitmap_short(vlen)) bad("validity buffer is too short");
      c.validity = static_cast<const uint8_t*>(vp);
    }
    const void* dp;
    int64_t dlen;
    if (c.type == DType::Utf8) {
      buffer(&dp, &dlen);
      if (dlen < 4 || dlen / 4 - 1 < len) bad("offsets buffer is too short");
      c.offsets = static_cast<const int32_t*>(dp);
      const void* sp;
      int64_t slen;
      buffer(&sp, &slen);
      // every offset in [0, data length] and non-decreasing, so each string
      // lies inside the data buffer
      int32_t prev;
      memcpy(&prev, c.offsets, 4);
      if (prev < 0) bad("string offsets lie outside the data buffer");
      for (int64_t i = 1; i <= len; ++i) {
        int32_t cur;
        memcpy(&cur, c.offsets + i, 4);
        if (cur < prev) bad("string offsets decrease at row " + std::to_string(i - 1));
        prev = cur;
      }
      if (prev > slen) bad("string offsets lie outside the data buffer");
      c.data = sp;
    } else {
      buffer(&dp, &dlen);
      const bool short_values = c.type == DType::Bool ? bitmap_short(dlen) : dlen / dtype_width(c.type) < len;
      if (short_values) bad("values buffer is too short");
      c.data = dp;
    }
    b.cols.push_back(c);
  }
  out.batches.push_back(std::move(b));
}

static Stream parse(const uint8_t* bytes, uint64_t n) {
  if (!bytes && n) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL IPC bytes");
  Stream out;
  bool have_schema = false;
  uint64_t p = 0;
  while (p + 4 <= n) {
    uint32_t w;
    memcpy(&w, bytes + p, 4);
    int32_t meta;
    if (w == 0xFFFFFFFFu) {
      if (p + 8 > n) bad("ends inside a message prefix");
      memcpy(&meta, bytes + p + 4, 4);
      p += 8;
    } else {
      meta = static_cast<int32_t>(w);
      p += 4;
    }
    if (meta == 0) break;  // end of stream
    if (meta < 0 || static_cast<uint64_t>(meta) > n - p) bad("message metadata runs past the end");
    const Fb fb{bytes + p, static_cast<size_t>(meta)};
    const Table msg(&fb, fb.get<uint32_t>(0));
    const uint8_t ht = msg.scalar<uint8_t>(1, 0);
    const int64_t body_len = msg.scalar<int64_t>(3, 0);
    p += static_cast<uint64_t>(meta);
    if (body_len < 0 || static_cast<uint64_t>(body_len) > n - p) bad("message body runs past the end");
    const uint8_t* body = bytes + p;
    if (ht == kHdrSchema) {
      if (have_schema) bad("has a second schema");
      read_schema(msg.table(2), out);
      have_schema = true;
    } else if (ht == kHdrRecordBatch) {
      if (!have_schema) bad("has a record batch before its schema");
      read_batch(msg.table(2), body, static_cast<uint64_t>(body_len), out);
    } else if (ht == kHdrDictionary) {
      bad("has a dictionary batch, which is not supported");
    } else {
      bad("has a message of type " + std::to_string(ht) + ", which is not supported");
    }
    p += static_cast<uint64_t>(body_len);
  }
  if (!have_schema) bad("has no schema");
  return out;
}

// ---- Arrow C Data export of the parsed stream (buffers stay in the bytes) ----
struct ColHold {
  const void* buffers[3];
};
struct StructHold {
  std::vector<ArrowArray*> children;
  const void* buffers[1] = {nullptr};
};
struct SchemaHold {
  std::string format, name;
  std::vector<ArrowSchema*> children;
};

static void rel_col(ArrowArray* a) {
  delete static_cast<ColHold*>(a->private_data);
  a->release = nullptr;
}
static void rel_struct(ArrowArray* a) {
  auto* h = static_cast<StructHold*>(a->private_data);
  for (ArrowArray* c : h->children) {
    if (c->release) c->release(c);
    delete c;
  }
  delete h;
  a->release = nullptr;
}
static void rel_schema(ArrowSchema* s) {
  auto* h = static_cast<SchemaHold*>(s->private_data);
  for (ArrowSchema* c : h->children) {
    if (c->release) c->release(c);
    delete c;
  }
  delete h;
  s->release = nullptr;
}

static void fill_schema(ArrowSchema* s, const std::string& format, const std::string& name, bool nullable) {
  auto* h = new SchemaHold();
  h->format = format;
  h->name = name;
  s->format = h->format.c_str();
  s->name = h->name.c_str();
  s->metadata = nullptr;
  s->flags = nullable ? ARROW_FLAG_NULLABLE : 0;
  s->n_children = 0;
  s->children = nullptr;
  s->dictionary = nullptr;
  s->release = rel_schema;
  s->private_data = h;
}

static void export_schema(const Stream& st, ArrowSchema* out) {
  fill_schema(out, "+s", "", false);
  auto* sh = static_cast<SchemaHold*>(out->private_data);
  for (const FieldInfo& f : st.fields) {
    auto* cs = new ArrowSchema();
    fill_schema(cs, dtype_format(f.type), f.name, f.nullable);
    sh->children.push_back(cs);
  }
  out->n_children = static_cast<int64_t>(st.fields.size());
  out->children = sh->children.data();
}

static const int32_t kZero[1] = {0};

static void export_batch_view(const Batch& b, ArrowArray* out) {
  auto* h = new StructHold();
  for (const HostColumn& c : b.cols) {
    auto* a = new ArrowArray();
    auto* ch = new ColHold();
    a->length = c.length;
    a->null_count = c.validity ? c.null_count : 0;
    a->offset = 0;
    a->n_children = 0;
    a->children = nullptr;
    a->dictionary = nullptr;
    ch->buffers[0] = c.validity;
    if (c.type == DType::Utf8) {
      ch->buffers[1] = c.offsets ? static_cast<const void*>(c.offsets) : kZero;
      ch->buffers[2] = c.data;
      a->n_buffers = 3;
    } else {
      ch->buffers[1] = c.data;
      a->n_buffers = 2;
    }
    a->buffers = ch->buffers;
    a->release = rel_col;
    a->private_data = ch;
    h->children.push_back(a);
  }
  out->length = b.n_rows;
  out->null_count = 0;
  out->offset = 0;
  out->n_buffers = 1;
  out->buffers = h->buffers;
  out->n_children = static_cast<int64_t>(b.cols.size());
  out->children = h->children.data();
  out->dictionary = nullptr;
  out->release = rel_struct;
  out->private_data = h;
}

}  // namespace ipc
}  // namespace prism

// Decoded stream resident in HBM.
struct prism_ipc_stream {
  prism::Ctx* ctx = nullptr;
  std::vector<prism::ipc::FieldInfo> fields;
  std::vector<int64_t> rows;
  std::vector<std::vector<prism::DevColumn>> cols;  // per batch, per field
};

using namespace prism;

extern "C" {

prism_status prism_ipc_read_host(const uint8_t* bytes, uint64_t n_bytes, struct ArrowArray* out_batches,
                                 uint32_t capacity, uint32_t* n_batches, struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!n_batches || !out_schema || (capacity && !out_batches))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL output");
    const ipc::Stream st = ipc::parse(bytes, n_bytes);
    if (st.batches.size() > capacity)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: the stream holds " + std::to_string(st.batches.size()) +
                                           " batches, the output room for " + std::to_string(capacity));
    ipc::export_schema(st, out_schema);
    for (size_t i = 0; i < st.batches.size(); ++i) ipc::export_batch_view(st.batches[i], &out_batches[i]);
    *n_batches = static_cast<uint32_t>(st.batches.size());
  });
}

prism_status prism_ipc_read_device(prism_ctx* ctx, const uint8_t* bytes, uint64_t n_bytes, prism_ipc_stream** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    const ipc::Stream st = ipc::parse(bytes, n_bytes);
    use_device(ctx);
    auto S = std::make_unique<prism_ipc_stream>();
    S->ctx = as_ctx(ctx);
    S->fields = st.fields;
    for (const ipc::Batch& b : st.batches) {
      S->rows.push_back(b.n_rows);
      S->cols.emplace_back();
      for (const HostColumn& c : b.cols) S->cols.back().push_back(upload(ctx, c));
    }
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    *out = S.release();
  });
}

prism_status prism_ipc_stream_info(const prism_ipc_stream* s, uint32_t* n_batches, uint32_t* n_columns) {
  return guarded([&] {
    if (!s) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL stream");
    if (n_batches) *n_batches = static_cast<uint32_t>(s->rows.size());
    if (n_columns) *n_columns = static_cast<uint32_t>(s->fields.size());
  });
}

prism_status prism_ipc_field(const prism_ipc_stream* s, uint32_t column, const char** name, int32_t* type,
                             int32_t* nullable) {
  return guarded([&] {
    if (!s || column >= s->fields.size()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such column");
    if (name) *name = s->fields[column].name.c_str();
    if (type) *type = static_cast<int32_t>(s->fields[column].type);
    if (nullable) *nullable = s->fields[column].nullable ? 1 : 0;
  });
}

prism_status prism_ipc_column(const prism_ipc_stream* s, uint32_t batch, uint32_t column, uint64_t* n_rows,
                              prism_column* out) {
  return guarded([&] {
    if (!s || batch >= s->rows.size() || column >= s->fields.size())
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such batch or column");
    if (n_rows) *n_rows = static_cast<uint64_t>(s->rows[batch]);
    if (out) *out = s->cols[batch][column].desc();
  });
}

void prism_ipc_stream_free(prism_ipc_stream* s) {
  if (!s) return;
  if (s->ctx) (void)hipSetDevice(s->ctx->device);
  delete s;
}

}  // extern "C"
