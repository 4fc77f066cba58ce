// Parquet column-chunk decode into HBM: the step upstream of the operators
// (prism-executor's parquet_scan reads files with the `parquet` crate into
// RecordBatches, parquet_scan.rs:55-258).  The file is decoded without a
// Parquet library:
//   file   = "PAR1" [column chunks ...] [FileMetaData] [u32 footer length] "PAR1"
//   chunk  = [dictionary page] [data pages ...], each page = [PageHeader] [payload]
// Metadata and page headers are Thrift compact-protocol structs.  Only flat
// schemas of BOOLEAN / INT32 (int32, date) / INT64 / DOUBLE / BYTE_ARRAY (utf8)
// leaves are decoded; nested or repeated fields, other physical or logical
// types, DELTA_* / BYTE_STREAM_SPLIT encodings and codecs other than
// UNCOMPRESSED / SNAPPY / GZIP are rejected with an Arrow error.
//
// Division of the work.  The host walks page headers, inflates payloads,
// turns definition levels into the Arrow validity bitmap (bit width 1 for a
// flat optional column), copies PLAIN values (length-prefixed strings are
// walked here: their offsets form a serial chain) and records the index
// stream of dictionary pages as a run table (RLE runs and bit-packed runs
// whose bytes travel unchanged).  The device then resolves every row to an
// entry of one source column (dictionary entries followed by PLAIN values)
// and gathers it: validity popcounts per 64-row word + scan give each row its
// value ordinal, a binary search over the run table its run, and the run its
// source index (RLE value, bit-packed field or PLAIN ordinal).
#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "arrow_io.hpp"
#include "gather.hpp"

namespace prism {
namespace pq {

[[noreturn]] static void bad(const std::string& m) { fail(PRISM_ERR_ARROW, "Arrow error: Parquet " + m); }

// ------------------------------------------------------------ Thrift compact
enum TType : int { T_STOP = 0, T_TRUE = 1, T_FALSE = 2, T_BYTE = 3, T_I16 = 4, T_I32 = 5, T_I64 = 6, T_DOUBLE = 7,
                   T_BINARY = 8, T_LIST = 9, T_SET = 10, T_MAP = 11, T_STRUCT = 12 };

struct Thrift {
  const uint8_t* p;
  const uint8_t* end;
  uint8_t byte() {
    if (p >= end) bad("metadata is truncated");
    return *p++;
  }
  uint64_t varint() {
    uint64_t v = 0;
    for (int s = 0; s < 64; s += 7) {
      const uint8_t b = byte();
      v |= static_cast<uint64_t>(b & 0x7f) << s;
      if (!(b & 0x80)) return v;
    }
    bad("metadata holds an over-long varint");
  }
  int64_t zigzag() {
    const uint64_t v = varint();
    return static_cast<int64_t>(v >> 1) ^ -static_cast<int64_t>(v & 1);
  }
  std::string binary() {
    const uint64_t n = varint();
    if (n > static_cast<uint64_t>(end - p)) bad("metadata is truncated");
    std::string s(reinterpret_cast<const char*>(p), n);
    p += n;
    return s;
  }
  // Next field of the current struct; false at its end.  Field ids are deltas
  // from the previous id (high nibble) or a zigzag i16 when the delta is 0.
  bool field(int16_t& last, int16_t& id, int& type) {
    const uint8_t h = byte();
    if (h == T_STOP) return false;
    type = h & 0x0f;
    const int d = h >> 4;
    id = d ? static_cast<int16_t>(last + d) : static_cast<int16_t>(zigzag());
    last = id;
    return true;
  }
  void list(int& elem, uint64_t& n) {
    const uint8_t h = byte();
    elem = h & 0x0f;
    n = h >> 4;
    if (n == 15) n = varint();
    if (n > static_cast<uint64_t>(end - p)) bad("metadata list is longer than the metadata");
  }
  int64_t integer(int type) {
    if (type == T_BYTE) return static_cast<int8_t>(byte());
    if (type == T_I16 || type == T_I32 || type == T_I64) return zigzag();
    bad("metadata field has type " + std::to_string(type) + " where an integer belongs");
  }
  void skip(int type, int depth = 0) {
    if (depth > 32) bad("metadata nests too deeply");
    switch (type) {
      case T_TRUE: case T_FALSE: return;  // the value is the field type
      case T_BYTE: byte(); return;
      case T_I16: case T_I32: case T_I64: varint(); return;
      case T_DOUBLE:
        if (end - p < 8) bad("metadata is truncated");
        p += 8;
        return;
      case T_BINARY: {
        const uint64_t n = varint();
        if (n > static_cast<uint64_t>(end - p)) bad("metadata is truncated");
        p += n;
        return;
      }
      case T_LIST: case T_SET: {
        int e;
        uint64_t n;
        list(e, n);
        for (uint64_t i = 0; i < n; ++i) elem(e, depth + 1);
        return;
      }
      case T_MAP: {
        const uint64_t n = varint();
        if (!n) return;
        const uint8_t kv = byte();
        for (uint64_t i = 0; i < n; ++i) {
          elem(kv >> 4, depth + 1);
          elem(kv & 0x0f, depth + 1);
        }
        return;
      }
      case T_STRUCT: {
        int16_t last = 0, id;
        int t;
        while (field(last, id, t)) skip(t, depth + 1);
        return;
      }
      default: bad("metadata holds unknown Thrift type " + std::to_string(type));
    }
  }
  // Container elements: booleans take one byte there.
  void elem(int type, int depth) {
    if (type == T_TRUE || type == T_FALSE) byte();
    else skip(type, depth);
  }
};

// ------------------------------------------------------------ metadata
enum Phys : int { P_BOOLEAN = 0, P_INT32 = 1, P_INT64 = 2, P_INT96 = 3, P_FLOAT = 4, P_DOUBLE = 5, P_BYTE_ARRAY = 6,
                  P_FIXED = 7 };
enum Codec : int { C_NONE = 0, C_SNAPPY = 1, C_GZIP = 2 };
enum PageType : int { PG_DATA = 0, PG_INDEX = 1, PG_DICT = 2, PG_DATA_V2 = 3 };
enum Enc : int { E_PLAIN = 0, E_PLAIN_DICT = 2, E_RLE = 3, E_RLE_DICT = 8 };

struct Stats {
  bool present = false, has_min = false, has_max = false;
  int64_t null_count = -1;
  std::string min, max;
};

struct ColMeta {
  int type = -1, codec = 0;
  int64_t num_values = 0, compressed = 0, data_off = -1, dict_off = -1;
  std::string path;  // last element of path_in_schema
  Stats stats;
};

struct RowGroup {
  std::vector<ColMeta> cols;
  int64_t num_rows = 0;
};

struct Leaf {
  std::string name;
  int phys = -1;
  bool optional = false;
  DType type = DType::Int64;
  bool supported = false;
  std::string why;  // reason when not supported
};

struct FileMeta {
  int64_t num_rows = 0;
  std::vector<Leaf> leaves;
  std::vector<RowGroup> rgs;
};

static Stats read_stats(Thrift& T) {
  Stats s;
  s.present = true;
  std::string omin, omax;
  bool has_omin = false, has_omax = false;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1 && t == T_BINARY) { omax = T.binary(); has_omax = true; }
    else if (id == 2 && t == T_BINARY) { omin = T.binary(); has_omin = true; }
    else if (id == 3) s.null_count = T.integer(t);
    else if (id == 5 && t == T_BINARY) { s.max = T.binary(); s.has_max = true; }
    else if (id == 6 && t == T_BINARY) { s.min = T.binary(); s.has_min = true; }
    else T.skip(t);
  }
  // min_value / max_value (fields 5, 6) supersede the deprecated min / max
  // (fields 1, 2), which older writers filled alone
  if (!s.has_min && has_omin) { s.min = omin; s.has_min = true; }
  if (!s.has_max && has_omax) { s.max = omax; s.has_max = true; }
  return s;
}

static ColMeta read_col_meta(Thrift& T) {
  ColMeta m;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1) m.type = static_cast<int>(T.integer(t));
    else if (id == 3 && t == T_LIST) {
      int e;
      uint64_t n;
      T.list(e, n);
      for (uint64_t i = 0; i < n; ++i) {
        if (e != T_BINARY) bad("path_in_schema is not a string list");
        m.path = T.binary();
      }
    } else if (id == 4) m.codec = static_cast<int>(T.integer(t));
    else if (id == 5) m.num_values = T.integer(t);
    else if (id == 7) m.compressed = T.integer(t);
    else if (id == 9) m.data_off = T.integer(t);
    else if (id == 11) m.dict_off = T.integer(t);
    else if (id == 12 && t == T_STRUCT) m.stats = read_stats(T);
    else T.skip(t);
  }
  return m;
}

static ColMeta read_col_chunk(Thrift& T) {
  ColMeta m;
  bool have = false;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1 && t == T_BINARY) {
      if (!T.binary().empty()) bad("column chunks in other files are not supported");
    } else if (id == 3 && t == T_STRUCT) {
      m = read_col_meta(T);
      have = true;
    } else T.skip(t);
  }
  if (!have) bad("column chunk without ColumnMetaData");
  return m;
}

static RowGroup read_row_group(Thrift& T) {
  RowGroup g;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1 && t == T_LIST) {
      int e;
      uint64_t n;
      T.list(e, n);
      for (uint64_t i = 0; i < n; ++i) {
        if (e != T_STRUCT) bad("row group columns are not structs");
        g.cols.push_back(read_col_chunk(T));
      }
    } else if (id == 3) g.num_rows = T.integer(t);
    else T.skip(t);
  }
  return g;
}

struct SchemaEl {
  int type = -1, repetition = 0, num_children = 0, converted = -1;
  int logical = -1;               // set member of the LogicalType union
  int int_bits = 0, int_signed = 1;  // LogicalType INTEGER
  std::string name;
};

static void read_logical(Thrift& T, SchemaEl& s) {
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    s.logical = id;
    if (id == 10 && t == T_STRUCT) {  // IntType {1: i8 bitWidth, 2: bool isSigned}
      int16_t l2 = 0, i2;
      int t2;
      while (T.field(l2, i2, t2)) {
        if (i2 == 1) s.int_bits = static_cast<int>(T.integer(t2));
        else if (i2 == 2 && (t2 == T_TRUE || t2 == T_FALSE)) s.int_signed = t2 == T_TRUE;
        else T.skip(t2);
      }
    } else T.skip(t);
  }
}

static SchemaEl read_schema_el(Thrift& T) {
  SchemaEl s;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1) s.type = static_cast<int>(T.integer(t));
    else if (id == 3) s.repetition = static_cast<int>(T.integer(t));
    else if (id == 4 && t == T_BINARY) s.name = T.binary();
    else if (id == 5) s.num_children = static_cast<int>(T.integer(t));
    else if (id == 6) s.converted = static_cast<int>(T.integer(t));
    else if (id == 10 && t == T_STRUCT) read_logical(T, s);
    else T.skip(t);
  }
  return s;
}

// Arrow type of a leaf, as arrow's Parquet reader maps it for this set.
static Leaf leaf_of(const SchemaEl& s) {
  Leaf L;
  L.name = s.name;
  L.phys = s.type;
  L.optional = s.repetition == 1;
  auto no = [&](const std::string& w) { L.supported = false; L.why = w; return L; };
  if (s.repetition == 2) return no("repeated field");
  L.supported = true;
  switch (s.type) {
    case P_BOOLEAN: L.type = DType::Bool; return L;
    case P_INT32:
      if (s.converted == 6 || s.logical == 6) { L.type = DType::Date32; return L; }
      if ((s.converted == -1 || s.converted == 17) &&
          (s.logical == -1 || (s.logical == 10 && s.int_bits == 32 && s.int_signed))) {
        L.type = DType::Int32;
        return L;
      }
      return no("INT32 with converted type " + std::to_string(s.converted) + " / logical type " +
                std::to_string(s.logical));
    case P_INT64:
      if ((s.converted == -1 || s.converted == 18) &&
          (s.logical == -1 || (s.logical == 10 && s.int_bits == 64 && s.int_signed))) {
        L.type = DType::Int64;
        return L;
      }
      return no("INT64 with converted type " + std::to_string(s.converted) + " / logical type " +
                std::to_string(s.logical));
    case P_DOUBLE: L.type = DType::Float64; return L;
    case P_BYTE_ARRAY:
      if (s.converted == 0 || s.logical == 1) { L.type = DType::Utf8; return L; }
      return no("BYTE_ARRAY without the UTF8 / STRING annotation");
    default: return no("physical type " + std::to_string(s.type));
  }
}

static FileMeta read_footer(const uint8_t* bytes, uint64_t n) {
  if (!bytes || n < 12) bad("file is shorter than its magic numbers");
  if (memcmp(bytes, "PAR1", 4) || memcmp(bytes + n - 4, "PAR1", 4)) bad("file does not start and end with PAR1");
  uint32_t len;
  memcpy(&len, bytes + n - 8, 4);
  if (len > n - 12) bad("footer length exceeds the file");
  Thrift T{bytes + n - 8 - len, bytes + n - 8};
  FileMeta F;
  std::vector<SchemaEl> schema;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 2 && t == T_LIST) {
      int e;
      uint64_t m;
      T.list(e, m);
      for (uint64_t i = 0; i < m; ++i) {
        if (e != T_STRUCT) bad("schema elements are not structs");
        schema.push_back(read_schema_el(T));
      }
    } else if (id == 3) F.num_rows = T.integer(t);
    else if (id == 4 && t == T_LIST) {
      int e;
      uint64_t m;
      T.list(e, m);
      for (uint64_t i = 0; i < m; ++i) {
        if (e != T_STRUCT) bad("row groups are not structs");
        F.rgs.push_back(read_row_group(T));
      }
    } else T.skip(t);
  }
  if (schema.empty()) bad("schema is empty");
  if (static_cast<size_t>(schema[0].num_children) != schema.size() - 1)
    bad("nested schemas are not supported (root has " + std::to_string(schema[0].num_children) + " children, " +
        std::to_string(schema.size() - 1) + " elements follow)");
  for (size_t i = 1; i < schema.size(); ++i) {
    if (schema[i].num_children) bad("nested field '" + schema[i].name + "' is not supported");
    F.leaves.push_back(leaf_of(schema[i]));
  }
  for (const RowGroup& g : F.rgs) {
    if (g.cols.size() != F.leaves.size()) bad("row group column count differs from the schema");
    if (g.num_rows < 0) bad("row group has a negative row count");
    for (size_t c = 0; c < g.cols.size(); ++c) {
      const ColMeta& m = g.cols[c];
      if (m.type != F.leaves[c].phys) bad("column chunk type differs from the schema");
      if (m.compressed < 0 || m.data_off < 0) bad("column chunk has a negative offset or size");
    }
  }
  return F;
}

// ------------------------------------------------------------ page headers
struct PageHeader {
  int type = -1;
  int64_t usize = 0, csize = 0;
  // data page (v1 fields; v2 adds the level lengths)
  int64_t num_values = 0, num_nulls = -1;
  int encoding = -1, def_encoding = -1;
  int64_t def_len = 0, rep_len = 0;
  bool v2_compressed = true;
  // dictionary page
  int64_t dict_values = 0;
  int dict_encoding = -1;
};

static PageHeader read_page_header(Thrift& T) {
  PageHeader h;
  int16_t last = 0, id;
  int t;
  while (T.field(last, id, t)) {
    if (id == 1) h.type = static_cast<int>(T.integer(t));
    else if (id == 2) h.usize = T.integer(t);
    else if (id == 3) h.csize = T.integer(t);
    else if (id == 5 && t == T_STRUCT) {
      int16_t l = 0, i;
      int tt;
      while (T.field(l, i, tt)) {
        if (i == 1) h.num_values = T.integer(tt);
        else if (i == 2) h.encoding = static_cast<int>(T.integer(tt));
        else if (i == 3) h.def_encoding = static_cast<int>(T.integer(tt));
        else T.skip(tt);
      }
    } else if (id == 7 && t == T_STRUCT) {
      int16_t l = 0, i;
      int tt;
      while (T.field(l, i, tt)) {
        if (i == 1) h.dict_values = T.integer(tt);
        else if (i == 2) h.dict_encoding = static_cast<int>(T.integer(tt));
        else T.skip(tt);
      }
    } else if (id == 8 && t == T_STRUCT) {
      int16_t l = 0, i;
      int tt;
      while (T.field(l, i, tt)) {
        if (i == 1) h.num_values = T.integer(tt);
        else if (i == 2) h.num_nulls = T.integer(tt);
        else if (i == 4) h.encoding = static_cast<int>(T.integer(tt));
        else if (i == 5) h.def_len = T.integer(tt);
        else if (i == 6) h.rep_len = T.integer(tt);
        else if (i == 7 && (tt == T_TRUE || tt == T_FALSE)) h.v2_compressed = tt == T_TRUE;
        else T.skip(tt);
      }
    } else T.skip(t);
  }
  if (h.usize < 0 || h.csize < 0 || h.num_values < 0 || h.dict_values < 0 || h.def_len < 0 || h.rep_len < 0)
    bad("page header has a negative size");
  return h;
}

// ------------------------------------------------------------ codecs
// Snappy block format (google/snappy format_description.txt): a varint of the
// uncompressed length, then elements whose tag's low two bits select a
// literal (00) or a copy with a 1-, 2- or 4-byte offset (01, 10, 11).
static void snappy(const uint8_t* in, size_t n, uint8_t* out, size_t out_n) {
  size_t i = 0;
  uint64_t len = 0;
  for (int s = 0;; s += 7) {
    if (i >= n || s > 28) bad("snappy block has a bad length header");
    const uint8_t b = in[i++];
    len |= static_cast<uint64_t>(b & 0x7f) << s;
    if (!(b & 0x80)) break;
  }
  if (len != out_n) bad("snappy block length differs from the page header");
  size_t o = 0;
  auto copy = [&](size_t off, size_t l) {
    if (off == 0 || off > o || l > out_n - o) bad("snappy copy is out of range");
    uint8_t* d = out + o;
    const uint8_t* s = d - off;
    if (off >= l) memcpy(d, s, l);
    else
      for (size_t k = 0; k < l; ++k) d[k] = s[k];  // overlapping: repeats the last `off` bytes
    o += l;
  };
  while (i < n) {
    const uint8_t tag = in[i++];
    switch (tag & 3) {
      case 0: {
        size_t l = tag >> 2;
        if (l >= 60) {
          const size_t nb = l - 59;
          if (nb > n - i) bad("snappy literal header is truncated");
          l = 0;
          for (size_t k = 0; k < nb; ++k) l |= static_cast<size_t>(in[i + k]) << (8 * k);
          i += nb;
        }
        l += 1;
        if (l > n - i || l > out_n - o) bad("snappy literal is out of range");
        memcpy(out + o, in + i, l);
        i += l;
        o += l;
        break;
      }
      case 1: {
        if (i >= n) bad("snappy copy is truncated");
        const size_t l = 4 + ((tag >> 2) & 7);
        const size_t off = (static_cast<size_t>(tag >> 5) << 8) | in[i++];
        copy(off, l);
        break;
      }
      case 2: {
        if (2 > n - i) bad("snappy copy is truncated");
        const size_t off = in[i] | (static_cast<size_t>(in[i + 1]) << 8);
        i += 2;
        copy(off, 1 + (tag >> 2));
        break;
      }
      default: {
        if (4 > n - i) bad("snappy copy is truncated");
        uint32_t off;
        memcpy(&off, in + i, 4);
        i += 4;
        copy(off, 1 + (tag >> 2));
        break;
      }
    }
  }
  if (o != out_n) bad("snappy block is shorter than its length header");
}

static void gunzip(const uint8_t* in, size_t n, uint8_t* out, size_t out_n) {
  z_stream z{};
  if (inflateInit2(&z, 15 + 32) != Z_OK) bad("gzip: inflateInit2 failed");
  z.next_in = const_cast<Bytef*>(in);
  z.avail_in = static_cast<uInt>(n);
  z.next_out = out;
  z.avail_out = static_cast<uInt>(out_n);
  const int r = inflate(&z, Z_FINISH);
  const size_t got = out_n - z.avail_out;
  inflateEnd(&z);
  if (r != Z_STREAM_END || got != out_n) bad("gzip page does not inflate to its header's size");
}

// Page payload after decompression: a view of the file bytes or an owned copy.
struct Payload {
  const uint8_t* p = nullptr;
  size_t n = 0;
  std::vector<uint8_t> own;
};

static void inflate_into(int codec, const uint8_t* in, size_t n, size_t out_n, Payload& out) {
  if (codec == C_NONE) {
    if (n != out_n) bad("uncompressed page sizes differ");
    out.p = in;
    out.n = n;
    return;
  }
  out.own.resize(out_n ? out_n : 1);
  if (codec == C_SNAPPY) snappy(in, n, out.own.data(), out_n);
  else if (codec == C_GZIP) gunzip(in, n, out.own.data(), out_n);
  else bad("compression codec " + std::to_string(codec) + " is not supported (UNCOMPRESSED, SNAPPY, GZIP are)");
  out.p = out.own.data();
  out.n = out_n;
}

// ------------------------------------------------------------ chunk decode (host part)
// A run of value ordinals [v0, next run's v0) and where their source index comes from.
enum : uint32_t { RUN_RLE = 0, RUN_BITS = 1, RUN_PLAIN = 2 };
struct Run {
  uint64_t v0;
  uint32_t info;  // kind | bit width << 8
  uint32_t data;  // RLE: index; BITS: first 32-bit word of the packed bytes; PLAIN: source index of v0
};
static_assert(sizeof(Run) == 16, "run layout");

// Values in source order: dictionary entries, then PLAIN page values.
struct Source {
  std::vector<uint8_t> fixed;
  std::vector<int32_t> off{0};
  std::vector<uint8_t> bytes;
  std::vector<uint8_t> bits;
  uint64_t n = 0;
};

struct HostChunk {
  Leaf leaf;
  int64_t rows = 0;
  int64_t null_count = 0;
  std::vector<uint8_t> validity;  // (rows + 7) / 8 bytes, rounded up to 8; empty without nulls
  Source src;
  std::vector<Run> runs;
  std::vector<uint32_t> words;  // bit-packed index runs, each word-aligned and followed by a zero word
  bool plain_only = true;
};

// Append nbits bits (LSB first) of src to dst at bit position pos.
static void put_bits(std::vector<uint8_t>& dst, uint64_t pos, const uint8_t* src, uint64_t nbits) {
  const uint64_t need = (pos + nbits + 7) / 8;
  if (dst.size() < need) dst.resize(need, 0);
  const unsigned sh = pos & 7;
  uint8_t* d = dst.data() + pos / 8;
  for (uint64_t k = 0; k * 8 < nbits; ++k) {
    uint8_t b = src[k];
    const uint64_t left = nbits - k * 8;
    if (left < 8) b &= static_cast<uint8_t>((1u << left) - 1);
    d[k] |= static_cast<uint8_t>(b << sh);
    if (sh && (b >> (8 - sh))) d[k + 1] |= static_cast<uint8_t>(b >> (8 - sh));
  }
}

static void set_bits(std::vector<uint8_t>& dst, uint64_t pos, uint64_t n) {
  if (!n) return;
  const uint64_t need = (pos + n + 7) / 8;
  if (dst.size() < need) dst.resize(need, 0);
  uint64_t i = pos;
  const uint64_t e = pos + n;
  for (; i < e && (i & 7); ++i) dst[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
  if (e - i >= 8) {
    memset(dst.data() + (i >> 3), 0xff, (e - i) >> 3);
    i += (e - i) & ~7ULL;
  }
  for (; i < e; ++i) dst[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
}

static uint64_t read_uleb(const uint8_t*& q, const uint8_t* e) {
  uint64_t v = 0;
  for (int s = 0; s < 64; s += 7) {
    if (q >= e) bad("RLE run header is truncated");
    const uint8_t b = *q++;
    v |= static_cast<uint64_t>(b & 0x7f) << s;
    if (!(b & 0x80)) return v;
  }
  bad("RLE run header is an over-long varint");
}

// Definition levels of a flat optional column (bit width 1, RLE / bit-packed
// hybrid): set the validity bits of rows [row0, row0 + nv); returns the number set.
static uint64_t def_levels(const uint8_t* q, const uint8_t* e, uint64_t nv, std::vector<uint8_t>& valid,
                           uint64_t row0) {
  uint64_t done = 0, set = 0;
  while (done < nv) {
    const uint64_t h = read_uleb(q, e);
    if (!(h & 1)) {
      const uint64_t cnt = std::min<uint64_t>(h >> 1, nv - done);
      if (q >= e) bad("RLE run value is truncated");
      const uint8_t v = *q++;
      if (v > 1) bad("definition level exceeds the maximum level 1");
      if (!(h >> 1)) bad("RLE run of length 0");
      if (v) {
        set_bits(valid, row0 + done, cnt);
        set += cnt;
      }
      done += cnt;
    } else {
      const uint64_t groups = h >> 1;
      if (!groups) bad("bit-packed run of length 0");
      if (groups > static_cast<uint64_t>(e - q)) bad("bit-packed definition levels are truncated");
      const uint64_t left = nv - done;
      const uint64_t cnt = groups >= (left + 7) / 8 ? left : groups * 8;  // no product overflow
      put_bits(valid, row0 + done, q, cnt);
      for (uint64_t k = 0; k * 8 < cnt; ++k) {
        uint8_t b = q[k];
        if (cnt - k * 8 < 8) b &= static_cast<uint8_t>((1u << (cnt - k * 8)) - 1);
        set += __builtin_popcount(b);
      }
      q += groups;
      done += cnt;
    }
  }
  return set;
}

// Dictionary indices of one data page (bit width byte already read): append runs.
static void index_runs(const uint8_t* q, const uint8_t* e, int bw, uint64_t count, uint64_t v0, uint64_t n_dict,
                       HostChunk& H) {
  if (bw > 32) bad("dictionary index bit width " + std::to_string(bw) + " exceeds 32");
  const int vbytes = (bw + 7) / 8;
  uint64_t done = 0;
  while (done < count) {
    const uint64_t h = read_uleb(q, e);
    if (!(h & 1)) {
      if (!(h >> 1)) bad("RLE run of length 0");
      const uint64_t cnt = std::min<uint64_t>(h >> 1, count - done);
      if (vbytes > e - q) bad("RLE run value is truncated");
      uint32_t v = 0;
      for (int k = 0; k < vbytes; ++k) v |= static_cast<uint32_t>(q[k]) << (8 * k);
      q += vbytes;
      if (v >= n_dict) bad("dictionary index " + std::to_string(v) + " exceeds the dictionary");
      H.runs.push_back({v0 + done, RUN_RLE, v});
      done += cnt;
    } else {
      const uint64_t groups = h >> 1;
      if (!groups) bad("bit-packed run of length 0");
      // bounded before the product: a forged header must not wrap it past the check
      if (bw && groups > static_cast<uint64_t>(e - q) / static_cast<uint64_t>(bw))
        bad("bit-packed dictionary indices are truncated");
      const uint64_t nbytes = groups * static_cast<uint64_t>(bw);
      const uint64_t left = count - done;
      const uint64_t cnt = groups >= (left + 7) / 8 ? left : groups * 8;  // no product overflow
      if (bw == 0) {  // every index is 0 and no bytes follow
        if (!n_dict) bad("dictionary index exceeds the empty dictionary");
        H.runs.push_back({v0 + done, RUN_RLE, 0u});
        done += cnt;
        continue;
      }
      const size_t w0 = H.words.size();
      if (w0 > 0xffffffffULL - (nbytes + 3) / 4 - 1) bad("column chunk index stream exceeds 16 GiB");
      H.words.resize(w0 + (nbytes + 3) / 4 + 1, 0);
      memcpy(H.words.data() + w0, q, nbytes);
      q += nbytes;
      H.runs.push_back({v0 + done, RUN_BITS | (static_cast<uint32_t>(bw) << 8), static_cast<uint32_t>(w0)});
      done += cnt;
    }
  }
}

// PLAIN values appended to the source (count values from q).
static void plain_values(const uint8_t* q, const uint8_t* e, uint64_t count, DType t, Source& S) {
  const uint64_t avail = static_cast<uint64_t>(e - q);
  switch (t) {
    case DType::Bool:
      if ((count + 7) / 8 > avail) bad("PLAIN booleans are truncated");
      put_bits(S.bits, S.n, q, count);
      break;
    case DType::Utf8:
      for (uint64_t i = 0; i < count; ++i) {
        if (4 > static_cast<uint64_t>(e - q)) bad("PLAIN byte array length is truncated");
        uint32_t l;
        memcpy(&l, q, 4);
        q += 4;
        if (l > static_cast<uint64_t>(e - q)) bad("PLAIN byte array is truncated");
        if (S.bytes.size() + l > 0x7fffffffULL) bad("column chunk strings exceed i32 offsets");
        S.bytes.insert(S.bytes.end(), q, q + l);
        S.off.push_back(static_cast<int32_t>(S.bytes.size()));
        q += l;
      }
      break;
    default: {
      const uint64_t w = static_cast<uint64_t>(dtype_width(t));
      if (count > avail / w) bad("PLAIN values are truncated");
      S.fixed.insert(S.fixed.end(), q, q + count * w);
      break;
    }
  }
  S.n += count;
}

// Walk one column chunk's pages.
static HostChunk decode_host(const uint8_t* bytes, uint64_t n, const FileMeta& F, uint32_t rg, uint32_t col) {
  const RowGroup& G = F.rgs[rg];
  const ColMeta& m = G.cols[col];
  HostChunk H;
  H.leaf = F.leaves[col];
  if (!H.leaf.supported) bad("column '" + H.leaf.name + "': " + H.leaf.why + " is not supported");
  H.rows = G.num_rows;
  if (H.rows > 0x7fffffffLL) bad("row group of more than 2^31 - 1 rows");
  const DType t = H.leaf.type;
  const bool opt = H.leaf.optional;
  uint64_t start = static_cast<uint64_t>(m.data_off);
  if (m.dict_off > 0 && static_cast<uint64_t>(m.dict_off) < start) start = static_cast<uint64_t>(m.dict_off);
  if (start > n || static_cast<uint64_t>(m.compressed) > n - start) bad("column chunk lies outside the file");
  const uint8_t* p = bytes + start;
  const uint8_t* end = p + m.compressed;
  uint64_t rows = 0, vals = 0, n_dict = 0;
  bool have_dict = false, any_data = false;
  if (opt) H.validity.assign((static_cast<uint64_t>(H.rows) + 63) / 64 * 8, 0);
  while (p < end && rows < static_cast<uint64_t>(H.rows)) {
    Thrift T{p, end};
    const PageHeader h = read_page_header(T);
    p = T.p;
    if (static_cast<uint64_t>(h.csize) > static_cast<uint64_t>(end - p)) bad("page exceeds its column chunk");
    const uint8_t* body = p;
    p += h.csize;
    if (h.type == PG_DICT) {
      if (have_dict || any_data) bad("dictionary page is not the chunk's first page");
      if (h.dict_encoding != E_PLAIN && h.dict_encoding != E_PLAIN_DICT)
        bad("dictionary page encoding " + std::to_string(h.dict_encoding) + " is not supported");
      Payload d;
      inflate_into(m.codec, body, h.csize, h.usize, d);
      plain_values(d.p, d.p + d.n, h.dict_values, t, H.src);
      n_dict = H.src.n;
      have_dict = true;
      continue;
    }
    if (h.type != PG_DATA && h.type != PG_DATA_V2) continue;  // index pages carry nothing to decode
    any_data = true;
    const uint64_t nv = h.num_values;
    if (nv > static_cast<uint64_t>(H.rows) - rows) bad("data pages hold more values than the row group has rows");
    Payload d;
    const uint8_t* lv = nullptr;
    const uint8_t* lv_end = nullptr;
    const uint8_t* q;
    const uint8_t* qe;
    if (h.type == PG_DATA) {
      inflate_into(m.codec, body, h.csize, h.usize, d);
      q = d.p;
      qe = d.p + d.n;
      if (opt) {
        if (h.def_encoding != E_RLE) bad("definition level encoding " + std::to_string(h.def_encoding) +
                                         " is not supported");
        if (4 > qe - q) bad("definition levels are truncated");
        uint32_t len;
        memcpy(&len, q, 4);
        q += 4;
        if (len > static_cast<uint64_t>(qe - q)) bad("definition levels are truncated");
        lv = q;
        lv_end = q + len;
        q += len;
      }
    } else {
      if (h.rep_len) bad("repetition levels are not supported");
      if (static_cast<uint64_t>(h.def_len) > static_cast<uint64_t>(h.csize) || h.def_len > h.usize)
        bad("definition levels exceed the page");
      lv = body;
      lv_end = body + h.def_len;
      const uint8_t* vb = body + h.def_len;
      const uint64_t vc = h.csize - h.def_len, vu = h.usize - h.def_len;
      inflate_into(h.v2_compressed ? m.codec : C_NONE, vb, vc, h.v2_compressed ? vu : vc, d);
      q = d.p;
      qe = d.p + d.n;
    }
    uint64_t nonnull = nv;
    if (opt) nonnull = def_levels(lv, lv_end, nv, H.validity, rows);
    else if (h.type == PG_DATA_V2 && h.def_len) bad("required column carries definition levels");
    if (h.type == PG_DATA_V2 && h.num_nulls >= 0 && static_cast<uint64_t>(h.num_nulls) != nv - nonnull)
      bad("page null count differs from its definition levels");
    switch (h.encoding) {
      case E_PLAIN:
        H.runs.push_back({vals, RUN_PLAIN, static_cast<uint32_t>(H.src.n)});
        plain_values(q, qe, nonnull, t, H.src);
        break;
      case E_RLE: {
        if (t != DType::Bool) bad("RLE value encoding on a non-boolean column");
        // booleans: 4-byte length, then a bit width 1 hybrid stream decoded here into bits
        if (4 > qe - q) bad("RLE booleans are truncated");
        uint32_t len;
        memcpy(&len, q, 4);
        q += 4;
        if (len > static_cast<uint64_t>(qe - q)) bad("RLE booleans are truncated");
        std::vector<uint8_t> bits((nonnull + 7) / 8 + 8, 0);
        def_levels(q, q + len, nonnull, bits, 0);  // the same bit-width-1 hybrid stream as the levels
        H.runs.push_back({vals, RUN_PLAIN, static_cast<uint32_t>(H.src.n)});
        put_bits(H.src.bits, H.src.n, bits.data(), nonnull);
        H.src.n += nonnull;
        break;
      }
      case E_PLAIN_DICT:
      case E_RLE_DICT: {
        if (!have_dict) bad("dictionary-encoded page without a dictionary page");
        H.plain_only = false;
        if (nonnull) {
          if (q >= qe) bad("dictionary index bit width is missing");
          const int bw = *q++;
          index_runs(q, qe, bw, nonnull, vals, n_dict, H);
        }
        break;
      }
      default:
        bad("column '" + H.leaf.name + "': value encoding " + std::to_string(h.encoding) + " is not supported");
    }
    rows += nv;
    vals += nonnull;
  }
  if (rows != static_cast<uint64_t>(H.rows)) bad("column chunk pages hold fewer rows than the row group");
  if (H.src.n > 0xffffffffULL) bad("column chunk holds more than 2^32 values");
  H.null_count = static_cast<int64_t>(rows - vals);
  if (!H.null_count) H.validity.clear();
  if (have_dict) H.plain_only = H.plain_only && n_dict == 0;
  return H;
}

// ------------------------------------------------------------ device part
constexpr int kPB = 256;

__global__ void k_pq_word_counts(const unsigned long long* __restrict__ valid, uint64_t n_words,
                                 uint32_t* __restrict__ cnt) {
  for (uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; w < n_words;
       w += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    cnt[w] = static_cast<uint32_t>(__popcll(valid[w]));
}

// Source index of every row: value ordinal from the validity prefix, run by
// binary search, index from the run.  Null rows get index 0 (masked by take).
__global__ __launch_bounds__(kPB) void k_pq_index(const Run* __restrict__ runs, uint32_t n_runs,
                                                  const uint32_t* __restrict__ words,
                                                  const unsigned long long* __restrict__ valid,
                                                  const uint32_t* __restrict__ prefix, uint64_t n_rows,
                                                  uint64_t n_src, uint32_t* __restrict__ idx,
                                                  uint32_t* __restrict__ err) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < n_rows;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    bool ok = true;
    uint64_t v = r;
    if (valid) {
      const unsigned long long w = valid[r >> 6];
      const unsigned s = r & 63;
      ok = (w >> s) & 1;
      v = prefix[r >> 6] + static_cast<uint64_t>(__popcll(w & ((1ULL << s) - 1)));
    }
    uint32_t out = 0;
    if (ok) {
      uint32_t lo = 0, hi = n_runs - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (runs[mid].v0 <= v) lo = mid;
        else hi = mid - 1;
      }
      const Run R = runs[lo];
      const uint32_t kind = R.info & 0xff, bw = R.info >> 8;
      const uint64_t k = v - R.v0;
      uint64_t x;
      if (kind == RUN_RLE) x = R.data;
      else if (kind == RUN_BITS) {
        const uint64_t bit = k * bw;
        const uint64_t wi = R.data + (bit >> 5);
        const uint64_t win = static_cast<uint64_t>(words[wi]) | (static_cast<uint64_t>(words[wi + 1]) << 32);
        x = (win >> (bit & 31)) & ((1ULL << bw) - 1);
      } else x = R.data + k;
      if (x >= n_src) {
        atomicOr(err, 1u);
        x = 0;
      }
      out = static_cast<uint32_t>(x);
    }
    idx[r] = out;
  }
}

static dim3 pgrid(prism_ctx* ctx, uint64_t n) {
  const uint64_t b = ceil_div(n ? n : 1, static_cast<uint64_t>(kPB));
  const uint64_t cap = static_cast<uint64_t>(ctx->num_cus) * 8;
  return dim3(static_cast<unsigned>(b < cap ? b : cap));
}

// Host part done: move the chunk to HBM and resolve it there.
// `err` (device) is set when a dictionary index exceeds the source; the caller
// reads the flags of all chunks back once.
static DevColumn decode_device(prism_ctx* ctx, HostChunk& H, uint32_t* err) {
  Ctx* c = as_ctx(ctx);
  hipStream_t st = ctx->stream;
  HostColumn src;
  src.type = H.leaf.type;
  src.name = H.leaf.name;
  src.nullable = false;
  src.length = static_cast<int64_t>(H.src.n);
  if (H.leaf.type == DType::Utf8) {
    src.data = H.src.bytes.data();
    src.offsets = H.src.off.data();
  } else if (H.leaf.type == DType::Bool) {
    H.src.bits.resize((H.src.n + 7) / 8 + 1, 0);
    src.data = H.src.bits.data();
  } else {
    H.src.fixed.resize(std::max<size_t>(H.src.fixed.size(), 8));
    src.data = H.src.fixed.data();
  }
  DevColumn S = upload(ctx, src);
  const uint64_t rows = static_cast<uint64_t>(H.rows);
  // PLAIN pages without nulls: the source is the column
  if (H.plain_only && !H.null_count && H.src.n == rows) {
    S.nullable = H.leaf.optional;
    return S;
  }
  const uint64_t n_words = (rows + 63) / 64;
  DeviceBuffer runs(c, std::max<size_t>(H.runs.size(), 1) * sizeof(Run));
  DeviceBuffer words(c, std::max<size_t>(H.words.size(), 2) * 4);
  DeviceBuffer valid, counts, prefix, tmp;
  DeviceBuffer idx(c, std::max<uint64_t>(rows, 1) * 4);
  staged_h2d(ctx, runs.ptr, H.runs.data(), H.runs.size() * sizeof(Run));
  staged_h2d(ctx, words.ptr, H.words.data(), H.words.size() * 4);
  if (H.null_count) {
    valid.alloc(c, n_words * 8);
    counts.alloc(c, n_words * 4);
    prefix.alloc(c, n_words * 4);
    staged_h2d(ctx, valid.ptr, H.validity.data(), n_words * 8);
    hipLaunchKernelGGL(k_pq_word_counts, pgrid(ctx, n_words), dim3(kPB), 0, st, valid.as<unsigned long long>(),
                       n_words, counts.as<uint32_t>());
    size_t tb = 0;
    PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<uint32_t>(), prefix.as<uint32_t>(),
                                               static_cast<int>(n_words), st));
    tmp.alloc(c, tb ? tb : 8);
    PRISM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, tb, counts.as<uint32_t>(), prefix.as<uint32_t>(),
                                               static_cast<int>(n_words), st));
  }
  if (H.runs.empty() && rows != static_cast<uint64_t>(H.null_count))
    fail(PRISM_ERR_INTERNAL, "Internal error: Parquet chunk without runs for its values");
  if (rows) {
    KernelTimer kt(ctx, "parquet_index");
    hipLaunchKernelGGL(k_pq_index, pgrid(ctx, rows), dim3(kPB), 0, st, runs.as<Run>(),
                       static_cast<uint32_t>(H.runs.size()), words.as<uint32_t>(),
                       H.null_count ? valid.as<unsigned long long>() : nullptr,
                       H.null_count ? prefix.as<uint32_t>() : nullptr, rows, H.src.n, idx.as<uint32_t>(), err);
    PRISM_CHECK_LAUNCH();
  }
  DevColumn out;
  {
    KernelTimer kt(ctx, "parquet_gather");
    out = take(ctx, S, idx.as<uint32_t>(), H.null_count ? valid.as<uint8_t>() : nullptr, rows, err);
  }
  out.name = H.leaf.name;
  out.nullable = H.leaf.optional;
  out.null_count = H.null_count;
  if (!H.null_count) out.validity = DeviceBuffer();
  return out;
}

}  // namespace pq
}  // namespace prism

struct prism_parquet_file {
  const uint8_t* bytes = nullptr;
  uint64_t n = 0;
  prism::pq::FileMeta meta;
};

struct prism_parquet_table {
  prism::Ctx* ctx = nullptr;
  std::vector<uint32_t> cols;  // file leaf of each output column
  std::vector<int64_t> rows;
  std::vector<std::vector<prism::DevColumn>> batches;
};

using namespace prism;

extern "C" {

prism_status prism_parquet_open(const uint8_t* bytes, uint64_t n_bytes, prism_parquet_file** out) {
  return guarded([&] {
    if (!out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL output");
    *out = nullptr;
    auto f = std::make_unique<prism_parquet_file>();
    f->meta = pq::read_footer(bytes, n_bytes);
    f->bytes = bytes;
    f->n = n_bytes;
    *out = f.release();
  });
}

prism_status prism_parquet_file_info(const prism_parquet_file* f, uint32_t* n_row_groups, uint32_t* n_columns,
                                     int64_t* n_rows) {
  return guarded([&] {
    if (!f) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL file");
    if (n_row_groups) *n_row_groups = static_cast<uint32_t>(f->meta.rgs.size());
    if (n_columns) *n_columns = static_cast<uint32_t>(f->meta.leaves.size());
    if (n_rows) *n_rows = f->meta.num_rows;
  });
}

prism_status prism_parquet_field(const prism_parquet_file* f, uint32_t column, const char** name, int32_t* type,
                                 int32_t* nullable, int32_t* physical_type) {
  return guarded([&] {
    if (!f || column >= f->meta.leaves.size()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such column");
    const pq::Leaf& L = f->meta.leaves[column];
    if (name) *name = L.name.c_str();
    if (type) *type = L.supported ? static_cast<int32_t>(L.type) : 0;
    if (nullable) *nullable = L.optional ? 1 : 0;
    if (physical_type) *physical_type = L.phys;
  });
}

prism_status prism_parquet_row_group(const prism_parquet_file* f, uint32_t row_group, int64_t* n_rows) {
  return guarded([&] {
    if (!f || row_group >= f->meta.rgs.size()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such row group");
    if (n_rows) *n_rows = f->meta.rgs[row_group].num_rows;
  });
}

prism_status prism_parquet_statistics(const prism_parquet_file* f, uint32_t row_group, uint32_t column,
                                      prism_parquet_stats* out) {
  return guarded([&] {
    if (!f || row_group >= f->meta.rgs.size() || column >= f->meta.leaves.size() || !out)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such row group or column");
    const pq::ColMeta& m = f->meta.rgs[row_group].cols[column];
    const pq::Stats& s = m.stats;
    *out = prism_parquet_stats{};
    out->present = s.present ? 1 : 0;
    out->physical_type = m.type;
    out->null_count = s.null_count;
    out->has_min = s.has_min ? 1 : 0;
    out->has_max = s.has_max ? 1 : 0;
    out->min = reinterpret_cast<const uint8_t*>(s.min.data());
    out->min_len = static_cast<uint32_t>(s.min.size());
    out->max = reinterpret_cast<const uint8_t*>(s.max.data());
    out->max_len = static_cast<uint32_t>(s.max.size());
  });
}

void prism_parquet_close(prism_parquet_file* f) { delete f; }

prism_status prism_parquet_read_device(prism_ctx* ctx, const prism_parquet_file* f, const uint32_t* row_groups,
                                       uint32_t n_row_groups, const uint32_t* columns, uint32_t n_columns,
                                       prism_parquet_table** out) {
  return guarded([&] {
    if (!ctx || !f || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context, file or output");
    *out = nullptr;
    const pq::FileMeta& F = f->meta;
    std::vector<uint32_t> rgs, cols;
    if (row_groups) rgs.assign(row_groups, row_groups + n_row_groups);
    else for (uint32_t i = 0; i < F.rgs.size(); ++i) rgs.push_back(i);
    if (columns) cols.assign(columns, columns + n_columns);
    else for (uint32_t i = 0; i < F.leaves.size(); ++i) cols.push_back(i);
    for (uint32_t g : rgs)
      if (g >= F.rgs.size()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no row group " + std::to_string(g));
    for (uint32_t k : cols)
      if (k >= F.leaves.size()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no column " + std::to_string(k));
    use_device(ctx);
    auto T = std::make_unique<prism_parquet_table>();
    T->ctx = as_ctx(ctx);
    T->cols = cols;
    // host part of the chunks on a pool of threads, a wave at a time, the next
    // wave's host part overlapping this wave's device part (chunk order kept)
    const size_t n_chunks = rgs.size() * cols.size();
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (uint32_t g : rgs) {
      T->rows.push_back(F.rgs[g].num_rows);
      T->batches.emplace_back();
    }
    struct Wave {
      size_t base = 0, m = 0;
      std::vector<pq::HostChunk> chunks;
      std::vector<prism_status> codes;
      std::vector<std::string> errs;
      std::vector<std::thread> th;
      void join() {
        for (auto& t : th) t.join();
        th.clear();
      }
      ~Wave() { join(); }
    };
    auto start = [&](Wave& w, size_t base) {
      w.base = base;
      w.m = std::min<size_t>(hw, n_chunks - base);
      w.chunks.assign(w.m, pq::HostChunk());
      w.codes.assign(w.m, PRISM_OK);
      w.errs.assign(w.m, std::string());
      for (size_t i = 0; i < w.m; ++i)
        w.th.emplace_back([&w, &F, &rgs, &cols, f, i] {
          const size_t k = w.base + i;
          try {
            w.chunks[i] = pq::decode_host(f->bytes, f->n, F, rgs[k / cols.size()], cols[k % cols.size()]);
          } catch (const Error& e) {
            w.codes[i] = e.status;
            w.errs[i] = e.what();
          } catch (const std::exception& e) {
            w.codes[i] = PRISM_ERR_INTERNAL;
            w.errs[i] = std::string("Internal error: ") + e.what();
          }
        });
    };
    DeviceBuffer flags(T->ctx, std::max<size_t>(n_chunks, 1) * 4);
    PRISM_HIP(hipMemsetAsync(flags.ptr, 0, std::max<size_t>(n_chunks, 1) * 4, ctx->stream));
    Wave a, b;
    Wave* cur = &a;
    Wave* nxt = &b;
    if (n_chunks) start(*cur, 0);
    for (size_t base = 0; base < n_chunks; base += hw) {
      cur->join();
      for (size_t i = 0; i < cur->m; ++i)
        if (cur->codes[i] != PRISM_OK) fail(cur->codes[i], cur->errs[i]);
      if (base + hw < n_chunks) start(*nxt, base + hw);
      for (size_t i = 0; i < cur->m; ++i) {
        const size_t k = base + i;
        T->batches[k / cols.size()].push_back(pq::decode_device(ctx, cur->chunks[i], flags.as<uint32_t>() + k));
        cur->chunks[i] = pq::HostChunk();
      }
      std::swap(cur, nxt);
    }
    std::vector<uint32_t> hf(n_chunks);
    if (n_chunks)
      PRISM_HIP(hipMemcpyAsync(hf.data(), flags.ptr, n_chunks * 4, hipMemcpyDeviceToHost, ctx->stream));
    PRISM_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < n_chunks; ++k)
      if (hf[k])
        fail(PRISM_ERR_ARROW, "Arrow error: Parquet column '" + F.leaves[cols[k % cols.size()]].name +
                                  "': a dictionary index exceeds the dictionary");
    *out = T.release();
  });
}

prism_status prism_parquet_table_info(const prism_parquet_table* t, uint32_t* n_batches, uint32_t* n_columns) {
  return guarded([&] {
    if (!t) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL table");
    if (n_batches) *n_batches = static_cast<uint32_t>(t->rows.size());
    if (n_columns) *n_columns = static_cast<uint32_t>(t->cols.size());
  });
}

prism_status prism_parquet_table_column(const prism_parquet_table* t, uint32_t batch, uint32_t column,
                                        uint64_t* n_rows, prism_column* out) {
  return guarded([&] {
    if (!t || batch >= t->rows.size() || column >= t->cols.size())
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such batch or column");
    if (n_rows) *n_rows = static_cast<uint64_t>(t->rows[batch]);
    if (out) *out = t->batches[batch][column].desc();
  });
}

prism_status prism_parquet_table_export(const prism_parquet_table* t, uint32_t batch, struct ArrowArray* out,
                                        struct ArrowSchema* out_schema) {
  return guarded([&] {
    if (!t || batch >= t->rows.size() || !out || !out_schema)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: no such batch or NULL output");
    use_device(t->ctx);
    std::vector<OutColumn> cols;
    for (const DevColumn& d : t->batches[batch]) cols.push_back(download(t->ctx, d));
    export_batch(std::move(cols), t->rows[batch], out, out_schema);
  });
}

void prism_parquet_table_free(prism_parquet_table* t) {
  if (!t) return;
  if (t->ctx) (void)hipSetDevice(t->ctx->device);
  delete t;
}

}  // extern "C"
