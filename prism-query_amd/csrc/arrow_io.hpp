// Arrow C Data Interface import/export for the operator layer, and the
// device-side column type that carries a column through HBM.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace prism {

// Arrow types the operators handle (the set prism-executor's operators match
// on: Int32 / Int64 / Float64 / Utf8 / Boolean / Date32, plus UInt32 indices).
enum class DType : int32_t { Int32 = PRISM_TYPE_INT32, Int64 = PRISM_TYPE_INT64, Float64 = PRISM_TYPE_FLOAT64,
                             Utf8 = PRISM_TYPE_UTF8, Bool = PRISM_TYPE_BOOL, Date32 = PRISM_TYPE_DATE32,
                             UInt32 = 100, Decimal128 = 101 };

const char* dtype_format(DType t);
std::string dtype_name(DType t);
int dtype_width(DType t);  // bytes per value (0 for Bool and Utf8)

// A column of an imported RecordBatch (host memory owned by the producer).
struct HostColumn {
  DType type;
  std::string name;
  bool nullable = true;
  int64_t length = 0;
  int64_t offset = 0;       // logical offset into the buffers
  int64_t null_count = 0;
  const uint8_t* validity = nullptr;
  const void* data = nullptr;       // values / utf8 bytes / bool bits
  const int32_t* offsets = nullptr; // utf8
};

struct HostBatch {
  std::vector<HostColumn> cols;
  int64_t n_rows = 0;
};

// Import a RecordBatch exported as a struct array ("+s").  The producer keeps
// ownership; nothing is released here.  `outputs`: also accept the types only
// operators produce (Decimal128(38,0)), for writers of operator results.
HostBatch import_batch(const ArrowArray* array, const ArrowSchema* schema, bool outputs = false);
// Schema only (for empty batch lists).
std::vector<HostColumn> import_schema(const ArrowSchema* schema);

// A column resident in HBM.  Data starts at element `offset & ~7` of the
// source so that `dev_offset` (< 8) indexes data and validity bits alike.
struct DevColumn {
  DType type = DType::Int64;
  std::string name;
  bool nullable = true;
  int64_t length = 0;
  int64_t dev_offset = 0;
  int64_t null_count = 0;
  DeviceBuffer data, offsets, validity;
  prism_column desc() const;
};

// Copy host bytes into HBM through the context's pinned staging (returns once
// the copies are done, so `src` may be reused).
void staged_h2d(prism_ctx* ctx, void* dst, const void* src, size_t bytes);

// Copy a host column into HBM (pinned staging, asynchronous on ctx->stream).
DevColumn upload(prism_ctx* ctx, const HostColumn& c);

// Output column assembled on the host before export.
struct OutColumn {
  DType type = DType::Int64;
  std::string name;
  bool nullable = true;
  int64_t length = 0;
  int64_t null_count = 0;
  int32_t precision = 38, scale = 0;  // Decimal128
  std::vector<uint8_t> validity;      // empty when there are no nulls
  std::vector<uint8_t> data;
  std::vector<int32_t> offsets;       // utf8
};

// Download a device column (offset 0 on device) into an OutColumn.
OutColumn download(prism_ctx* ctx, const DevColumn& c);

// Export columns as a struct array + schema.  Takes ownership of the vectors.
void export_batch(std::vector<OutColumn>&& cols, int64_t n_rows, ArrowArray* out, ArrowSchema* out_schema);
// Export one column as a plain array + schema.
void export_array(OutColumn&& col, ArrowArray* out, ArrowSchema* out_schema);

// Host helpers.
inline bool host_valid(const HostColumn& c, int64_t row) {
  if (!c.validity) return true;
  int64_t i = c.offset + row;
  return (c.validity[i >> 3] >> (i & 7)) & 1;
}

}  // namespace prism
