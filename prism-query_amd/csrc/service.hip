// Device-resident operator service with continuous batching.
//
// Requests (group-by over a registered plan, probe of a registered resident
// build table, filter by a registered predicate) arrive from any thread.  A
// dispatcher thread admits them in arrival order into batches, as in-flight
// slots and HBM free up, and launches each batch on its slot's stream while
// other batches are still running, so the work of requests that arrive while
// others are on the GPU overlaps on the device.
//
// One batch = one H2D copy of its control block (request descriptors and the
// (request, row-range) tiles), then per request kind one persistent kernel
// whose workgroups pull tiles from a queue until it is empty, one kernel that
// finishes the group-by requests (collect + order groups by first row), one
// D2H copy of the results, one event.  Launch sequences recur for a plan
// shape, so they are captured once into a hipGraph and replayed.
//
// Memory: resident build tables are rebuilt from their host copy on demand and
// evicted least-recently-used when the HBM budget needs room (never while a
// batch in flight uses them); every slot owns a fixed working arena, and a
// request joins a batch only while its footprint fits the arena.
//
// Results are identical to the one-request-per-launch operators: group-by
// accumulates in the same exact int128 / f64-order-preserving words, probes
// sum payloads exactly, filters evaluate the same row programs.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gb_jit.hpp"
#include "groupby.hpp"
#include "jit.hpp"
#include "join.hpp"
#include "like_scan.hpp"
#include "program.hpp"

namespace prism {
namespace svc {

using clk = std::chrono::steady_clock;

constexpr int kSB = 256;               // threads per workgroup (group-by kernels assume kBlock)
// Probe workgroups and filter waves add their totals to one of kAccCopies
// copies of the request's accumulator (workgroup or wave index modulo
// kAccCopies), once per request: one address would serialise the atomics.
constexpr int kAccCopies = 16;
// Group-by plans with at most this many slots flush a tile's groups in no time.
constexpr uint64_t kSmallGroupCap = 64;
static_assert(kSB == gb::kBlock, "service group-by tiles use the group-by workgroup size");
constexpr int kMaxReq = 256;           // requests per batch
constexpr int kMaxTiles = 32768;       // tiles per batch
using gb::kFinishCap;                  // group-table slots a finish workgroup orders in LDS
constexpr int kVariants = 10;          // persistent-kernel variants (see Variant)
constexpr int kJitQueues = 16;         // queues of generated group-by tile kernels per batch
constexpr int kGraphMinLaunches = 10;  // launches per batch from which a graph replay pays
constexpr size_t kResultsBytes = 48u << 20;

enum Variant : int {
  V_GB_444P = 0, V_GB_444S, V_GB_448P, V_GB_448S, V_GB_288P, V_GB_288S,
  V_PROBE_DENSE, V_PROBE_HASHED, V_FILT_4, V_FILT_2
};
enum Kind : int { K_GB = 0, K_PROBE = 1, K_FILT = 2 };
static Kind kind_of(int v) { return v <= V_GB_288S ? K_GB : v <= V_PROBE_HASHED ? K_PROBE : K_FILT; }

struct FilterProg {
  int32_t len;
  int32_t like_col;  // >= 0: the program is one LIKE '%literal%' of this column, run as the byte scan
  prism_instr code[kMaxCode];
  StrTable strs;
  LikeArg like;
};

using GbDesc = gb::TileReq;

struct ProbeDesc {
  join::TableView T;
  const int64_t* keys;
  const int64_t* pay;
  uint64_t n;
  unsigned long long* acc;  // kAccCopies x {matches, probe sum lo/hi, build sum lo/hi}
  // pairs (optional): each tile writes its (probe row, build row) pairs in
  // probe order to stage[tile start ...]; a scan over the request's tile
  // counts and a move put them at their final offsets in out_p / out_b
  uint2* stage;
  uint32_t* out_p;
  uint32_t* out_b;
  uint64_t cap;
  uint32_t t_first, t_count;  // the request's tiles in the batch's tile array
};

struct FiltDesc {
  const FilterProg* P;
  const prism_column* cols;
  uint64_t n;
  unsigned long long* mask;   // optional caller bitmap
  unsigned long long* count;  // kAccCopies selected-row counts, then error flags
};

using Tile = gb::RowTile;
using Queue = gb::TileQueue;

struct InitDesc {
  unsigned long long* dst;
  uint64_t n_words;
  const unsigned long long* pattern;  // dst[i] = pattern[i % period], or 0 when NULL
  uint32_t period;
  uint32_t pad;
};

// One request of a batch: its descriptor (by kind), the words its batch
// clears or fills before the tile kernels run, and its column descriptors.
struct ReqRec {
  union {
    GbDesc gb;  // first: the generated group-by tile kernels read it at gb::kTileReqStride steps
    ProbeDesc pr;
    FiltDesc fd;
  };
  InitDesc init[2];  // n_words 0: unused
  prism_column cols[kMaxCols];
};
static_assert(sizeof(ReqRec) <= gb::kTileReqStride, "a request record fits its stride");
static_assert(offsetof(ReqRec, gb) == 0, "the group-by descriptor leads the record");

// The records of a batch's n requests, then its tiles, packed from the start
// of the pool: tile j of the batch is tile (first_tile(n) + j) of the pool
// seen as a tile array, so one host -> device copy of the header and the used
// part of the pool carries the whole batch.
constexpr uint32_t first_tile(uint32_t n_req) {
  return static_cast<uint32_t>((static_cast<uint64_t>(n_req) * gb::kTileReqStride + sizeof(Tile) - 1) / sizeof(Tile));
}
constexpr uint32_t kPoolTiles = first_tile(kMaxReq) + kMaxTiles;

// Control block (device mirror of a pinned host staging block).
struct Control {
  Queue q[kVariants + kJitQueues];  // prebuilt variants, then generated group-by kernels
  uint32_t n_req, n_gb, n_pair, pad0;
  uint32_t gb_idx[kMaxReq];    // group-by requests
  uint32_t pair_req[kMaxReq];  // probe requests that want pairs
  alignas(256) unsigned char pool[static_cast<size_t>(kPoolTiles) * sizeof(Tile)];
  // written on the device only (not part of the host -> device copy)
  uint32_t tile_cnt[kPoolTiles];  // pairs of each probe tile
  uint64_t tile_off[kPoolTiles];  // their first output offset

  __host__ __device__ ReqRec& req(uint32_t i) {
    return *reinterpret_cast<ReqRec*>(pool + static_cast<size_t>(i) * gb::kTileReqStride);
  }
  __host__ __device__ Tile* tiles() { return reinterpret_cast<Tile*>(pool); }
};

// ------------------------------------------------------------------ kernels
using gb::next_tile;

__global__ __launch_bounds__(kSB) void k_svc_init(Control* C) {
  const uint32_t n = 2 * C->n_req;
  for (uint32_t d = blockIdx.x; d < n; d += gridDim.x) {
    const InitDesc D = C->req(d >> 1).init[d & 1];
    for (uint64_t i = threadIdx.x; i < D.n_words; i += kSB) D.dst[i] = D.pattern ? D.pattern[i % D.period] : 0ULL;
  }
}

template <int kRows, int kDepth, int kPref, bool kPriv>
__global__ __launch_bounds__(kSB) void k_svc_groupby(Control* C, int v) {
  extern __shared__ unsigned long long lds[];
  __shared__ uint32_t s_t;
  Queue* q = &C->q[v];
  const uint32_t n_tiles = q->n_tiles;
  bool first = true;  // first tile: the workgroup's own index
  while (true) {
    // uniform per workgroup: scalar loads of the tile and its request descriptor
    const uint32_t t = __builtin_amdgcn_readfirstlane(next_tile(q, &s_t, first));
    if (t >= n_tiles) break;
    const Tile tl = C->tiles()[q->offset + t];
    const GbDesc& d = C->req(tl.req).gb;
    gb::gb_init<kPriv>(d.P, lds, d.lds_slots);
    gb::gb_rows<kRows, kDepth, kPref, kPriv>(d.P, d.cols, tl.start, static_cast<uint64_t>(kSB) * kRows, tl.end, 0,
                                             d.G, lds, d.lds_slots, d.lds_shift, d.flags);
    gb::gb_flush<kPriv>(d.P, d.G, lds, d.lds_slots, d.flags);
    __syncthreads();
  }
}

__device__ __forceinline__ void wave_i128(__int128& v) {
  for (int d = 32; d > 0; d >>= 1) {
    const uint64_t lo = static_cast<uint64_t>(v), hi = static_cast<uint64_t>(v >> 64);
    const uint64_t olo = __shfl_xor(lo, d, 64), ohi = __shfl_xor(hi, d, 64);
    v += static_cast<__int128>((static_cast<unsigned __int128>(ohi) << 64) | olo);
  }
}

// Sums of one request's tiles stay in registers while the workgroup's tiles
// belong to that request; they are added to the request's accumulator copy
// (chosen by workgroup) when the request changes and at the end, so that
// contended carry-propagating atomics are paid once per workgroup and request.
__device__ __forceinline__ void svc_probe_flush(unsigned long long* acc_base, unsigned long long& m, __int128& ps,
                                                __int128& bs, unsigned long long* s_m, __int128* s_p, __int128* s_b) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 32; k > 0; k >>= 1) m += __shfl_xor(m, k, 64);
  wave_i128(ps);
  wave_i128(bs);
  if (lane == 0) {
    s_m[wave] = m;
    s_p[wave] = ps;
    s_b[wave] = bs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    m = 0;
    ps = 0;
    bs = 0;
    for (int w = 0; w < kSB / 64; ++w) {
      m += s_m[w];
      ps += s_p[w];
      bs += s_b[w];
    }
    unsigned long long* acc = acc_base + 5 * (blockIdx.x % kAccCopies);
    if (m) atomicAdd(&acc[0], m);
    if (ps) atomic_add_i128(&acc[1], &acc[2], static_cast<uint64_t>(ps), static_cast<int64_t>(ps >> 64));
    if (bs) atomic_add_i128(&acc[3], &acc[4], static_cast<uint64_t>(bs), static_cast<int64_t>(bs >> 64));
  }
  __syncthreads();
  m = 0;
  ps = 0;
  bs = 0;
}

template <int KIND>
__global__ __launch_bounds__(kSB) void k_svc_probe(Control* C, int v) {
  constexpr int U = 8;  // independent lookups in flight per thread
  constexpr int kW = kSB / 64;
  __shared__ uint32_t s_t;
  __shared__ unsigned long long s_m[kSB / 64];
  __shared__ __int128 s_p[kSB / 64], s_b[kSB / 64];
  __shared__ uint32_t s_cnt[U * kW], s_pre[U * kW], s_tot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  Queue* q = &C->q[v];
  const uint32_t n_tiles = q->n_tiles;
  uint32_t cur = 0xFFFFFFFFu;  // request whose sums the registers hold (uniform per workgroup)
  unsigned long long m = 0;
  __int128 ps = 0, bs = 0;
  bool first = true;  // first tile: the workgroup's own index
  while (true) {
    // uniform per workgroup: scalar loads of the tile and its request descriptor
    const uint32_t t = __builtin_amdgcn_readfirstlane(next_tile(q, &s_t, first));
    if (t >= n_tiles) break;
    const Tile tl = C->tiles()[q->offset + t];
    // the descriptor in registers (a reference into the control block would be
    // re-read after every store), its buffers as global pointers
    const ProbeDesc d = C->req(tl.req).pr;
    if (tl.req != cur) {
      if (cur != 0xFFFFFFFFu) svc_probe_flush(C->req(cur).pr.acc, m, ps, bs, s_m, s_p, s_b);
      cur = tl.req;
    }
    const PRISM_GLOBAL int64_t* keys = (const PRISM_GLOBAL int64_t*)(d.keys);
    const PRISM_GLOBAL int64_t* pay_src = (const PRISM_GLOBAL int64_t*)(d.pay ? d.pay : d.keys);
    PRISM_GLOBAL uint64_t* stage = (PRISM_GLOBAL uint64_t*)(d.stage);  // {probe row, build row} pairs
    if (d.stage) {
      // Pairs: the tile's matches in row order.  Row c0 + u * kSB + tid has
      // rank (u, wave, lane) in the chunk; per-(u, wave) ballot counts are
      // scanned by one wave, so each hit knows its place.
      uint32_t run = 0;
      for (uint64_t c0 = tl.start; c0 < tl.end; c0 += static_cast<uint64_t>(kSB) * U) {
        int64_t key[U], x[U];
        bool hit[U];
        uint32_t brow[U];
        // unconditional loads (rows past the tile read its last row and are
        // masked), so the compiler waits for exactly the loads each use needs
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint64_t r = c0 + static_cast<uint64_t>(u) * kSB + threadIdx.x;
          const uint64_t rc = r < tl.end ? r : tl.end - 1;
          hit[u] = r < tl.end;
          key[u] = __builtin_nontemporal_load(keys + rc);
          x[u] = __builtin_nontemporal_load(pay_src + rc);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          uint32_t val = 0, cnt;
          int32_t pay = 0;
          if (!d.pay) x[u] = 0;
          hit[u] = join::lookup<KIND, true>(d.T, key[u], val, cnt, pay) && hit[u];
          brow[u] = val;
          if (hit[u]) {
            ++m;
            ps += x[u];
            if (d.T.has_payload) bs += pay;
          }
        }
        unsigned long long bal[U];
#pragma unroll
        for (int u = 0; u < U; ++u) bal[u] = __ballot(hit[u]);
        if (lane == 0) {
#pragma unroll
          for (int u = 0; u < U; ++u) s_cnt[u * kW + wave] = static_cast<uint32_t>(__popcll(bal[u]));
        }
        __syncthreads();
        if (wave == 0) {
          const uint32_t own = lane < U * kW ? s_cnt[lane] : 0u;
          uint32_t incl = own;
          for (int k = 1; k < 64; k <<= 1) {
            const uint32_t o = __shfl_up(incl, k, 64);
            if (lane >= k) incl += o;
          }
          if (lane < U * kW) s_pre[lane] = incl - own;
          if (lane == 63) s_tot = incl;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (hit[u]) {
            const uint32_t pos = run + s_pre[u * kW + wave] + static_cast<uint32_t>(__popcll(bal[u] & lt));
            const uint64_t r = c0 + static_cast<uint64_t>(u) * kSB + threadIdx.x;
            stage[tl.start + pos] = static_cast<uint32_t>(r) | (static_cast<uint64_t>(brow[u]) << 32);
          }
        run += s_tot;
        __syncthreads();  // s_pre / s_tot are rewritten by the next chunk
      }
      if (threadIdx.x == 0) C->tile_cnt[q->offset + t] = run;
      continue;
    }
    for (uint64_t r0 = tl.start + threadIdx.x; r0 < tl.end; r0 += static_cast<uint64_t>(kSB) * U) {
      int64_t key[U], x[U];
      bool in[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t r = r0 + static_cast<uint64_t>(u) * kSB;
        const uint64_t rc = r < tl.end ? r : tl.end - 1;
        in[u] = r < tl.end;
        // streamed once: non-temporal, so the table's lines keep their place in
        // L2 / MALL; unconditional (masked) loads, with the key: no dependent load
        key[u] = __builtin_nontemporal_load(keys + rc);
        x[u] = __builtin_nontemporal_load(pay_src + rc);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        uint32_t val, cnt;
        int32_t pay = 0;
        if (!d.pay) x[u] = 0;
        if (join::lookup<KIND, true>(d.T, key[u], val, cnt, pay) && in[u]) {
          ++m;
          ps += x[u];
          if (d.T.has_payload) bs += pay;
        }
      }
    }
  }
  if (cur != 0xFFFFFFFFu) svc_probe_flush(C->req(cur).pr.acc, m, ps, bs, s_m, s_p, s_b);
}

// Exclusive scan of each pairs request's tile counts (one workgroup per request).
__global__ __launch_bounds__(kSB) void k_svc_pair_scan(Control* C) {
  using BS = hipcub::BlockScan<uint64_t, kSB>;
  __shared__ typename BS::TempStorage tmp;
  const uint32_t n_pair = C->n_pair;
  for (uint32_t k = blockIdx.x; k < n_pair; k += gridDim.x) {
    const ProbeDesc& d = C->req(C->pair_req[k]).pr;
    uint64_t carry = 0;
    for (uint32_t b = 0; b < d.t_count; b += kSB) {
      const uint32_t i = b + threadIdx.x;
      const uint64_t v = i < d.t_count ? C->tile_cnt[d.t_first + i] : 0;
      uint64_t ex, agg;
      BS(tmp).ExclusiveSum(v, ex, agg);
      if (i < d.t_count) C->tile_off[d.t_first + i] = carry + ex;
      carry += agg;
      __syncthreads();
    }
  }
}

// Staged pairs to their final offsets (the first `cap` pairs of a request).
// The probe queues' tiles are adjacent in the tile array.
__global__ __launch_bounds__(kSB) void k_svc_pair_move(Control* C) {
  const uint32_t g0 = C->q[V_PROBE_DENSE].offset;
  const uint32_t g1 = C->q[V_PROBE_HASHED].offset + C->q[V_PROBE_HASHED].n_tiles;
  for (uint32_t g = g0 + blockIdx.x; g < g1; g += gridDim.x) {
    const Tile tl = C->tiles()[g];
    const ProbeDesc& d = C->req(tl.req).pr;
    if (!d.stage) continue;
    const uint32_t cnt = C->tile_cnt[g];
    const uint64_t off = C->tile_off[g];
    for (uint32_t i = threadIdx.x; i < cnt; i += kSB) {
      const uint64_t o = off + i;
      if (o >= d.cap) break;
      const uint2 p = d.stage[tl.start + i];
      d.out_p[o] = p.x;
      d.out_b[o] = p.y;
    }
  }
}

// Filter tiles start on 64-row boundaries: every wave step covers whole words.
template <int R, int DEPTH>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(4))) void k_svc_filter(Control* C, int v) {
  static_assert(kSB == kLsThreads, "the LIKE scan uses the filter workgroup");
  __shared__ uint32_t s_t;
  __shared__ uint32_t s_bits[kLsBitWords];
  Queue* q = &C->q[v];
  const uint32_t n_tiles = q->n_tiles;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t copy = (blockIdx.x * (kSB / 64) + wave) % kAccCopies;
  const uint64_t pre[1][R] = {};
  bool first = true;  // first tile: the workgroup's own index
  // selected rows of one request, counted per wave across that request's
  // tiles and added to its count copy when the request changes and at the end
  uint32_t cur = 0xFFFFFFFFu;
  unsigned long long sel = 0;
  while (true) {
    // uniform per workgroup: scalar loads of the tile and its request descriptor
    const uint32_t t = __builtin_amdgcn_readfirstlane(next_tile(q, &s_t, first));
    if (t >= n_tiles) break;
    const Tile tl = C->tiles()[q->offset + t];
    const FiltDesc& d = C->req(tl.req).fd;
    if (tl.req != cur) {
      if (cur != 0xFFFFFFFFu && lane == 0 && sel) atomicAdd(C->req(cur).fd.count + copy, sel);
      sel = 0;
      cur = tl.req;
    }
    if (d.P->like_col >= 0) {  // uniform: one LIKE '%literal%', as a byte scan
      unsigned long long nulls = 0;
      for (uint64_t r0 = tl.start; r0 < tl.end; r0 += kLsRows)
        like_scan_tile(d.cols[d.P->like_col], r0, r0 + kLsRows < tl.end ? r0 + kLsRows : tl.end, d.P->like, s_bits,
                       d.mask, nullptr, nulls, sel);
      continue;
    }
    const uint64_t w_end = ceil_div(tl.end, 64);
    for (uint64_t w0 = tl.start / 64 + static_cast<uint64_t>(wave) * R; w0 < w_end; w0 += (kSB / 64) * R) {
      uint64_t row[R];
      bool in[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        row[r] = (w0 + r) * 64 + lane;
        in[r] = row[r] < tl.end;
      }
      uint64_t val[R];
      uint32_t valid;
      run_program<R, DEPTH, 1>(d.P->code, d.P->len, d.cols, &d.P->strs, row, in, pre, 0u, val, valid,
                               reinterpret_cast<uint32_t*>(d.count + kAccCopies));
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (w0 + r >= w_end) break;
        const unsigned long long bw = __ballot(in[r] && ((valid >> r) & 1) && (val[r] & 1));
        sel += __popcll(bw);
        if (lane == 0 && d.mask) d.mask[w0 + r] = bw;
      }
    }
  }
  if (cur != 0xFFFFFFFFu && lane == 0 && sel) atomicAdd(C->req(cur).fd.count + copy, sel);
}

// Group-by finish: one workgroup per request collects the occupied slots,
// orders them by first row (bitonic sort in LDS) and writes the rows.
__global__ __launch_bounds__(gb::kFinishThreads) void k_svc_gb_finish(Control* C) {
  __shared__ unsigned long long s_first[kFinishCap];
  __shared__ uint32_t s_slot[kFinishCap];
  __shared__ uint32_t s_n;
  const uint32_t n_gb = C->n_gb;
  for (uint32_t k = blockIdx.x; k < n_gb; k += gridDim.x) {
    const GbDesc& d = C->req(C->gb_idx[k]).gb;
    const uint32_t n = gb::order_groups(d.G, d.out + 2, s_first, s_slot, &s_n);
    if (threadIdx.x == 0) {
      d.out[0] = n;
      d.out[1] = *d.flags;
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- host side
struct Table {
  uint32_t id = 0;
  std::vector<int64_t> keys;
  std::vector<int32_t> payload;
  prism_join_table* t = nullptr;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  int pins = 0;
};

struct GbPlan {
  prism_groupby* g = nullptr;
  int variant = 0;
  bool priv = false;
  int lds_slots = 0, lds_shift = 0;
  size_t lds_bytes = 0;
  uint64_t cap = 0;
  // generated tile kernels (one per set of nullable columns): their form,
  // rows per lane and shared-form LDS table
  int jit_form = 0, jit_rows = 4;
  int jit_slots = 0, jit_shift = 0;
  size_t jit_lds = 0;
};

struct FiltPlan {
  DeviceBuffer prog;
  int variant = V_FILT_4;
  std::vector<int32_t> col_types;
};

struct Req {
  uint64_t ticket = 0;
  prism_request r{};
  std::vector<prism_column> cols;
  clk::time_point t_submit, t_launch;
  prism_request_result res{};
  std::vector<unsigned long long> rows;  // group-by result rows
  GbPlan* gbp = nullptr;
  Table* table = nullptr;
  unsigned long long* d_out = nullptr;   // device results of this request (in the slot's results area)
  size_t out_words = 0;
  size_t res_off = 0;                    // word offset in the slot's results area
  bool done = false;
};

struct Cand {
  uint32_t tile_rows;
  uint32_t bpc;  // workgroups per CU
};

struct Tuner {
  std::vector<Cand> cands;
  std::vector<double> rows_per_us;
  std::vector<int> samples;
  int best = -1;
  int next = 0;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  hipEvent_t ev_start = nullptr;
  clk::time_point t_admit, t_launched;
  uint64_t batch_rows = 0;
  DeviceBuffer arena;      // results area, then group tables
  DeviceBuffer control;    // Control
  Control* h_control = nullptr;
  unsigned long long* h_results = nullptr;
  size_t results_words = 0, table_words = 0;
  std::vector<Req*> reqs;
  bool busy = false;
  uint64_t batch_id = 0;
  std::map<std::string, hipGraphExec_t> graphs;
  std::map<std::string, int> shapes_seen;  // launch shapes issued without a graph so far
  // tuning: (variant, events, rows, candidate)
  struct Timed {
    int variant;
    hipEvent_t a, b;
    uint64_t rows;
    int cand;
  };
  std::vector<Timed> timed;
  hipEvent_t ev_kern = nullptr;       // after the kernels, before the results copy
  Slot() = default;
  Slot(Slot&&) noexcept {}
};

}  // namespace svc
}  // namespace prism

struct prism_service {
  void wake();
  prism::Ctx* ctx = nullptr;           // owned context: table builds, allocations
  prism_service_config cfg{};
  std::vector<prism::svc::Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  // How long an idle dispatcher polls for submissions before it sleeps
  // (PRISM_SVC_IDLE_SPIN_US, default 5 ms; 0 = sleep at once).  Polling keeps
  // one host core busy while requests come at least every 5 ms and saves the
  // sleeping thread's wake-up: served p50 0.249/0.237/0.239 ms ->
  // 0.236/0.229/0.232 ms at 200/500/1000 req/s.
  uint64_t idle_spin_us = [] {
    const char* v = std::getenv("PRISM_SVC_IDLE_SPIN_US");
    return v ? std::strtoull(v, nullptr, 10) : 5000ULL;
  }();
  std::deque<prism::svc::Req*> pending;
  std::unordered_map<uint64_t, prism::svc::Req*> reqs;
  std::map<uint32_t, prism::svc::Table> tables;
  std::map<std::string, uint32_t> table_keys;  // store keys (PartitionStore style) -> table ids
  uint32_t next_keyed_id = 0x80000000u;        // ids given to keyed tables
  std::map<uint32_t, prism::svc::GbPlan> gb_plans;
  std::map<uint32_t, prism::svc::FiltPlan> filt_plans;
  prism::svc::Tuner tuner[3];
  uint64_t next_ticket = 1, use_clock = 0, batches = 0, built = 0, evicted = 0, graph_launches = 0;
  uint64_t completed = 0, submitted = 0, batch_reqs = 0;
  uint64_t resident_bytes = 0;
  std::vector<double> latencies;
  double sum_gpu_ms = 0, max_gpu_ms = 0, sum_launch_ms = 0, sum_detect_ms = 0, sum_batch_rows = 0, sum_kern_ms = 0;
  bool stop = false;
  std::mutex wake_mu;
  bool woke = false;
  std::thread worker;
  std::string fatal;
};

void prism_service::wake() {
  { std::lock_guard<std::mutex> l(wake_mu); woke = true; }
  cv_work.notify_one();
}

namespace prism {
namespace svc {

static double ms_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

static int gb_variant(const prism_groupby* g, bool priv) {
  const bool shallow = g->max_depth <= 4;
  if (shallow && g->plan.n_pref <= 4) return priv ? V_GB_444P : V_GB_444S;
  if (shallow) return priv ? V_GB_448P : V_GB_448S;
  return priv ? V_GB_288P : V_GB_288S;
}

static Cand cand_of(prism_service* S, Kind k) {
  Tuner& T = S->tuner[k];
  if (T.best >= 0) return T.cands[T.best];
  if (!S->cfg.autotune) return T.cands[0];
  return T.cands[T.next % T.cands.size()];
}

static int cand_index(prism_service* S, Kind k) {
  Tuner& T = S->tuner[k];
  if (T.best >= 0) return T.best;
  if (!S->cfg.autotune) return 0;
  return static_cast<int>(T.next % T.cands.size());
}

// Make a table resident: evict least-recently-used unpinned tables until it fits.
static bool make_resident(prism_service* S, Table& T) {
  T.last_use = ++S->use_clock;
  if (T.t) return true;
  const uint64_t need = T.keys.size() * 16 + (64u << 20);  // generous bound of the device table
  const uint64_t arenas = S->slots.size() * (S->cfg.arena_bytes + sizeof(Control));
  while (S->resident_bytes + need + arenas > S->cfg.hbm_budget) {
    Table* victim = nullptr;
    for (auto& kv : S->tables)
      if (kv.second.t && kv.second.pins == 0 && (!victim || kv.second.last_use < victim->last_use)) victim = &kv.second;
    if (!victim) return false;  // wait until a batch releases its tables
    S->resident_bytes -= victim->bytes;
    prism_join_table_free(victim->t);
    victim->t = nullptr;
    victim->bytes = 0;
    ++S->evicted;
  }
  Ctx* c = S->ctx;
  DeviceBuffer k(c, T.keys.size() * 8 + 8), p;
  PRISM_HIP(hipMemcpyAsync(k.ptr, T.keys.data(), T.keys.size() * 8, hipMemcpyHostToDevice, c->stream));
  if (!T.payload.empty()) {
    p.alloc(c, T.payload.size() * 4 + 4);
    PRISM_HIP(hipMemcpyAsync(p.ptr, T.payload.data(), T.payload.size() * 4, hipMemcpyHostToDevice, c->stream));
  }
  prism_join_table* t = nullptr;
  prism_status st = T.payload.empty()
                        ? prism_join_build_i64(c, k.as<int64_t>(), T.keys.size(), PRISM_TABLE_AUTO, &t)
                        : prism_join_build_i64_payload(c, k.as<int64_t>(), p.as<int32_t>(), T.keys.size(),
                                                       PRISM_TABLE_AUTO, &t);
  if (st != PRISM_OK) fail(st, prism_last_error());
  PRISM_HIP(hipStreamSynchronize(c->stream));
  if (!t->info.unique) {
    prism_join_table_free(t);
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: service build tables need unique keys");
  }
  T.t = t;
  T.bytes = t->info.device_bytes;
  S->resident_bytes += T.bytes;
  ++S->built;
  return true;
}

// Words a request needs in the slot's results area and table area.
static void footprint(const Req* R, size_t& res_words, size_t& table_words) {
  res_words = 0;
  table_words = 0;
  switch (R->r.kind) {
    case PRISM_REQ_GROUPBY: {
      const uint64_t stride = static_cast<uint64_t>(R->gbp->g->plan.stride);
      res_words = 2 + (R->gbp->cap + 1) * stride;
      table_words = (R->gbp->cap + 1) * stride + 2;  // + n_groups counter, flags
      break;
    }
    case PRISM_REQ_PROBE:
      res_words = 5 * kAccCopies;
      if (R->r.probe_idx) table_words = R->r.n_rows + 32;  // staged pairs, 8 B per row
      break;
    default:
      res_words = kAccCopies + 1;
  }
}

// Grid sizes rounded up to a power of two: batches of similar size share a
// launch shape (and its graph); the extra workgroups find the queue empty.
static uint64_t pow2_grid(uint64_t n) { return n <= 1 ? 1 : next_pow2(n); }

static std::string shape_key(const int (&grid)[kVariants], const size_t (&lds)[kVariants], int n_init_grid,
                             int n_fin_grid, const std::vector<hipFunction_t>& jf, const std::vector<int>& jgrid,
                             const std::vector<size_t>& jlds) {
  std::string k;
  for (int v = 0; v < kVariants; ++v) k += std::to_string(grid[v]) + ":" + std::to_string(lds[v]) + ",";
  for (size_t i = 0; i < jf.size(); ++i)
    k += std::to_string(reinterpret_cast<uintptr_t>(jf[i])) + "/" + std::to_string(jgrid[i]) + ":" +
         std::to_string(jlds[i]) + ",";
  return k + std::to_string(n_init_grid) + "," + std::to_string(n_fin_grid);
}

static void launch_variant(int v, dim3 grid, size_t lds, hipStream_t st, Control* dC) {
  switch (v) {
    case V_GB_444P: hipLaunchKernelGGL((k_svc_groupby<4, 4, 4, true>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_GB_444S: hipLaunchKernelGGL((k_svc_groupby<4, 4, 4, false>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_GB_448P: hipLaunchKernelGGL((k_svc_groupby<4, 4, 8, true>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_GB_448S: hipLaunchKernelGGL((k_svc_groupby<4, 4, 8, false>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_GB_288P: hipLaunchKernelGGL((k_svc_groupby<2, kStackDepth, 8, true>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_GB_288S: hipLaunchKernelGGL((k_svc_groupby<2, kStackDepth, 8, false>), grid, dim3(kSB), lds, st, dC, v); break;
    case V_PROBE_DENSE: hipLaunchKernelGGL(k_svc_probe<PRISM_TABLE_DENSE>, grid, dim3(kSB), 0, st, dC, v); break;
    case V_PROBE_HASHED: hipLaunchKernelGGL(k_svc_probe<PRISM_TABLE_HASHED>, grid, dim3(kSB), 0, st, dC, v); break;
    case V_FILT_4: hipLaunchKernelGGL((k_svc_filter<4, 4>), grid, dim3(kSB), 0, st, dC, v); break;
    default: hipLaunchKernelGGL((k_svc_filter<2, kStackDepth>), grid, dim3(kSB), 0, st, dC, v); break;
  }
}


// Build the batch's control block, launch it on the slot's stream.
static void launch_batch(prism_service* S, Slot& sl) {
  Control* H = sl.h_control;
  Control* dC = sl.control.as<Control>();
  memset(H, 0, offsetof(Control, pool));
  const int n_cu = S->ctx->num_cus;
  std::vector<Tile> vt[kVariants];
  uint64_t vrows[kVariants] = {};
  // Tile rows per kind: the tuned size, shrunk for small batches so that every
  // workgroup slot of the persistent grid gets several tiles.
  uint64_t kind_rows[3] = {};
  for (Req* R : sl.reqs)
    kind_rows[R->r.kind == PRISM_REQ_GROUPBY ? K_GB : R->r.kind == PRISM_REQ_PROBE ? K_PROBE : K_FILT] += R->r.n_rows;
  uint64_t tile_of[3];
  // the shared LDS form flushes every group slot per tile: keep its tiles
  // larger when a plan has many slots (a few slots flush in no time)
  bool gb_big_flush = false;
  for (Req* R : sl.reqs)
    gb_big_flush = gb_big_flush || (R->r.kind == PRISM_REQ_GROUPBY && (!R->gbp->priv || jit::enabled()) &&
                                    R->gbp->cap > kSmallGroupCap);
  // probe tiles may be small (a workgroup flushes its sums once per request,
  // not per tile): small requests still spread over every workgroup
  const uint64_t min_tile[3] = {gb_big_flush ? 65536u : 16384u, 1024, 4096};
  const uint64_t per_wg[3] = {4, 2, 4};
  for (int k = 0; k < 3; ++k) {
    const Cand cd = cand_of(S, static_cast<Kind>(k));
    const uint64_t want = kind_rows[k] / (static_cast<uint64_t>(n_cu) * cd.bpc * per_wg[k]);
    tile_of[k] = std::min<uint64_t>(cd.tile_rows, std::max<uint64_t>(min_tile[k], want));
  }
  size_t lds[kVariants] = {};
  // generated group-by tile kernels of this batch: one queue per kernel
  std::vector<hipFunction_t> jf;
  std::vector<std::vector<Tile>> jt;
  std::vector<uint64_t> jrows;
  std::vector<size_t> jlds;
  std::vector<uint64_t> jcap;  // largest group capacity per generated kernel
  unsigned long long* res_base = sl.arena.as<unsigned long long>();
  unsigned long long* tab_base = res_base + kResultsBytes / 8;
  size_t res_used = 0, tab_used = 0;
  std::vector<int> pair_variant;  // probe queue of each pairs request
  const uint32_t n_req = static_cast<uint32_t>(sl.reqs.size());
  H->n_req = n_req;
  for (uint32_t k = 0; k < n_req; ++k) {
    Req* R = sl.reqs[k];
    size_t rw, tw;
    footprint(R, rw, tw);
    R->res_off = res_used;
    R->d_out = res_base + res_used;
    R->out_words = rw;
    res_used += rw;
    ReqRec& rec = H->req(k);
    memset(&rec, 0, sizeof(ReqRec));
    // device address of the record's column descriptors
    const prism_column* d_cols = dC->req(k).cols;
    switch (R->r.kind) {
      case PRISM_REQ_GROUPBY: {
        const GbPlan& gp = *R->gbp;
        GbDesc& d = rec.gb;
        const int stride = gp.g->plan.stride;
        unsigned long long* table = tab_base + tab_used;
        tab_used += tw;
        memcpy(rec.cols, R->cols.data(), R->cols.size() * sizeof(prism_column));
        d.P = gp.g->d_plan.as<gb::DevPlan>();
        d.cols = d_cols;
        d.n = R->r.n_rows;
        d.G.w = table;
        d.G.cap = gp.cap;
        d.G.shift = 64 - log2_pow2(gp.cap);
        d.G.stride = stride;
        d.G.n_groups = table + (gp.cap + 1) * stride;
        d.flags = reinterpret_cast<uint32_t*>(table + (gp.cap + 1) * stride + 1);
        d.out = R->d_out;
        uint32_t nullable = 0;
        for (size_t c = 0; c < R->cols.size(); ++c)
          if (R->cols[c].validity) nullable |= 1u << c;
        hipFunction_t tf = gb::jit_kernels(gp.g, gp.jit_form, gp.jit_rows, nullable).tiles;
        int qi = -1;
        if (tf) {
          for (size_t i = 0; i < jf.size(); ++i)
            if (jf[i] == tf) qi = static_cast<int>(i);
          if (qi < 0 && jf.size() < static_cast<size_t>(kJitQueues)) {
            qi = static_cast<int>(jf.size());
            jf.push_back(tf);
            jt.emplace_back();
            jrows.push_back(0);
            jlds.push_back(0);
            jcap.push_back(0);
          }
          if (qi >= 0) jcap[qi] = std::max<uint64_t>(jcap[qi], gp.cap);
        }
        d.lds_slots = qi >= 0 ? gp.jit_slots : gp.lds_slots;
        d.lds_shift = qi >= 0 ? gp.jit_shift : gp.lds_shift;
        H->gb_idx[H->n_gb++] = k;
        InitDesc& a = rec.init[0];
        a.dst = table;
        a.n_words = (gp.cap + 1) * stride;
        a.pattern = reinterpret_cast<const unsigned long long*>(gp.g->d_plan.as<char>() + offsetof(gb::DevPlan, tmpl));
        a.period = static_cast<uint32_t>(stride);
        InitDesc& b = rec.init[1];
        b.dst = table + (gp.cap + 1) * stride;
        b.n_words = 2;
        const uint64_t step =
            static_cast<uint64_t>(kSB) * (qi >= 0 ? gp.jit_rows : (gp.variant >= V_GB_288P ? 2 : 4));
        const uint64_t tr = std::max<uint64_t>(step, tile_of[K_GB] / step * step);
        std::vector<Tile>& dst = qi >= 0 ? jt[qi] : vt[gp.variant];
        for (uint64_t s0 = 0; s0 < R->r.n_rows; s0 += tr)
          dst.push_back(Tile{k, 0, s0, std::min<uint64_t>(R->r.n_rows, s0 + tr)});
        if (qi >= 0) {
          jrows[qi] += R->r.n_rows;
          jlds[qi] = std::max(jlds[qi], gp.jit_lds);
        } else {
          vrows[gp.variant] += R->r.n_rows;
          lds[gp.variant] = std::max(lds[gp.variant], gp.lds_bytes);
        }
        break;
      }
      case PRISM_REQ_PROBE: {
        ProbeDesc& d = rec.pr;
        d.T = R->table->t->view;
        d.keys = R->r.probe_keys;
        d.pay = R->r.probe_payload;
        d.n = R->r.n_rows;
        d.acc = R->d_out;
        d.stage = nullptr;
        d.out_p = d.out_b = nullptr;
        d.cap = 0;
        InitDesc& a = rec.init[0];
        a.dst = R->d_out;
        a.n_words = 5 * kAccCopies;
        const int v = R->table->t->info.kind == PRISM_TABLE_DENSE ? V_PROBE_DENSE : V_PROBE_HASHED;
        const uint64_t tr = std::max<uint64_t>(kSB, tile_of[K_PROBE]);
        d.t_first = static_cast<uint32_t>(vt[v].size());  // local to the queue until offsets are known
        for (uint64_t s0 = 0; s0 < R->r.n_rows; s0 += tr)
          vt[v].push_back(Tile{k, 0, s0, std::min<uint64_t>(R->r.n_rows, s0 + tr)});
        d.t_count = static_cast<uint32_t>(vt[v].size()) - d.t_first;
        if (R->r.probe_idx && R->r.n_rows) {
          d.stage = reinterpret_cast<uint2*>(tab_base + tab_used);
          tab_used += tw;
          d.out_p = R->r.probe_idx;
          d.out_b = R->r.build_idx;
          d.cap = R->r.pair_capacity;
          pair_variant.push_back(v);
          H->pair_req[H->n_pair++] = k;
        }
        vrows[v] += R->r.n_rows;
        break;
      }
      default: {
        const FiltPlan& fp = S->filt_plans[R->r.target];
        FiltDesc& d = rec.fd;
        memcpy(rec.cols, R->cols.data(), R->cols.size() * sizeof(prism_column));
        d.P = fp.prog.as<FilterProg>();
        d.cols = d_cols;
        d.n = R->r.n_rows;
        d.mask = R->r.out_mask;
        d.count = R->d_out;
        InitDesc& a = rec.init[0];
        a.dst = R->d_out;
        a.n_words = kAccCopies + 1;
        const uint64_t tr = std::max<uint64_t>(64 * 64, tile_of[K_FILT] / 64 * 64);
        for (uint64_t s0 = 0; s0 < R->r.n_rows; s0 += tr)
          vt[fp.variant].push_back(Tile{k, 0, s0, std::min<uint64_t>(R->r.n_rows, s0 + tr)});
        vrows[fp.variant] += R->r.n_rows;
      }
    }
  }
  // tiles follow the request records in the pool
  const uint32_t t0 = first_tile(n_req);
  uint32_t off = t0;
  int grid[kVariants] = {};
  for (int v = 0; v < kVariants; ++v) {
    H->q[v].n_tiles = static_cast<uint32_t>(vt[v].size());
    memset(H->q[v].next, 0, sizeof(H->q[v].next));
    H->q[v].offset = off;
    if (!vt[v].empty()) {
      if (off - t0 + vt[v].size() > static_cast<size_t>(kMaxTiles))
        fail(PRISM_ERR_INTERNAL, "Internal error: batch tile overflow");
      memcpy(&H->tiles()[off], vt[v].data(), vt[v].size() * sizeof(Tile));
      off += static_cast<uint32_t>(vt[v].size());
      const Cand cd = cand_of(S, kind_of(v));
      uint32_t bpc = cd.bpc;
      if (lds[v]) bpc = std::min<uint32_t>(bpc, std::max<uint32_t>(1, static_cast<uint32_t>((160u << 10) / lds[v])));
      // tiles past the first round are claimed per residue of the workgroup
      // index mod kQueueLanes: such grids hold every residue
      grid[v] = static_cast<int>(std::min<uint64_t>(pow2_grid(vt[v].size()),
                                                    std::max<uint64_t>(static_cast<uint64_t>(n_cu) * bpc,
                                                                       gb::kQueueLanes)));
    }
  }
  std::vector<int> jgrid(jf.size(), 0);
  for (size_t i = 0; i < jf.size(); ++i) {
    Queue& q = H->q[kVariants + i];
    q.n_tiles = static_cast<uint32_t>(jt[i].size());
    memset(q.next, 0, sizeof(q.next));
    q.offset = off;
    if (off - t0 + jt[i].size() > static_cast<size_t>(kMaxTiles))
      fail(PRISM_ERR_INTERNAL, "Internal error: batch tile overflow");
    memcpy(&H->tiles()[off], jt[i].data(), jt[i].size() * sizeof(Tile));
    off += static_cast<uint32_t>(jt[i].size());
    const Cand cd = cand_of(S, K_GB);
    uint32_t bpc = cd.bpc;
    if (jcap[i] <= kSmallGroupCap) bpc = std::max<uint32_t>(bpc, 4);  // cheap flushes: fill the CUs
    if (jlds[i]) bpc = std::min<uint32_t>(bpc, std::max<uint32_t>(1, static_cast<uint32_t>((160u << 10) / jlds[i])));
    jgrid[i] = static_cast<int>(std::min<uint64_t>(pow2_grid(jt[i].size()),
                                                   std::max<uint64_t>(static_cast<uint64_t>(n_cu) * bpc,
                                                                      gb::kQueueLanes)));
  }
  for (uint32_t i = 0; i < H->n_pair; ++i)  // tile indices of pairs requests, now global
    H->req(H->pair_req[i]).pr.t_first += H->q[pair_variant[i]].offset;
  const int pair_grid = H->n_pair ? static_cast<int>(pow2_grid(H->n_pair)) : 0;
  const uint32_t probe_tiles = H->q[V_PROBE_DENSE].n_tiles + H->q[V_PROBE_HASHED].n_tiles;
  const int move_grid =
      H->n_pair ? static_cast<int>(std::min<uint64_t>(pow2_grid(probe_tiles), static_cast<uint64_t>(n_cu) * 8)) : 0;
  // the header, the request records and the tiles after them
  const size_t ctl_bytes = offsetof(Control, pool) + static_cast<size_t>(off) * sizeof(Tile);
  hipStream_t st = sl.st;
  uint64_t brows = 0;
  for (int v = 0; v < kVariants; ++v) brows += vrows[v];
  for (uint64_t r : jrows) brows += r;
  sl.batch_rows = brows;
  PRISM_HIP(hipEventRecord(sl.ev_start, st));
  PRISM_HIP(hipMemcpyAsync(dC, H, ctl_bytes, hipMemcpyHostToDevice, st));
  const int init_grid = static_cast<int>(std::min<uint64_t>(pow2_grid(2 * n_req), static_cast<uint64_t>(n_cu) * 4));
  const int fin_grid = static_cast<int>(std::min<uint64_t>(pow2_grid(H->n_gb), static_cast<uint64_t>(n_cu) * 2));
  // kinds launched by this batch; only a batch holding a kind still being
  // tuned is timed per launch (and kept out of graphs)
  unsigned kinds = jf.empty() ? 0u : 1u << K_GB;
  for (int v = 0; v < kVariants; ++v)
    if (grid[v]) kinds |= 1u << kind_of(v);
  bool timed = false;
  if (S->cfg.autotune)
    for (int k = 0; k < 3; ++k)
      if (((kinds >> k) & 1) && S->tuner[k].best < 0) timed = true;
  auto issue = [&](bool with_events) {
    hipLaunchKernelGGL(k_svc_init, dim3(init_grid), dim3(kSB), 0, st, dC);
    for (int v = 0; v < kVariants; ++v) {
      if (!grid[v]) continue;
      hipEvent_t a = nullptr, b = nullptr;
      if (with_events) {
        PRISM_HIP(hipEventCreate(&a));
        PRISM_HIP(hipEventCreate(&b));
        PRISM_HIP(hipEventRecord(a, st));
      }
      launch_variant(v, dim3(grid[v]), lds[v], st, dC);
      if (with_events) {
        PRISM_HIP(hipEventRecord(b, st));
        sl.timed.push_back(Slot::Timed{v, a, b, vrows[v], cand_index(S, kind_of(v))});
      }
    }
    Tile* d_tiles = dC->tiles();
    GbDesc* d_gb = &dC->req(0).gb;
    for (size_t i = 0; i < jf.size(); ++i) {
      hipEvent_t a = nullptr, b = nullptr;
      if (with_events) {
        PRISM_HIP(hipEventCreate(&a));
        PRISM_HIP(hipEventCreate(&b));
        PRISM_HIP(hipEventRecord(a, st));
      }
      Queue* d_q = &dC->q[kVariants + i];
      void* args[] = {&d_q, &d_tiles, &d_gb};
      PRISM_HIP(hipModuleLaunchKernel(jf[i], static_cast<unsigned>(jgrid[i]), 1, 1, kSB, 1, 1,
                                      static_cast<unsigned>(jlds[i]), st, args, nullptr));
      if (with_events) {
        PRISM_HIP(hipEventRecord(b, st));
        sl.timed.push_back(Slot::Timed{V_GB_444S, a, b, jrows[i], cand_index(S, K_GB)});
      }
    }
    if (H->n_gb) hipLaunchKernelGGL(k_svc_gb_finish, dim3(fin_grid), dim3(gb::kFinishThreads), 0, st, dC);
    if (pair_grid) {
      hipLaunchKernelGGL(k_svc_pair_scan, dim3(pair_grid), dim3(kSB), 0, st, dC);
      hipLaunchKernelGGL(k_svc_pair_move, dim3(move_grid), dim3(kSB), 0, st, dC);
    }
  };
  // A launch sequence of at least kGraphMinLaunches kernels is captured into
  // a graph the second time its shape comes (a shape met once would pay
  // capture + instantiation for nothing).  Fewer launches go directly: a
  // graph launch costs the host ~29 us, a kernel launch ~2.6 us (median HIP
  // API times of the served stream), so a batch of one or two requests
  // (2-4 kernels) launched from a graph waited longer.
  int n_launch = 1 + static_cast<int>(jf.size()) + (H->n_gb ? 1 : 0) + (pair_grid ? 2 : 0);
  for (int v = 0; v < kVariants; ++v) n_launch += grid[v] ? 1 : 0;
  const std::string key = S->cfg.use_graphs && !timed && (S->cfg.use_graphs >= 2 || n_launch >= kGraphMinLaunches)
                              ? shape_key(grid, lds, init_grid, fin_grid, jf, jgrid, jlds) + "," +
                                    std::to_string(pair_grid) + "," + std::to_string(move_grid)
                              : std::string();
  const bool recurring = !key.empty() && (sl.graphs.count(key) || ++sl.shapes_seen[key] >= 2);
  if (recurring) {
    auto it = sl.graphs.find(key);
    if (it == sl.graphs.end()) {
      hipGraph_t g;
      PRISM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      issue(false);
      PRISM_HIP(hipStreamEndCapture(st, &g));
      hipGraphExec_t ex;
      PRISM_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      PRISM_HIP(hipGraphDestroy(g));
      it = sl.graphs.emplace(key, ex).first;
    }
    PRISM_HIP(hipGraphLaunch(it->second, st));
    ++S->graph_launches;
  } else {
    issue(timed);
  }
  PRISM_CHECK_LAUNCH();
  PRISM_HIP(hipEventRecord(sl.ev_kern, st));
  if (res_used) PRISM_HIP(hipMemcpyAsync(sl.h_results, res_base, res_used * 8, hipMemcpyDeviceToHost, st));
  sl.results_words = res_used;
  sl.table_words = tab_used;
  PRISM_HIP(hipEventRecord(sl.ev, st));
  sl.t_launched = clk::now();
  if (S->cfg.autotune)
    for (int k = 0; k < 3; ++k)
      if (((kinds >> k) & 1) && S->tuner[k].best < 0) ++S->tuner[k].next;
}

// Finish a completed batch (called with the service lock held).
static void finalize(prism_service* S, Slot& sl) {
  const auto now = clk::now();
  float gms = 0;
  if (hipEventElapsedTime(&gms, sl.ev_start, sl.ev) == hipSuccess) {
    S->sum_gpu_ms += gms;
    S->max_gpu_ms = std::max<double>(S->max_gpu_ms, gms);
  }
  S->sum_launch_ms += ms_between(sl.t_admit, sl.t_launched);
  // host wall time from the launch call's return to this point beyond the
  // batch's device time (the device may start before the launch call returns)
  S->sum_detect_ms += std::max(0.0, ms_between(sl.t_launched, now) - gms);
  float kms = 0;
  if (hipEventElapsedTime(&kms, sl.ev_start, sl.ev_kern) == hipSuccess) S->sum_kern_ms += kms;
  S->sum_batch_rows += static_cast<double>(sl.batch_rows);
  for (Req* R : sl.reqs) {
    const unsigned long long* o = sl.h_results + R->res_off;
    prism_request_result& res = R->res;
    res.status = PRISM_OK;
    res.kind = R->r.kind;
    res.n_rows = R->r.n_rows;
    res.batch_id = sl.batch_id;
    switch (R->r.kind) {
      case PRISM_REQ_GROUPBY: {
        const uint64_t n = o[0];
        const uint32_t flags = static_cast<uint32_t>(o[1]);
        res.count = n;
        if (flags & gb::kFlagOverflow) res.status = PRISM_ERR_EXECUTION;
        else if (flags & (gb::kFlagDropped | gb::kFlagGrow)) res.status = PRISM_ERR_RESOURCE;
        const size_t stride = static_cast<size_t>(R->gbp->g->plan.stride);
        R->rows.assign(o + 2, o + 2 + n * stride);
        break;
      }
      case PRISM_REQ_PROBE: {
        unsigned long long cnt = 0;
        unsigned __int128 ps = 0, bs = 0;
        for (int c = 0; c < kAccCopies; ++c) {
          const unsigned long long* a = o + 5 * c;
          cnt += a[0];
          ps += (static_cast<unsigned __int128>(a[2]) << 64) | a[1];
          bs += (static_cast<unsigned __int128>(a[4]) << 64) | a[3];
        }
        res.count = cnt;
        res.probe_sum_lo = static_cast<uint64_t>(ps);
        res.probe_sum_hi = static_cast<int64_t>(static_cast<uint64_t>(ps >> 64));
        res.build_sum_lo = static_cast<uint64_t>(bs);
        res.build_sum_hi = static_cast<int64_t>(static_cast<uint64_t>(bs >> 64));
        if (R->table) --R->table->pins;
        break;
      }
      default: {
        unsigned long long cnt = 0;
        for (int c = 0; c < kAccCopies; ++c) cnt += o[c];
        res.count = cnt;
        if (o[kAccCopies] & gb::kFlagOverflow) res.status = PRISM_ERR_EXECUTION;
      }
    }
    res.latency_ms = ms_between(R->t_submit, now);
    res.queue_ms = ms_between(R->t_submit, R->t_launch);
    S->latencies.push_back(res.latency_ms);
    R->done = true;
    ++S->completed;
  }
  for (auto& tm : sl.timed) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, tm.a, tm.b) == hipSuccess && ms > 0) {
      Tuner& T = S->tuner[kind_of(tm.variant)];
      if (T.best < 0 && tm.cand >= 0 && tm.cand < static_cast<int>(T.cands.size())) {
        T.rows_per_us[tm.cand] += tm.rows / (ms * 1e3);
        ++T.samples[tm.cand];
        bool all = true;
        for (int s : T.samples) all = all && s >= 2;
        if (all) {
          int best = 0;
          for (size_t i = 1; i < T.cands.size(); ++i)
            if (T.rows_per_us[i] / T.samples[i] > T.rows_per_us[best] / T.samples[best]) best = static_cast<int>(i);
          T.best = best;
        }
      }
    }
    (void)hipEventDestroy(tm.a);
    (void)hipEventDestroy(tm.b);
  }
  sl.timed.clear();
  sl.reqs.clear();
  sl.busy = false;
}

static bool admit_into(prism_service* S, Slot& sl) {
  size_t res_used = 0, tab_used = 0;
  uint64_t rows = 0;
  uint32_t tiles = 0;
  const size_t res_cap = kResultsBytes / 8, tab_cap = (S->cfg.arena_bytes - kResultsBytes) / 8;
  while (!S->pending.empty() && sl.reqs.size() < S->cfg.max_batch_requests) {
    Req* R = S->pending.front();
    size_t rw, tw;
    footprint(R, rw, tw);
    if (rw > res_cap || tw > tab_cap) {  // can never fit: fail it
      S->pending.pop_front();
      R->res.status = PRISM_ERR_RESOURCE;
      R->res.kind = R->r.kind;
      R->done = true;
      ++S->completed;
      continue;
    }
    const uint64_t est_tiles = R->r.n_rows / 65536 + 1;
    if (!sl.reqs.empty() && (res_used + rw > res_cap || tab_used + tw > tab_cap ||
                             rows + R->r.n_rows > S->cfg.max_batch_rows || tiles + est_tiles > kMaxTiles / 4))
      break;
    if (R->r.kind == PRISM_REQ_PROBE) {
      Table& T = S->tables[R->r.target];
      bool ok = false;
      try {
        ok = make_resident(S, T);
      } catch (const Error& e) {
        S->pending.pop_front();
        R->res.status = e.status;
        R->res.kind = R->r.kind;
        R->done = true;
        ++S->completed;
        continue;
      }
      if (!ok) {
        bool busy = false;
        for (const Slot& o : S->slots) busy = busy || o.busy;
        if (busy || !sl.reqs.empty()) break;  // memory held by batches in flight: retry when they finish
        S->pending.pop_front();                // nothing to wait for: the table can never fit
        R->res.status = PRISM_ERR_RESOURCE;
        R->res.kind = R->r.kind;
        R->done = true;
        ++S->completed;
        continue;
      }
      ++T.pins;
      R->table = &T;
    }
    S->pending.pop_front();
    res_used += rw;
    tab_used += tw;
    rows += R->r.n_rows;
    tiles += static_cast<uint32_t>(est_tiles);
    R->t_launch = clk::now();
    sl.reqs.push_back(R);
  }
  return !sl.reqs.empty();
}

static void dispatcher(prism_service* S) {
  (void)hipSetDevice(S->ctx->device);
  std::unique_lock<std::mutex> lock(S->mu);
  while (true) {
    bool progress = false;
    for (Slot& sl : S->slots)
      if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
        PRISM_HIP(hipEventSynchronize(sl.ev));
        finalize(S, sl);
        progress = true;
      }
    if (progress) S->cv_done.notify_all();
    for (Slot& sl : S->slots) {
      if (sl.busy || S->pending.empty()) continue;
      try {
        sl.t_admit = clk::now();
        if (admit_into(S, sl)) {
          sl.batch_id = ++S->batches;
          S->batch_reqs += sl.reqs.size();
          launch_batch(S, sl);
          sl.busy = true;
          progress = true;
        }
      } catch (const std::exception& e) {
        // End a capture the throw interrupted and let whatever was queued on
        // the slot's stream finish before the requests' buffers go back to
        // their callers.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(sl.st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(sl.st, &g);
          if (g) (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
        (void)hipStreamSynchronize(sl.st);
        const Error* pe = dynamic_cast<const Error*>(&e);
        const prism_status code = pe ? pe->status : PRISM_ERR_INTERNAL;
        for (Req* R : sl.reqs) {
          R->res.status = code;
          R->res.kind = R->r.kind;
          R->done = true;
          if (R->table) --R->table->pins;
          ++S->completed;
        }
        sl.reqs.clear();
        // only a failed HIP call stops the service; a batch the host refused
        // (validation) fails alone
        if (!pe || pe->status == PRISM_ERR_EXECUTION) S->fatal = e.what();
        S->cv_done.notify_all();
      }
    }
    bool any_busy = false;
    for (Slot& sl : S->slots) any_busy = any_busy || sl.busy;
    if (S->stop && S->pending.empty() && !any_busy) break;
    if (!progress) {
      lock.unlock();
      if (any_busy) {
        // Batches in flight: poll their events (0.1 us a query) and the submit
        // flag, yielding between polls, for up to 200 us.  A completion
        // callback on the stream (hipLaunchHostFunc) cost 13-24 us of the
        // dispatcher's time per batch and woke it no sooner.
        const auto until = clk::now() + std::chrono::microseconds(200);
        bool ready = false;
        while (!ready && clk::now() < until) {
          {
            std::lock_guard<std::mutex> wl(S->wake_mu);
            ready = S->woke;
            S->woke = false;
          }
          for (Slot& sl : S->slots)  // busy and ev are the dispatcher's own
            ready = ready || (sl.busy && hipEventQuery(sl.ev) == hipSuccess);
          if (!ready) std::this_thread::yield();
        }
      } else {
        // idle: poll the submit flag for spin_us, then sleep until submit()
        // wakes the dispatcher (a sleeping thread's wake-up costs tens of us)
        const auto until = clk::now() + std::chrono::microseconds(S->idle_spin_us);
        bool ready = false;
        while (!ready && S->idle_spin_us && clk::now() < until) {
          {
            std::lock_guard<std::mutex> wl(S->wake_mu);
            ready = S->woke;
            S->woke = false;
          }
          if (!ready) std::this_thread::yield();
        }
        if (!ready) {
          std::unique_lock<std::mutex> wl(S->wake_mu);
          S->cv_work.wait_for(wl, std::chrono::milliseconds(5), [&] { return S->woke; });
          S->woke = false;
        }
      }
      lock.lock();
    }
  }
}

}  // namespace svc
}  // namespace prism

using namespace prism;
using namespace prism::svc;

extern "C" {

prism_status prism_service_create(prism_ctx* ctx, const prism_service_config* cfg, prism_service** out) {
  return guarded([&] {
    if (!ctx || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL context or output");
    *out = nullptr;
    auto S = std::make_unique<prism_service>();
    prism_service_config c{};
    if (cfg) c = *cfg;
    if (!c.hbm_budget) c.hbm_budget = 64ULL << 30;
    if (!c.max_inflight) c.max_inflight = 3;
    if (c.max_inflight > 8) c.max_inflight = 8;
    if (!c.max_batch_requests || c.max_batch_requests > static_cast<uint32_t>(kMaxReq)) c.max_batch_requests = kMaxReq;
    if (!c.max_batch_rows) c.max_batch_rows = 256ULL << 20;
    if (!c.arena_bytes) c.arena_bytes = 256ULL << 20;
    if (c.arena_bytes < kResultsBytes + (16u << 20)) c.arena_bytes = kResultsBytes + (16u << 20);
    S->cfg = c;
    prism_ctx* own = nullptr;
    prism_status st = prism_ctx_create(ctx->device, &own);
    if (st != PRISM_OK) fail(st, prism_last_error());
    S->ctx = as_ctx(own);
    use_device(own);
    S->slots.resize(c.max_inflight);
    for (Slot& sl : S->slots) {
      PRISM_HIP(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
      PRISM_HIP(hipEventCreate(&sl.ev));
      PRISM_HIP(hipEventCreate(&sl.ev_start));
      PRISM_HIP(hipEventCreate(&sl.ev_kern));
      sl.arena.alloc(S->ctx, c.arena_bytes);
      sl.control.alloc(S->ctx, sizeof(Control));
      PRISM_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_control), sizeof(Control), hipHostMallocDefault));
      PRISM_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_results), kResultsBytes, hipHostMallocDefault));
    }
    S->tuner[K_GB].cands = {{131072, 2}, {524288, 2}, {2097152, 4}};
    S->tuner[K_PROBE].cands = {{262144, 4}, {1048576, 8}, {4194304, 8}};
    S->tuner[K_FILT].cands = {{131072, 4}, {524288, 8}, {2097152, 8}};
    for (auto& t : S->tuner) {
      t.rows_per_us.assign(t.cands.size(), 0.0);
      t.samples.assign(t.cands.size(), 0);
    }
    prism_service* raw = S.get();
    S->worker = std::thread([raw] { dispatcher(raw); });
    *out = S.release();
  });
}

void prism_service_destroy(prism_service* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->wake();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->ctx->device);
  for (Slot& sl : s->slots) {
    (void)hipStreamSynchronize(sl.st);
    for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    if (sl.h_control) (void)hipHostFree(sl.h_control);
    if (sl.h_results) (void)hipHostFree(sl.h_results);
    sl.arena.release();
    sl.control.release();
    (void)hipEventDestroy(sl.ev);
    (void)hipEventDestroy(sl.ev_start);
    (void)hipEventDestroy(sl.ev_kern);
    (void)hipStreamDestroy(sl.st);
  }
  for (auto& kv : s->tables)
    if (kv.second.t) prism_join_table_free(kv.second.t);
  for (auto& kv : s->gb_plans) prism_groupby_destroy(kv.second.g);
  s->filt_plans.clear();
  for (auto& kv : s->reqs) delete kv.second;
  prism_ctx* own = s->ctx;
  delete s;
  prism_ctx_destroy(own);
}

// A table no request waits for or runs on (callers hold s->mu).
static void check_unused(prism_service* s, uint32_t id) {
  auto it = s->tables.find(id);
  if (it != s->tables.end() && it->second.pins)
    fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: table in use by requests in flight");
  for (const prism::svc::Req* R : s->pending)
    if (R->r.kind == PRISM_REQ_PROBE && R->r.target == id)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: table in use by requests waiting to run");
}

static void drop_table(prism_service* s, uint32_t id) {
  auto it = s->tables.find(id);
  if (it == s->tables.end()) return;
  if (it->second.t) {
    s->resident_bytes -= it->second.bytes;
    prism_join_table_free(it->second.t);
  }
  s->tables.erase(it);
}

prism_status prism_service_register_table(prism_service* s, uint32_t table_id, const int64_t* h_keys,
                                          const int32_t* h_payload, uint64_t n) {
  return guarded([&] {
    if (!s || (n && !h_keys)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service or keys");
    if (table_id >= 0x80000000u)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: table ids from 2^31 up belong to keyed tables");
    if (n >= 0xFFFFFFFFULL) fail(PRISM_ERR_JOIN, "Join error: build side exceeds u32 row indices");
    std::lock_guard<std::mutex> lock(s->mu);
    Table& T = s->tables[table_id];
    if (T.t && T.pins) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: table in use by requests in flight");
    if (T.t) {
      s->resident_bytes -= T.bytes;
      prism_join_table_free(T.t);
      T.t = nullptr;
    }
    T.id = table_id;
    T.keys.assign(h_keys, h_keys + n);
    T.payload.clear();
    if (h_payload) T.payload.assign(h_payload, h_payload + n);
    use_device(s->ctx);
    make_resident(s, T);
  });
}

prism_status prism_service_put_table(prism_service* s, const char* key, const int64_t* h_keys, const int32_t* h_payload,
                                     uint64_t n, uint32_t* table_id) {
  return guarded([&] {
    if (!s || !key || (n && !h_keys)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service, key or keys");
    std::lock_guard<std::mutex> lock(s->mu);
    auto kt = s->table_keys.find(key);
    const bool fresh = kt == s->table_keys.end();
    uint32_t id = 0;
    if (fresh) {
      if (s->next_keyed_id == 0xFFFFFFFFu) fail(PRISM_ERR_RESOURCE, "Resource error: keyed table ids exhausted");
      id = s->next_keyed_id++;
    } else {
      id = kt->second;
      check_unused(s, id);
    }
    prism::svc::Table& T = s->tables[id];
    if (!fresh && !T.keys.empty() && T.payload.empty() != (h_payload == nullptr))
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: a batch put under a key must match the key's payload column");
    if (T.keys.size() + n >= 0xFFFFFFFFULL) fail(PRISM_ERR_JOIN, "Join error: build side exceeds u32 row indices");
    const size_t old_n = T.keys.size();
    if (T.t) {
      s->resident_bytes -= T.bytes;
      prism_join_table_free(T.t);
      T.t = nullptr;
      T.bytes = 0;
    }
    T.id = id;
    T.keys.insert(T.keys.end(), h_keys, h_keys + n);
    if (h_payload) T.payload.insert(T.payload.end(), h_payload, h_payload + n);
    use_device(s->ctx);
    try {
      make_resident(s, T);
    } catch (...) {  // e.g. the batch repeats a key: the key keeps its earlier rows
      T.keys.resize(old_n);
      if (h_payload) T.payload.resize(old_n);
      if (fresh) s->tables.erase(id);
      throw;
    }
    if (fresh) s->table_keys[key] = id;
    if (table_id) *table_id = id;
  });
}

prism_status prism_service_table_id(prism_service* s, const char* key, uint32_t* table_id) {
  return guarded([&] {
    if (!s || !key || !table_id) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service, key or output");
    std::lock_guard<std::mutex> lock(s->mu);
    auto kt = s->table_keys.find(key);
    if (kt == s->table_keys.end()) fail(PRISM_ERR_INVALID_ARGUMENT, std::string("Invalid argument: no table under key '") + key + "'");
    *table_id = kt->second;
  });
}

prism_status prism_service_clear_table(prism_service* s, const char* key) {
  return guarded([&] {
    if (!s || !key) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service or key");
    std::lock_guard<std::mutex> lock(s->mu);
    auto kt = s->table_keys.find(key);
    if (kt == s->table_keys.end()) return;  // clearing an absent key is a no-op
    check_unused(s, kt->second);
    drop_table(s, kt->second);
    s->table_keys.erase(kt);
  });
}

prism_status prism_service_drop_exchange(prism_service* s, const char* exchange_id) {
  return guarded([&] {
    if (!s || !exchange_id) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service or exchange id");
    std::lock_guard<std::mutex> lock(s->mu);
    const std::string prefix = std::string("exchange/") + exchange_id + "/";
    std::vector<std::map<std::string, uint32_t>::iterator> hit;
    for (auto it = s->table_keys.lower_bound(prefix); it != s->table_keys.end() && it->first.compare(0, prefix.size(), prefix) == 0; ++it)
      hit.push_back(it);
    for (auto& it : hit) check_unused(s, it->second);  // all or nothing
    for (auto& it : hit) {
      drop_table(s, it->second);
      s->table_keys.erase(it);
    }
  });
}

prism_status prism_service_register_groupby(prism_service* s, uint32_t plan_id, const prism_groupby_plan* plan,
                                            uint64_t max_groups) {
  return guarded([&] {
    if (!s || !plan) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service or plan");
    if (max_groups == 0 || max_groups > 1024)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: service group-by plans hold 1..1024 groups");
    prism_groupby* g = nullptr;
    prism_status st = prism_groupby_create(s->ctx, plan, max_groups, &g);
    if (st != PRISM_OK) fail(st, prism_last_error());
    GbPlan P;
    P.g = g;
    P.cap = std::max<uint64_t>(64, next_pow2(2 * max_groups));
    const uint64_t pw = static_cast<uint64_t>(g->plan.stride - gb::kW_COUNT);
    const uint64_t ps = next_pow2(max_groups > 2 ? max_groups : 2);
    const size_t priv_bytes = (2 * ps + ps * pw * kSB) * 8;
    if (ps <= 16 && priv_bytes <= gb::kPrivBudget) {
      P.priv = true;
      P.lds_slots = static_cast<int>(ps);
      P.lds_shift = 64 - log2_pow2(ps);
      P.lds_bytes = priv_bytes;
    } else {
      const size_t sb = static_cast<size_t>(g->plan.stride) * 8;
      uint64_t sl = next_pow2(2 * max_groups);
      while (sl * sb > gb::kLdsBudget && sl > 16) sl >>= 1;
      P.priv = false;
      P.lds_slots = static_cast<int>(sl);
      P.lds_shift = 64 - log2_pow2(sl);
      P.lds_bytes = sl * sb;
    }
    P.variant = gb_variant(g, P.priv);
    P.jit_form = g->plan.n_keys == 0 ? gb::kJitWave : gb::kJitShared;
    P.jit_rows = g->plan.n_pref > 6 ? 2 : 4;
    P.jit_lds = gb::shared_slots(g, max_groups, P.jit_slots, P.jit_shift);
    (void)gb::jit_kernels(g, P.jit_form, P.jit_rows, 0);  // null-free columns: compiled now
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->gb_plans.find(plan_id);
    if (it != s->gb_plans.end()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: plan id already registered");
    s->gb_plans[plan_id] = P;
  });
}

prism_status prism_service_register_filter(prism_service* s, uint32_t plan_id, const prism_instr* code, uint32_t n_code,
                                           const prism_string* strings, uint32_t n_strings, const int32_t* column_types,
                                           uint32_t n_columns) {
  return guarded([&] {
    if (!s || !code || (n_columns && !column_types)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL argument");
    if (n_columns > static_cast<uint32_t>(kMaxCols)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: at most 16 columns");
    int depth = 1;
    const ValType vt = check_program(code, n_code, column_types, n_columns, n_strings, &depth);
    if (vt != VT_B && vt != VT_NULL) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: filter program is not boolean");
    auto P = std::make_unique<FilterProg>();
    memset(P.get(), 0, sizeof(FilterProg));
    P->len = static_cast<int32_t>(n_code);
    memcpy(P->code, code, n_code * sizeof(prism_instr));
    build_str_table(strings, n_strings, P->strs);
    P->like_col = like_scan_prepare(code, n_code, P->strs, P->like);
    use_device(s->ctx);
    FiltPlan F;
    F.prog.alloc(s->ctx, sizeof(FilterProg));
    PRISM_HIP(hipMemcpyAsync(F.prog.ptr, P.get(), sizeof(FilterProg), hipMemcpyHostToDevice, s->ctx->stream));
    PRISM_HIP(hipStreamSynchronize(s->ctx->stream));
    F.variant = depth <= 4 ? V_FILT_4 : V_FILT_2;
    F.col_types.assign(column_types, column_types + n_columns);
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->filt_plans.count(plan_id)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: plan id already registered");
    s->filt_plans[plan_id] = std::move(F);
  });
}

prism_status prism_service_submit(prism_service* s, const prism_request* req, uint64_t* ticket) {
  return guarded([&] {
    if (!s || !req || !ticket) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service/request/ticket");
    auto R = std::make_unique<Req>();
    R->r = *req;
    R->t_submit = clk::now();
    std::lock_guard<std::mutex> lock(s->mu);
    if (!s->fatal.empty()) fail(PRISM_ERR_EXECUTION, "Execution error: service stopped: " + s->fatal);
    switch (req->kind) {
      case PRISM_REQ_GROUPBY: {
        auto it = s->gb_plans.find(req->target);
        if (it == s->gb_plans.end()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown group-by plan");
        const size_t nc = it->second.g->col_types.size();
        if (nc && !req->columns) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL columns");
        for (size_t i = 0; i < nc; ++i)
          if (req->columns[i].type != it->second.g->col_types[i])
            fail(PRISM_ERR_SCHEMA_MISMATCH, "Schema mismatch: column " + std::to_string(i) + " type differs from the plan");
        R->cols.assign(req->columns, req->columns + nc);
        R->gbp = &it->second;
        break;
      }
      case PRISM_REQ_PROBE:
        if (!s->tables.count(req->target)) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown table");
        if (req->n_rows && !req->probe_keys) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL probe keys");
        if (req->n_rows >= 0xFFFFFFFFULL) fail(PRISM_ERR_JOIN, "Join error: probe rows exceed u32 row indices");
        if ((req->probe_idx == nullptr) != (req->build_idx == nullptr))
          fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: give both pair outputs or neither");
        break;
      case PRISM_REQ_FILTER: {
        auto it = s->filt_plans.find(req->target);
        if (it == s->filt_plans.end()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown filter plan");
        const size_t nc = it->second.col_types.size();
        if (nc && !req->columns) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL columns");
        for (size_t i = 0; i < nc; ++i)
          if (req->columns[i].type != it->second.col_types[i])
            fail(PRISM_ERR_SCHEMA_MISMATCH, "Schema mismatch: column " + std::to_string(i) + " type differs from the plan");
        R->cols.assign(req->columns, req->columns + nc);
        break;
      }
      default:
        fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown request kind");
    }
    R->ticket = s->next_ticket++;
    *ticket = R->ticket;
    Req* raw = R.release();
    s->reqs[raw->ticket] = raw;
    s->pending.push_back(raw);
    s->wake();
    ++s->submitted;
  });
}

static Req* find_req(prism_service* s, uint64_t ticket) {
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: unknown ticket");
  return it->second;
}

prism_status prism_service_poll(prism_service* s, uint64_t ticket, int32_t* done) {
  return guarded([&] {
    if (!s || !done) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service/output");
    std::lock_guard<std::mutex> lock(s->mu);
    *done = find_req(s, ticket)->done ? 1 : 0;
  });
}

prism_status prism_service_wait(prism_service* s, uint64_t ticket, prism_request_result* result) {
  return guarded([&] {
    if (!s || !result) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service/result");
    std::unique_lock<std::mutex> lock(s->mu);
    Req* R = find_req(s, ticket);
    s->cv_done.wait(lock, [&] { return R->done; });
    *result = R->res;
  });
}

prism_status prism_service_read_groups(prism_service* s, uint64_t ticket, const prism_groupby_output* out) {
  return guarded([&] {
    if (!s || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service/output");
    std::lock_guard<std::mutex> lock(s->mu);
    Req* R = find_req(s, ticket);
    if (!R->done || R->r.kind != PRISM_REQ_GROUPBY)
      fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: not a finished group-by request");
    gb::convert_rows(R->gbp->g, R->rows.data(), R->res.count, out);
  });
}

prism_status prism_service_release(prism_service* s, uint64_t ticket) {
  return guarded([&] {
    if (!s) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service");
    std::lock_guard<std::mutex> lock(s->mu);
    Req* R = find_req(s, ticket);
    if (!R->done) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: request still running");
    s->reqs.erase(ticket);
    delete R;
  });
}

prism_status prism_service_drain(prism_service* s) {
  return guarded([&] {
    if (!s) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service");
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv_done.wait(lock, [&] { return s->completed >= s->submitted; });
  });
}

prism_status prism_service_stats_get(prism_service* s, prism_service_stats* out) {
  return guarded([&] {
    if (!s || !out) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service/output");
    std::lock_guard<std::mutex> lock(s->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->batches = s->batches;
    out->tables_built = s->built;
    out->tables_evicted = s->evicted;
    out->resident_table_bytes = s->resident_bytes;
    out->graph_launches = s->graph_launches;
    out->mean_batch_requests = s->batches ? static_cast<double>(s->batch_reqs) / s->batches : 0.0;
    const double nb = s->batches ? static_cast<double>(s->batches) : 1.0;
    out->mean_batch_gpu_ms = s->sum_gpu_ms / nb;
    out->max_batch_gpu_ms = s->max_gpu_ms;
    out->mean_batch_launch_ms = s->sum_launch_ms / nb;
    out->mean_batch_detect_ms = s->sum_detect_ms / nb;
    out->mean_batch_rows = s->sum_batch_rows / nb;
    out->mean_batch_kernel_ms = s->sum_kern_ms / nb;
    std::vector<double> l = s->latencies;
    if (!l.empty()) {
      auto pct = [&](double p) {
        size_t k = static_cast<size_t>(p * (l.size() - 1) + 0.5);
        std::nth_element(l.begin(), l.begin() + k, l.end());
        return l[k];
      };
      out->p50_ms = pct(0.50);
      out->p99_ms = pct(0.99);
    }
    for (int k = 0; k < 3; ++k) {
      const Tuner& T = s->tuner[k];
      const int i = T.best >= 0 ? T.best : 0;
      out->tile_rows[k] = T.cands[i].tile_rows;
      out->blocks_per_cu[k] = T.cands[i].bpc;
      out->tuned[k] = T.best >= 0 ? 1 : 0;
    }
  });
}

prism_status prism_service_reset_stats(prism_service* s) {
  return guarded([&] {
    if (!s) fail(PRISM_ERR_INVALID_ARGUMENT, "Invalid argument: NULL service");
    std::lock_guard<std::mutex> lock(s->mu);
    s->latencies.clear();
    s->batches = 0;
    s->batch_reqs = 0;
    s->sum_gpu_ms = s->max_gpu_ms = s->sum_launch_ms = s->sum_detect_ms = s->sum_batch_rows = s->sum_kern_ms = 0;
  });
}

}  // extern "C"
