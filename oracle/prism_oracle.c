/*
 * CPU restatement of the reference's hash join / hash aggregation / LIKE
 * (skn-dd/prism-query, native/prism-executor/src/).  TEST INFRASTRUCTURE ONLY:
 * the checker for tests/, smoke() and the CPU-baseline leg of bench.py.
 */
#include "prism_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

#define FX_K 0x517cc1b727220a95ULL

/* ------------------------------------------------------------------------- */
/* Open-addressing map u64 -> int64 (grows at 1/2 load).                      */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint64_t* keys;
  int64_t* vals;
  uint8_t* used;
  int64_t cap, n;
} map64;

static void map_init(map64* m, int64_t cap) {
  int64_t c = 16;
  while (c < cap * 2) c <<= 1;
  m->cap = c;
  m->n = 0;
  m->keys = (uint64_t*)calloc((size_t)c, 8);
  m->vals = (int64_t*)calloc((size_t)c, 8);
  m->used = (uint8_t*)calloc((size_t)c, 1);
}
static void map_free(map64* m) {
  free(m->keys);
  free(m->vals);
  free(m->used);
}
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static int64_t* map_get(const map64* m, uint64_t k) {
  int64_t s = (int64_t)(mix64(k) & (uint64_t)(m->cap - 1));
  while (m->used[s]) {
    if (m->keys[s] == k) return &m->vals[s];
    s = (s + 1) & (m->cap - 1);
  }
  return NULL;
}
static void map_grow(map64* m);
/* returns pointer to value; *fresh = 1 when inserted */
static int64_t* map_put(map64* m, uint64_t k, int* fresh) {
  if ((m->n + 1) * 2 > m->cap) map_grow(m);
  int64_t s = (int64_t)(mix64(k) & (uint64_t)(m->cap - 1));
  while (m->used[s]) {
    if (m->keys[s] == k) {
      *fresh = 0;
      return &m->vals[s];
    }
    s = (s + 1) & (m->cap - 1);
  }
  m->used[s] = 1;
  m->keys[s] = k;
  m->vals[s] = 0;
  m->n++;
  *fresh = 1;
  return &m->vals[s];
}
static void map_grow(map64* m) {
  map64 o = *m;
  map_init(m, o.cap);
  for (int64_t i = 0; i < o.cap; ++i)
    if (o.used[i]) {
      int f;
      *map_put(m, o.keys[i], &f) = o.vals[i];
    }
  map_free(&o);
}

/* ------------------------------------------------------------------------- */
/* Column access.                                                            */
/* ------------------------------------------------------------------------- */
static inline int is_valid(const orc_column* c, int64_t r) { return c->valid == NULL || c->valid[r]; }

static inline int64_t int_at(const orc_column* c, int64_t r) {
  switch (c->type) {
    case ORC_INT64: return ((const int64_t*)c->data)[r];
    case ORC_BOOL: return ((const uint8_t*)c->data)[r] ? 1 : 0;
    default: return ((const int32_t*)c->data)[r];
  }
}

/* arrow cast(col, Float64) followed by .value(row): the values buffer is read
 * even on null slots (hash_aggregate.rs:227-245,264-268). */
static inline double f64_at(const orc_column* c, int64_t r) {
  if (c->type == ORC_FLOAT64) return ((const double*)c->data)[r];
  return (double)int_at(c, r);
}

/* FxHasher (hash_aggregate.rs:24-75) + hash_array_value (492-504). */
static inline uint64_t fx(uint64_t st, uint64_t w) { return st * FX_K + w; }

static uint64_t hash_value(const orc_column* c, int64_t r, uint64_t st) {
  if (!is_valid(c, r)) return fx(st, 0); /* 0u8.hash -> write_u8 -> write_u64(0) */
  switch (c->type) {
    case ORC_INT64: return fx(st, (uint64_t)((const int64_t*)c->data)[r]);
    case ORC_INT32: return fx(st, (uint64_t)(int64_t)((const int32_t*)c->data)[r]); /* i32 as u64: sign-extend */
    case ORC_FLOAT64: {
      uint64_t b;
      memcpy(&b, &((const double*)c->data)[r], 8);
      return fx(st, b);
    }
    case ORC_UTF8: { /* str::hash: write(bytes) in 8-byte LE chunks, then write_u8(0xff) */
      const uint8_t* s = (const uint8_t*)c->data + c->offsets[r];
      int32_t n = c->offsets[r + 1] - c->offsets[r];
      for (int32_t p = 0; p < n; p += 8) {
        uint64_t w = 0;
        int32_t m = n - p < 8 ? n - p : 8;
        for (int32_t k = 0; k < m; ++k) w |= (uint64_t)s[p + k] << (8 * k);
        st = fx(st, w);
      }
      return fx(st, 0xff);
    }
    /* Date32 / Boolean hash a Debug string in the reference (format!, 502); only
     * the partition into groups is observable, and any injective word that
     * avoids the null word 0 gives the same partition. */
    case ORC_DATE32: return fx(st, (uint64_t)(uint32_t)((const int32_t*)c->data)[r] | (1ULL << 32));
    case ORC_BOOL: return fx(st, 2 + (((const uint8_t*)c->data)[r] ? 1 : 0));
  }
  return st;
}

uint64_t orc_group_hash(const orc_batch* b, const int32_t* key_cols, int32_t n_keys, int64_t row) {
  uint64_t st = 0;
  for (int32_t k = 0; k < n_keys; ++k) st = hash_value(&b->columns[key_cols[k]], row, st);
  return st;
}

/* ------------------------------------------------------------------------- */
/* COUNT DISTINCT sets (DistinctValue, hash_aggregate.rs:112-121,457-481).    */
/* ------------------------------------------------------------------------- */
typedef struct {
  int64_t group;
  int32_t tag;
  int32_t slen;
  uint64_t v;
  const uint8_t* s;
} dentry;

typedef struct {
  dentry* e;
  uint8_t* used;
  int64_t cap, n;
} dset;

static void dset_init(dset* d) {
  d->cap = 1024;
  d->n = 0;
  d->e = (dentry*)calloc((size_t)d->cap, sizeof(dentry));
  d->used = (uint8_t*)calloc((size_t)d->cap, 1);
}
static uint64_t dhash(const dentry* x) {
  uint64_t h = mix64((uint64_t)x->group * 31 + (uint64_t)x->tag);
  if (x->s) {
    for (int32_t i = 0; i < x->slen; ++i) h = mix64(h ^ x->s[i]);
    h = mix64(h ^ (uint64_t)x->slen);
  } else {
    h = mix64(h ^ x->v);
  }
  return h;
}
static int deq(const dentry* a, const dentry* b) {
  if (a->group != b->group || a->tag != b->tag) return 0;
  if (a->s || b->s) return a->slen == b->slen && memcmp(a->s, b->s, (size_t)a->slen) == 0;
  return a->v == b->v;
}
static int dset_insert(dset* d, const dentry* x);
static void dset_grow(dset* d) {
  dset o = *d;
  d->cap *= 2;
  d->n = 0;
  d->e = (dentry*)calloc((size_t)d->cap, sizeof(dentry));
  d->used = (uint8_t*)calloc((size_t)d->cap, 1);
  for (int64_t i = 0; i < o.cap; ++i)
    if (o.used[i]) dset_insert(d, &o.e[i]);
  free(o.e);
  free(o.used);
}
/* returns 1 if new */
static int dset_insert(dset* d, const dentry* x) {
  if ((d->n + 1) * 2 > d->cap) dset_grow(d);
  int64_t s = (int64_t)(dhash(x) & (uint64_t)(d->cap - 1));
  while (d->used[s]) {
    if (deq(&d->e[s], x)) return 0;
    s = (s + 1) & (d->cap - 1);
  }
  d->used[s] = 1;
  d->e[s] = *x;
  d->n++;
  return 1;
}

/* distinct_value (hash_aggregate.rs:457-481): None for null, ±0 and NaN folded. */
static int distinct_of(const orc_column* c, int64_t r, int64_t group, dentry* out) {
  if (!is_valid(c, r)) return 0;
  memset(out, 0, sizeof(*out));
  out->group = group;
  out->tag = c->type;
  switch (c->type) {
    case ORC_FLOAT64: {
      double v = ((const double*)c->data)[r];
      uint64_t b;
      if (v == 0.0) v = 0.0;
      if (isnan(v)) b = 0x7ff8000000000000ULL;
      else memcpy(&b, &v, 8);
      out->v = b;
      break;
    }
    case ORC_UTF8:
      out->s = (const uint8_t*)c->data + c->offsets[r];
      out->slen = c->offsets[r + 1] - c->offsets[r];
      break;
    default:
      out->v = (uint64_t)int_at(c, r);
  }
  return 1;
}

/* ------------------------------------------------------------------------- */
/* hash_aggregate_batches (hash_aggregate.rs:182-339).                       */
/* ------------------------------------------------------------------------- */
typedef struct {
  double sum, min, max;
  uint64_t count;
  u128 exact; /* exact integer sum (two's complement) */
} acc_t;

static void acc_init(acc_t* a) {
  a->sum = 0.0;
  a->count = 0;
  a->min = INFINITY;
  a->max = -INFINITY;
  a->exact = 0;
}
/* GroupAccumulator::accumulate (135-140) */
static inline void acc_add(acc_t* a, double v) {
  a->sum += v;
  a->count += 1;
  if (v < a->min) a->min = v;
  if (v > a->max) a->max = v;
}
/* GroupAccumulator::merge (149-159) */
static inline void acc_merge(acc_t* a, const acc_t* o) {
  a->sum += o->sum;
  a->count += o->count;
  if (o->min < a->min) a->min = o->min;
  if (o->max > a->max) a->max = o->max;
  a->exact += o->exact;
}

typedef struct {
  int64_t bi, ri;
  uint64_t gh;
  int64_t acc; /* offset of this group's accumulators in its pool */
} group_t;

/* Growable arrays of groups and accumulators. */
typedef struct {
  group_t* g;
  acc_t* a;
  int64_t n, cap;
  int32_t n_aggs;
} gvec;

static void gvec_init(gvec* v, int32_t n_aggs) {
  v->n = 0;
  v->cap = 1024;
  v->n_aggs = n_aggs;
  v->g = (group_t*)malloc(sizeof(group_t) * (size_t)v->cap);
  v->a = (acc_t*)malloc(sizeof(acc_t) * (size_t)v->cap * (size_t)(n_aggs ? n_aggs : 1));
}
static int64_t gvec_push(gvec* v, int64_t bi, int64_t ri, uint64_t gh) {
  if (v->n == v->cap) {
    v->cap *= 2;
    v->g = (group_t*)realloc(v->g, sizeof(group_t) * (size_t)v->cap);
    v->a = (acc_t*)realloc(v->a, sizeof(acc_t) * (size_t)v->cap * (size_t)(v->n_aggs ? v->n_aggs : 1));
  }
  group_t* g = &v->g[v->n];
  g->bi = bi;
  g->ri = ri;
  g->gh = gh;
  g->acc = v->n * v->n_aggs;
  for (int32_t a = 0; a < v->n_aggs; ++a) acc_init(&v->a[g->acc + a]);
  return v->n++;
}
static void gvec_free(gvec* v) {
  free(v->g);
  free(v->a);
}

static int global_aggregate(const orc_batch* batches, int32_t n_batches, const orc_agg* aggs, int32_t n_aggs,
                            orc_agg_out* out);

int orc_hash_aggregate(const orc_batch* batches, int32_t n_batches, const int32_t* group_by, int32_t n_gb,
                       const orc_agg* aggs, int32_t n_aggs, int32_t n_threads, orc_agg_out* out) {
  int any_rows = 0;
  for (int32_t i = 0; i < n_batches; ++i)
    if (batches[i].n_rows > 0) any_rows = 1;
  for (int32_t a = 0; a < n_aggs; ++a) {
    if (aggs[a].func == ORC_COUNT_DISTINCT || (n_gb == 0 && aggs[a].func == ORC_COUNT)) continue;
    for (int32_t i = 0; i < n_batches; ++i) {
      if (batches[i].n_rows == 0) continue;
      if (aggs[a].column < 0 || aggs[a].column >= batches[i].n_columns) return -1;
      int t = batches[i].columns[aggs[a].column].type;
      if (t != ORC_INT32 && t != ORC_INT64 && t != ORC_FLOAT64 && t != ORC_BOOL) return -1;
    }
  }
  if (n_gb == 0 && any_rows) return global_aggregate(batches, n_batches, aggs, n_aggs, out);
  out->n_groups = 0;
  if (!any_rows) return 0; /* build_empty_result (188-203) */

  /* non-empty batches, chunked by current_num_threads (209-218) */
  int32_t* ne = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_batches);
  int32_t n_ne = 0;
  for (int32_t i = 0; i < n_batches; ++i)
    if (batches[i].n_rows > 0) ne[n_ne++] = i;
  int32_t threads = n_threads > 0 ? n_threads : 1;
  int32_t chunk = n_ne / threads;
  if (chunk < 1) chunk = 1;

  map64 gmap;
  map_init(&gmap, 1024);
  gvec G, L;
  gvec_init(&G, n_aggs);
  gvec_init(&L, n_aggs);

  for (int32_t c0 = 0; c0 < n_ne; c0 += chunk) {
    int32_t c1 = c0 + chunk < n_ne ? c0 + chunk : n_ne;
    /* per-chunk map and first-appearance order (221-272) */
    map64 lmap;
    map_init(&lmap, 1024);
    L.n = 0;
    for (int32_t ci = c0; ci < c1; ++ci) {
      const orc_batch* b = &batches[ne[ci]];
      for (int64_t r = 0; r < b->n_rows; ++r) {
        uint64_t gh = orc_group_hash(b, group_by, n_gb, r);
        int fresh;
        int64_t* slot = map_put(&lmap, gh, &fresh);
        if (fresh) *slot = gvec_push(&L, ne[ci], r, gh);
        acc_t* acc = &L.a[L.g[*slot].acc];
        for (int32_t a = 0; a < n_aggs; ++a) {
          if (aggs[a].func == ORC_COUNT_DISTINCT) continue;
          const orc_column* col = &b->columns[aggs[a].column];
          acc_add(&acc[a], f64_at(col, r));
          if (col->type != ORC_FLOAT64) acc[a].exact += (u128)(i128)int_at(col, r);
        }
      }
    }
    /* merge in chunk order (277-296) */
    for (int64_t i = 0; i < L.n; ++i) {
      int fresh;
      int64_t* slot = map_put(&gmap, L.g[i].gh, &fresh);
      if (fresh) *slot = gvec_push(&G, L.g[i].bi, L.g[i].ri, L.g[i].gh);
      for (int32_t a = 0; a < n_aggs; ++a) acc_merge(&G.a[G.g[*slot].acc + a], &L.a[L.g[i].acc + a]);
    }
    map_free(&lmap);
  }
  const int64_t n_g = G.n;
  /* COUNT DISTINCT: per-group sets (merge = union, 154-158) */
  int64_t* dcount = (int64_t*)calloc((size_t)n_g * (size_t)(n_aggs ? n_aggs : 1), sizeof(int64_t));
  for (int32_t a = 0; a < n_aggs; ++a) {
    if (aggs[a].func != ORC_COUNT_DISTINCT) continue;
    dset ds;
    dset_init(&ds);
    for (int32_t ci = 0; ci < n_ne; ++ci) {
      const orc_batch* b = &batches[ne[ci]];
      const orc_column* col = &b->columns[aggs[a].column];
      for (int64_t r = 0; r < b->n_rows; ++r) {
        uint64_t gh = orc_group_hash(b, group_by, n_gb, r);
        int64_t g = *map_get(&gmap, gh);
        dentry e;
        if (distinct_of(col, r, g, &e) && dset_insert(&ds, &e)) dcount[a * n_g + g]++;
      }
    }
    free(ds.e);
    free(ds.used);
  }
  /* output (298-338) */
  out->n_groups = n_g;
  for (int64_t g = 0; g < n_g && g < out->cap; ++g) {
    out->first_batch[g] = G.g[g].bi;
    out->first_row[g] = G.g[g].ri;
    out->key_hash[g] = G.g[g].gh;
    for (int32_t a = 0; a < n_aggs; ++a) {
      const acc_t* x = &G.a[G.g[g].acc + a];
      double v = 0.0;
      switch (aggs[a].func) { /* GroupAccumulator::result (161-170) */
        case ORC_SUM: v = x->sum; break;
        case ORC_COUNT: v = (double)x->count; break;
        case ORC_AVG: v = x->count > 0 ? x->sum / (double)x->count : 0.0; break;
        case ORC_MIN: v = x->count > 0 ? x->min : 0.0; break;
        case ORC_MAX: v = x->count > 0 ? x->max : 0.0; break;
        case ORC_COUNT_DISTINCT: v = (double)dcount[a * n_g + g]; break;
      }
      out->f64[a * out->cap + g] = v;
      out->i64[a * out->cap + g] = (int64_t)v;
      out->sum_lo[a * out->cap + g] = (uint64_t)x->exact;
      out->sum_hi[a * out->cap + g] = (int64_t)(x->exact >> 64);
    }
  }
  free(ne);
  free(dcount);
  gvec_free(&G);
  gvec_free(&L);
  map_free(&gmap);
  return 0;
}

/* global_aggregate_batches (hash_aggregate.rs:343-447).  Per-batch arrow sums
 * are restated as plain left-to-right sums over the non-null values. */
static int global_aggregate(const orc_batch* batches, int32_t n_batches, const orc_agg* aggs, int32_t n_aggs,
                            orc_agg_out* out) {
  out->n_groups = 1;
  out->first_batch[0] = 0;
  out->first_row[0] = 0;
  out->key_hash[0] = 0;
  for (int32_t a = 0; a < n_aggs; ++a) {
    double result = 0.0;
    u128 exact = 0;
    int32_t col = aggs[a].column;
    switch (aggs[a].func) {
      case ORC_SUM:
      case ORC_AVG: {
        double total = 0.0;
        uint64_t cnt = 0;
        for (int32_t i = 0; i < n_batches; ++i) {
          const orc_batch* b = &batches[i];
          if (b->n_rows == 0) continue;
          const orc_column* c = &b->columns[col];
          double s = 0.0;
          for (int64_t r = 0; r < b->n_rows; ++r) {
            if (!is_valid(c, r)) continue;
            s += f64_at(c, r);
            cnt++;
            if (c->type != ORC_FLOAT64) exact += (u128)(i128)int_at(c, r);
          }
          total += s;
        }
        result = aggs[a].func == ORC_SUM ? total : (cnt > 0 ? total / (double)cnt : 0.0);
        break;
      }
      case ORC_COUNT: {
        uint64_t cnt = 0;
        for (int32_t i = 0; i < n_batches; ++i) {
          const orc_batch* b = &batches[i];
          if (col < 0 || col >= b->n_columns) {
            cnt += (uint64_t)b->n_rows; /* COUNT(*) on a zero-column batch (367-369) */
            continue;
          }
          const orc_column* c = &b->columns[col];
          for (int64_t r = 0; r < b->n_rows; ++r) cnt += is_valid(c, r) ? 1 : 0;
        }
        result = (double)cnt;
        break;
      }
      case ORC_MIN:
      case ORC_MAX: {
        int is_min = aggs[a].func == ORC_MIN;
        double acc = is_min ? INFINITY : -INFINITY;
        for (int32_t i = 0; i < n_batches; ++i) {
          const orc_batch* b = &batches[i];
          if (b->n_rows == 0) continue;
          const orc_column* c = &b->columns[col];
          /* arrow min/max: NaN orders above every number; None when all null */
          int have = 0, nan = 0;
          double m = 0.0;
          for (int64_t r = 0; r < b->n_rows; ++r) {
            if (!is_valid(c, r)) continue;
            double v = f64_at(c, r);
            if (isnan(v)) {
              nan = 1;
              continue;
            }
            if (!have || (is_min ? v < m : v > m)) m = v;
            have = 1;
          }
          double part;
          if (is_min) {
            if (!have && !nan) continue;
            part = have ? m : NAN;
            acc = fmin(acc, part); /* f64::min ignores NaN */
          } else {
            if (!have && !nan) continue;
            part = nan ? NAN : m;
            acc = fmax(acc, part);
          }
        }
        if (is_min) result = acc == INFINITY ? 0.0 : acc;
        else result = acc == -INFINITY ? 0.0 : acc;
        break;
      }
      case ORC_COUNT_DISTINCT: {
        dset ds;
        dset_init(&ds);
        int64_t n = 0;
        for (int32_t i = 0; i < n_batches; ++i) {
          const orc_batch* b = &batches[i];
          const orc_column* c = &b->columns[col];
          for (int64_t r = 0; r < b->n_rows; ++r) {
            dentry e;
            if (distinct_of(c, r, 0, &e) && dset_insert(&ds, &e)) n++;
          }
        }
        free(ds.e);
        free(ds.used);
        result = (double)n;
        break;
      }
    }
    out->f64[a * out->cap] = result;
    out->i64[a * out->cap] = (int64_t)result;
    out->sum_lo[a * out->cap] = (uint64_t)exact;
    out->sum_hi[a * out->cap] = (int64_t)(exact >> 64);
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* HashTable::build / probe (hash_join.rs:87-200).                            */
/* ------------------------------------------------------------------------- */
static uint64_t row_key_hash(const orc_batch* b, const int32_t* keys, int32_t n_keys, int64_t r) {
  /* Placement hash only (the reference uses SipHash here, hash_join.rs:386);
   * equality is decided by keys_equal below. */
  uint64_t h = 0x243F6A8885A308D3ULL;
  for (int32_t k = 0; k < n_keys; ++k) {
    const orc_column* c = &b->columns[keys[k]];
    if (!is_valid(c, r)) {
      h = mix64(h ^ 0x1234567ULL);
      continue;
    }
    switch (c->type) {
      case ORC_UTF8: {
        const uint8_t* s = (const uint8_t*)c->data + c->offsets[r];
        int32_t n = c->offsets[r + 1] - c->offsets[r];
        for (int32_t i = 0; i < n; ++i) h = mix64(h ^ s[i]);
        h = mix64(h ^ (uint64_t)n ^ 0xffULL << 40);
        break;
      }
      case ORC_FLOAT64: {
        uint64_t bits;
        memcpy(&bits, &((const double*)c->data)[r], 8);
        h = mix64(h ^ bits);
        break;
      }
      default:
        h = mix64(h ^ (uint64_t)int_at(c, r));
    }
  }
  return h;
}

/* keys_equal (hash_join.rs:628-670) plus the null handling that the hash
 * imposes: a null hashes as 0u8, so it only meets another null, and then the
 * raw values are compared. */
static int keys_equal(const orc_batch* a, const int32_t* ka, int64_t ra, const orc_batch* b, const int32_t* kb,
                      int64_t rb, int32_t n_keys) {
  for (int32_t k = 0; k < n_keys; ++k) {
    const orc_column* x = &a->columns[ka[k]];
    const orc_column* y = &b->columns[kb[k]];
    int vx = is_valid(x, ra), vy = is_valid(y, rb);
    if (vx != vy) return 0;
    switch (x->type) {
      case ORC_FLOAT64: {
        double p = ((const double*)x->data)[ra], q = ((const double*)y->data)[rb];
        uint64_t bp, bq;
        memcpy(&bp, &p, 8);
        memcpy(&bq, &q, 8);
        if (bp != bq || !(p == q)) return 0; /* hash on bits, equality by == */
        break;
      }
      case ORC_UTF8: {
        int32_t nx = x->offsets[ra + 1] - x->offsets[ra], ny = y->offsets[rb + 1] - y->offsets[rb];
        if (nx != ny) return 0;
        if (memcmp((const uint8_t*)x->data + x->offsets[ra], (const uint8_t*)y->data + y->offsets[rb], (size_t)nx))
          return 0;
        break;
      }
      default:
        if (int_at(x, ra) != int_at(y, rb)) return 0;
    }
  }
  return 1;
}

int64_t orc_join_probe(const orc_batch* build, const int32_t* build_keys, const orc_batch* probe,
                       const int32_t* probe_keys, int32_t n_keys, uint32_t* out_probe, uint32_t* out_build,
                       int64_t cap) {
  const int fast = n_keys == 1 && build->columns[build_keys[0]].type == ORC_INT64 &&
                   probe->columns[probe_keys[0]].type == ORC_INT64;
  const int64_t nb = build->n_rows;
  /* key hash -> row list, rows ascending (the serial scatter keeps order, 116-118) */
  uint64_t* bh = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nb ? nb : 1));
  for (int64_t r = 0; r < nb; ++r)
    bh[r] = fast ? (uint64_t)((const int64_t*)build->columns[build_keys[0]].data)[r] * FX_K
                 : row_key_hash(build, build_keys, n_keys, r);
  map64 m;
  map_init(&m, nb + 16);
  int64_t n_keys_distinct = 0;
  for (int64_t r = 0; r < nb; ++r) {
    int fresh;
    int64_t* v = map_put(&m, bh[r], &fresh);
    if (fresh) *v = n_keys_distinct++;
  }
  int64_t* cnt = (int64_t*)calloc((size_t)(n_keys_distinct + 1), sizeof(int64_t));
  int64_t* gid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb ? nb : 1));
  for (int64_t r = 0; r < nb; ++r) {
    gid[r] = *map_get(&m, bh[r]);
    cnt[gid[r] + 1]++;
  }
  for (int64_t i = 0; i < n_keys_distinct; ++i) cnt[i + 1] += cnt[i];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_keys_distinct + 1));
  memcpy(fill, cnt, sizeof(int64_t) * (size_t)(n_keys_distinct + 1));
  uint32_t* rows = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nb ? nb : 1));
  for (int64_t r = 0; r < nb; ++r) rows[fill[gid[r]]++] = (uint32_t)r;

  int64_t n_out = 0;
  for (int64_t p = 0; p < probe->n_rows; ++p) {
    uint64_t h = fast ? (uint64_t)((const int64_t*)probe->columns[probe_keys[0]].data)[p] * FX_K
                      : row_key_hash(probe, probe_keys, n_keys, p);
    int64_t* g = map_get(&m, h);
    if (!g) continue;
    for (int64_t i = cnt[*g]; i < cnt[*g + 1]; ++i) {
      uint32_t b = rows[i];
      if (!fast && !keys_equal(probe, probe_keys, p, build, build_keys, b, n_keys)) continue;
      if (n_out < cap) {
        out_probe[n_out] = (uint32_t)p;
        out_build[n_out] = b;
      }
      n_out++;
    }
  }
  free(bh);
  free(cnt);
  free(gid);
  free(fill);
  free(rows);
  map_free(&m);
  return n_out;
}

/* ------------------------------------------------------------------------- */
/* LIKE (string_ops.rs:16-25 -> arrow_string::like).                         */
/* ------------------------------------------------------------------------- */
static int utf8_len(uint8_t c) {
  return c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 1;
}
static uint8_t fold(uint8_t c) { return (c >= 'A' && c <= 'Z') ? (uint8_t)(c + 32) : c; }

/* Recursive definition of SQL LIKE over UTF-8 characters (exponential in the
 * worst case, fine for test sizes). */
static int like_rec(const uint8_t* s, int32_t ns, const uint8_t* p, int32_t np, int ci) {
  if (np == 0) return ns == 0;
  if (p[0] == '%') {
    for (int32_t i = 0; i <= ns;) {
      if (like_rec(s + i, ns - i, p + 1, np - 1, ci)) return 1;
      if (i == ns) break;
      i += utf8_len(s[i]);
    }
    return 0;
  }
  if (ns == 0) return 0;
  if (p[0] == '_') {
    int32_t k = utf8_len(s[0]);
    if (k > ns) k = ns;
    return like_rec(s + k, ns - k, p + 1, np - 1, ci);
  }
  int32_t pk = 0;
  uint8_t pc = p[0];
  if (pc == '\\' && np > 1) {
    pk = 1;
    pc = p[1];
  }
  if (ci ? fold(s[0]) != fold(pc) : s[0] != pc) return 0;
  return like_rec(s + 1, ns - 1, p + pk + 1, np - pk - 1, ci);
}

int orc_like(const uint8_t* s, int32_t ns, const uint8_t* p, int32_t np, int32_t ci) {
  return like_rec(s, ns, p, np, ci);
}

void orc_like_column(const orc_column* c, const uint8_t* p, int32_t np, int32_t ci, uint8_t* out) {
  for (int64_t r = 0; r < c->length; ++r) {
    const uint8_t* s = (const uint8_t*)c->data + c->offsets[r];
    out[r] = (uint8_t)like_rec(s, c->offsets[r + 1] - c->offsets[r], p, np, ci);
  }
}

/* ------------------------------------------------------------------------- */
/* Eager join -> aggregate.                                                   */
/* ------------------------------------------------------------------------- */
int orc_join_agg_eager(const int64_t* build_key, const int32_t* build_attr, int64_t n_build,
                       const int64_t* probe_key, const int64_t* probe_payload, int64_t n_probe,
                       int64_t n_groups, uint64_t* count_out, uint64_t* sum_lo_out, int64_t* sum_hi_out) {
  map64 m;
  map_init(&m, n_build + 16);
  int64_t nk = 0;
  for (int64_t i = 0; i < n_build; ++i) {
    int fresh;
    int64_t* v = map_put(&m, (uint64_t)build_key[i], &fresh);
    if (fresh) *v = nk++;
  }
  uint64_t* pc = (uint64_t*)calloc((size_t)(nk + 1), 8);
  u128* ps = (u128*)calloc((size_t)(nk + 1), sizeof(u128));
  for (int64_t i = 0; i < n_probe; ++i) {
    int64_t* k = map_get(&m, (uint64_t)probe_key[i]);
    if (!k) continue;
    pc[*k]++;
    ps[*k] += (u128)(i128)probe_payload[i];
  }
  u128* gs = (u128*)calloc((size_t)n_groups, sizeof(u128));
  for (int64_t g = 0; g < n_groups; ++g) count_out[g] = 0;
  for (int64_t b = 0; b < n_build; ++b) {
    int64_t k = *map_get(&m, (uint64_t)build_key[b]);
    count_out[build_attr[b]] += pc[k];
    gs[build_attr[b]] += ps[k];
  }
  for (int64_t g = 0; g < n_groups; ++g) {
    sum_lo_out[g] = (uint64_t)gs[g];
    sum_hi_out[g] = (int64_t)(gs[g] >> 64);
  }
  free(pc);
  free(ps);
  free(gs);
  map_free(&m);
  return 0;
}
