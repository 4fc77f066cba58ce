/*
 * The reference's CPU operator path with its own data structures and
 * parallel decomposition (see ref_cpu.h).  TEST INFRASTRUCTURE ONLY: timed
 * as bench.py's CPU baseline and pinned against the oracle by tests/.
 */
#include "ref_cpu.h"

#include <emmintrin.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define FX_K 0x517cc1b727220a95ULL
#define N_PART 64 /* HASH_PARTITIONS (hash_join.rs:44-45) */

/* ------------------------------------------------------------------------- */
/* Worker pool: tasks taken from a shared counter.                            */
/* ------------------------------------------------------------------------- */
typedef struct {
  void (*fn)(void*, int64_t);
  void* arg;
  int64_t n_tasks;
  atomic_llong next;
} job_t;

static void* work(void* p) {
  job_t* j = (job_t*)p;
  for (;;) {
    const int64_t t = atomic_fetch_add(&j->next, 1);
    if (t >= j->n_tasks) break;
    j->fn(j->arg, t);
  }
  return NULL;
}

static void run_tasks(int32_t threads, int64_t n_tasks, void (*fn)(void*, int64_t), void* arg) {
  job_t j;
  j.fn = fn;
  j.arg = arg;
  j.n_tasks = n_tasks;
  atomic_init(&j.next, 0);
  int64_t nt = threads < n_tasks ? threads : n_tasks;
  if (nt > 256) nt = 256;
  pthread_t th[256];
  for (int64_t i = 1; i < nt; ++i) pthread_create(&th[i], NULL, work, &j);
  work(&j);
  for (int64_t i = 1; i < nt; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------------- */
/* Growable arrays (Vec<T>): capacity 4 at the first push, doubling.           */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t* p;
  uint32_t len, cap;
} vec_u32;

static inline void vpush(vec_u32* v, uint32_t x) {
  if (v->len == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 4;
    v->p = (uint32_t*)realloc(v->p, (size_t)v->cap * 4);
  }
  v->p[v->len++] = x;
}

typedef struct {
  uint32_t* p;
  int64_t len, cap;
} vec_big;

static inline void bpush(vec_big* v, uint32_t x) {
  if (v->len == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 4;
    v->p = (uint32_t*)realloc(v->p, (size_t)v->cap * 4);
  }
  v->p[v->len++] = x;
}

/* ------------------------------------------------------------------------- */
/* SwissTable map u64 -> fixed-size value (the hashbrown layout): 16-byte      */
/* control groups, 7-bit tags from the hash's top bits, bucket = hash & mask, */
/* triangular probing, load limit 7/8, no deletions.                         */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint8_t* ctrl; /* buckets + 16 */
  uint64_t* keys;
  uint8_t* vals;
  int64_t buckets, items, growth_left;
  int32_t vsize;
  int32_t sip; /* map hash: 0 = the key itself (PassThroughU64), 1 = SipHash-1-3 (std RandomState) */
} swiss;

/* SipHash-1-3 of a u64 written as 8 bytes (Aumasson & Bernstein; std's
 * default hasher).  Fixed keys stand in for RandomState's random ones. */
#define ROTL(x, b) (uint64_t)(((x) << (b)) | ((x) >> (64 - (b))))
#define SIPROUND                     \
  do {                               \
    v0 += v1;                        \
    v1 = ROTL(v1, 13);               \
    v1 ^= v0;                        \
    v0 = ROTL(v0, 32);               \
    v2 += v3;                        \
    v3 = ROTL(v3, 16);               \
    v3 ^= v2;                        \
    v0 += v3;                        \
    v3 = ROTL(v3, 21);               \
    v3 ^= v0;                        \
    v2 += v1;                        \
    v1 = ROTL(v1, 17);               \
    v1 ^= v2;                        \
    v2 = ROTL(v2, 32);               \
  } while (0)

static inline uint64_t sip13_u64(uint64_t m) {
  const uint64_t k0 = 0x0706050403020100ULL, k1 = 0x0f0e0d0c0b0a0908ULL;
  uint64_t v0 = k0 ^ 0x736f6d6570736575ULL, v1 = k1 ^ 0x646f72616e646f6dULL;
  uint64_t v2 = k0 ^ 0x6c7967656e657261ULL, v3 = k1 ^ 0x7465646279746573ULL;
  v3 ^= m;
  SIPROUND;
  v0 ^= m;
  const uint64_t b = 8ULL << 56;
  v3 ^= b;
  SIPROUND;
  v0 ^= b;
  v2 ^= 0xff;
  SIPROUND;
  SIPROUND;
  SIPROUND;
  return v0 ^ v1 ^ v2 ^ v3;
}

static inline uint64_t map_hash(const swiss* m, uint64_t key) { return m->sip ? sip13_u64(key) : key; }

static int64_t next_pow2(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

static int64_t buckets_for(int64_t cap) {
  if (cap < 8) return cap < 4 ? 4 : 8;
  return next_pow2(cap * 8 / 7);
}

static int64_t full_capacity(int64_t buckets) { return buckets < 8 ? buckets - 1 : buckets / 8 * 7; }

static void swiss_alloc(swiss* m, int64_t buckets) {
  m->buckets = buckets;
  m->items = 0;
  m->growth_left = full_capacity(buckets);
  m->ctrl = (uint8_t*)malloc((size_t)buckets + 16);
  memset(m->ctrl, 0xFF, (size_t)buckets + 16);
  m->keys = (uint64_t*)malloc((size_t)buckets * 8);
  m->vals = (uint8_t*)malloc((size_t)buckets * (size_t)m->vsize);
}

/* with_capacity(cap); cap = 0 leaves the map unallocated (HashMap::new()). */
static void swiss_init(swiss* m, int64_t cap, int32_t vsize, int32_t sip) {
  memset(m, 0, sizeof(*m));
  m->vsize = vsize;
  m->sip = sip;
  if (cap > 0) swiss_alloc(m, buckets_for(cap));
}

static void swiss_free(swiss* m) {
  free(m->ctrl);
  free(m->keys);
  free(m->vals);
  memset(m, 0, sizeof(*m));
}

static inline void set_ctrl(swiss* m, int64_t i, uint8_t v) {
  m->ctrl[i] = v;
  m->ctrl[((i - 16) & (m->buckets - 1)) + 16] = v;
}

static inline int64_t swiss_find(const swiss* m, uint64_t hash, uint64_t key) {
  if (!m->buckets) return -1;
  const uint64_t mask = (uint64_t)m->buckets - 1;
  const __m128i tag = _mm_set1_epi8((char)(hash >> 57));
  const __m128i empty = _mm_set1_epi8((char)0xFF);
  uint64_t pos = hash & mask, stride = 0;
  for (;;) {
    const __m128i g = _mm_loadu_si128((const __m128i*)(m->ctrl + pos));
    unsigned bits = (unsigned)_mm_movemask_epi8(_mm_cmpeq_epi8(g, tag));
    while (bits) {
      const uint64_t idx = (pos + (uint64_t)__builtin_ctz(bits)) & mask;
      if (m->keys[idx] == key) return (int64_t)idx;
      bits &= bits - 1;
    }
    if (_mm_movemask_epi8(_mm_cmpeq_epi8(g, empty))) return -1;
    stride += 16;
    pos = (pos + stride) & mask;
  }
}

static int64_t insert_slot(const swiss* m, uint64_t hash) {
  const uint64_t mask = (uint64_t)m->buckets - 1;
  uint64_t pos = hash & mask, stride = 0;
  for (;;) {
    const __m128i g = _mm_loadu_si128((const __m128i*)(m->ctrl + pos));
    const unsigned e = (unsigned)_mm_movemask_epi8(g); /* EMPTY has its top bit set */
    if (e) {
      int64_t idx = (int64_t)((pos + (uint64_t)__builtin_ctz(e)) & mask);
      if (!(m->ctrl[idx] & 0x80)) { /* small table: the group ran past the end */
        idx = __builtin_ctz((unsigned)_mm_movemask_epi8(_mm_loadu_si128((const __m128i*)m->ctrl)));
      }
      return idx;
    }
    stride += 16;
    pos = (pos + stride) & mask;
  }
}

static void swiss_grow(swiss* m) {
  swiss o = *m;
  const int64_t want = (o.buckets ? full_capacity(o.buckets) : 0) + 1;
  swiss_alloc(m, buckets_for(want > o.items + 1 ? want : o.items + 1));
  for (int64_t i = 0; i < o.buckets; ++i)
    if (!(o.ctrl[i] & 0x80)) {
      const uint64_t h = map_hash(m, o.keys[i]);
      const int64_t s = insert_slot(m, h);
      set_ctrl(m, s, (uint8_t)(h >> 57));
      m->keys[s] = o.keys[i];
      memcpy(m->vals + (size_t)s * (size_t)m->vsize, o.vals + (size_t)i * (size_t)o.vsize, (size_t)m->vsize);
      m->items++;
      m->growth_left--;
    }
  free(o.ctrl);
  free(o.keys);
  free(o.vals);
}

/* entry(key).or_insert(zeroed value): index of the key's slot. */
static int64_t swiss_entry(swiss* m, uint64_t key, int* fresh) {
  const uint64_t h = map_hash(m, key);
  const int64_t f = swiss_find(m, h, key);
  if (f >= 0) {
    *fresh = 0;
    return f;
  }
  if (m->growth_left == 0) swiss_grow(m);
  const int64_t s = insert_slot(m, h);
  set_ctrl(m, s, (uint8_t)(h >> 57));
  m->keys[s] = key;
  memset(m->vals + (size_t)s * (size_t)m->vsize, 0, (size_t)m->vsize);
  m->items++;
  m->growth_left--;
  *fresh = 1;
  return s;
}

static inline void* swiss_val(const swiss* m, int64_t i) { return m->vals + (size_t)i * (size_t)m->vsize; }

/* ------------------------------------------------------------------------- */
/* HashTable::build / probe_matches_chunked (hash_join.rs:87-200, 286-309).   */
/* ------------------------------------------------------------------------- */
struct rc_table {
  swiss part[N_PART];
  int64_t n_rows;
};

typedef struct {
  uint64_t h;
  uint32_t row;
} hrow;

typedef struct {
  const int64_t* keys;
  uint64_t* hashes;
  int64_t n;
} hash_job;

static void hash_task(void* a, int64_t t) {
  hash_job* j = (hash_job*)a;
  const int64_t b = t * 65536, e = b + 65536 < j->n ? b + 65536 : j->n;
  for (int64_t i = b; i < e; ++i) j->hashes[i] = (uint64_t)j->keys[i] * FX_K; /* hash_i64_fast */
}

typedef struct {
  rc_table* t;
  hrow* rows[N_PART];
  int64_t len[N_PART];
} part_job;

static void part_task(void* a, int64_t p) {
  part_job* j = (part_job*)a;
  swiss* m = &j->t->part[p];
  swiss_init(m, j->len[p], (int32_t)sizeof(vec_u32), 0);
  for (int64_t i = 0; i < j->len[p]; ++i) {
    int fresh;
    const int64_t s = swiss_entry(m, j->rows[p][i].h, &fresh);
    vpush((vec_u32*)swiss_val(m, s), j->rows[p][i].row);
  }
}

rc_table* rc_join_build(const int64_t* keys, int64_t n, int32_t threads) {
  rc_table* t = (rc_table*)calloc(1, sizeof(rc_table));
  t->n_rows = n;
  hash_job hj = {keys, (uint64_t*)malloc((size_t)(n ? n : 1) * 8), n};
  run_tasks(threads, (n + 65535) / 65536, hash_task, &hj);
  /* serial scatter into the partitions, each reserved n / 64 + 1 */
  part_job pj;
  pj.t = t;
  int64_t cap[N_PART];
  for (int p = 0; p < N_PART; ++p) {
    cap[p] = n / N_PART + 1;
    pj.rows[p] = (hrow*)malloc((size_t)cap[p] * sizeof(hrow));
    pj.len[p] = 0;
  }
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t h = hj.hashes[i];
    const int p = (int)(h & (N_PART - 1));
    if (pj.len[p] == cap[p]) {
      cap[p] *= 2;
      pj.rows[p] = (hrow*)realloc(pj.rows[p], (size_t)cap[p] * sizeof(hrow));
    }
    pj.rows[p][pj.len[p]].h = h;
    pj.rows[p][pj.len[p]].row = (uint32_t)i;
    pj.len[p]++;
  }
  free(hj.hashes);
  run_tasks(threads, N_PART, part_task, &pj);
  for (int p = 0; p < N_PART; ++p) free(pj.rows[p]);
  return t;
}

void rc_join_free(rc_table* t) {
  if (!t) return;
  for (int p = 0; p < N_PART; ++p) {
    swiss* m = &t->part[p];
    for (int64_t i = 0; i < m->buckets; ++i)
      if (!(m->ctrl[i] & 0x80)) free(((vec_u32*)swiss_val(m, i))->p);
    swiss_free(m);
  }
  free(t);
}

/* HashTable::probe, Int64 fast path (hash_join.rs:152-173): pairs of one batch. */
static void probe_batch(const rc_table* t, const int64_t* keys, int64_t n, vec_big* pi, vec_big* bi) {
  pi->cap = bi->cap = n;
  pi->len = bi->len = 0;
  pi->p = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
  bi->p = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
  for (int64_t r = 0; r < n; ++r) {
    const uint64_t h = (uint64_t)keys[r] * FX_K;
    const swiss* m = &t->part[h & (N_PART - 1)];
    const int64_t s = swiss_find(m, h, h);
    if (s < 0) continue;
    const vec_u32* v = (const vec_u32*)swiss_val(m, s);
    for (uint32_t k = 0; k < v->len; ++k) {
      bpush(pi, (uint32_t)r);
      bpush(bi, v->p[k]);
    }
  }
}

typedef struct {
  const rc_table* t;
  const int64_t* keys;
  int64_t n, batch_rows;
  int64_t* pairs;
  uint64_t *pck, *bck;
} probe_job;

static void probe_task(void* a, int64_t b) {
  probe_job* j = (probe_job*)a;
  const int64_t lo = b * j->batch_rows;
  const int64_t hi = lo + j->batch_rows < j->n ? lo + j->batch_rows : j->n;
  vec_big pi, bi;
  probe_batch(j->t, j->keys + lo, hi - lo, &pi, &bi);
  uint64_t ps = 0, bs = 0;
  for (int64_t i = 0; i < pi.len; ++i) {
    ps += pi.p[i];
    bs += bi.p[i];
  }
  j->pairs[b] = pi.len;
  j->pck[b] = ps;
  j->bck[b] = bs;
  free(pi.p);
  free(bi.p);
}

int64_t rc_probe_matches_chunked(const rc_table* t, const int64_t* keys, int64_t n, int64_t batch_rows,
                                 int32_t threads, uint64_t* probe_checksum, uint64_t* build_checksum) {
  const int64_t nb = (n + batch_rows - 1) / batch_rows;
  probe_job j = {t, keys, n, batch_rows, (int64_t*)calloc((size_t)(nb ? nb : 1), 8),
                 (uint64_t*)calloc((size_t)(nb ? nb : 1), 8), (uint64_t*)calloc((size_t)(nb ? nb : 1), 8)};
  run_tasks(threads, nb, probe_task, &j);
  int64_t total = 0;
  uint64_t ps = 0, bs = 0;
  for (int64_t b = 0; b < nb; ++b) {
    total += j.pairs[b];
    ps += j.pck[b];
    bs += j.bck[b];
  }
  free(j.pairs);
  free(j.pck);
  free(j.bck);
  if (probe_checksum) *probe_checksum = ps;
  if (build_checksum) *build_checksum = bs;
  return total;
}

/* ------------------------------------------------------------------------- */
/* hash_aggregate_batches, grouped path (hash_aggregate.rs:182-339).          */
/* ------------------------------------------------------------------------- */
typedef struct {
  double sum;
  uint64_t count;
  double min, max;
  void* distinct;
} acc_t; /* GroupAccumulator (hash_aggregate.rs:104-171) */

typedef struct {
  int64_t batch, row;
  acc_t* accs; /* Vec<GroupAccumulator> */
} gval;

enum { KEY_I64 = 0, KEY_UTF8x2 = 1, KEY_I32 = 2 };

typedef struct {
  int64_t n;
  int32_t key_kind;
  const int64_t* k64;
  const int32_t* k32;
  const uint8_t *s0, *s1;
  const int32_t *o0, *o1;
  const double* vals[4]; /* f64 value array of each aggregate */
  int64_t row_base;
} abatch;

static inline uint64_t fx(uint64_t st, uint64_t w) { return st * FX_K + w; }

static inline uint64_t fx_str(uint64_t st, const uint8_t* s, int32_t n) {
  for (int32_t p = 0; p < n; p += 8) {
    uint64_t w = 0;
    const int32_t m = n - p < 8 ? n - p : 8;
    for (int32_t k = 0; k < m; ++k) w |= (uint64_t)s[p + k] << (8 * k);
    st = fx(st, w);
  }
  return fx(st, 0xff);
}

/* hash_group_keys (hash_aggregate.rs:483-490) for the key shapes used here. */
static inline uint64_t group_hash(const abatch* b, int64_t r) {
  switch (b->key_kind) {
    case KEY_I64: return fx(0, (uint64_t)b->k64[r]);
    case KEY_I32: return fx(0, (uint64_t)(int64_t)b->k32[r]);
    default: {
      const uint64_t st = fx_str(0, b->s0 + b->o0[r], b->o0[r + 1] - b->o0[r]);
      return fx_str(st, b->s1 + b->o1[r], b->o1[r + 1] - b->o1[r]);
    }
  }
}

static inline void accumulate(acc_t* a, double v) {
  a->sum += v;
  a->count += 1;
  if (v < a->min) a->min = v;
  if (v > a->max) a->max = v;
}

typedef struct {
  swiss map;
  uint64_t* order;
  int64_t n_order, cap_order;
} gmap;

static acc_t* new_accs(int32_t n_aggs) {
  acc_t* a = (acc_t*)malloc(sizeof(acc_t) * (size_t)(n_aggs ? n_aggs : 1));
  for (int32_t i = 0; i < n_aggs; ++i) {
    a[i].sum = 0.0;
    a[i].count = 0;
    a[i].min = __builtin_inf();
    a[i].max = -__builtin_inf();
    a[i].distinct = NULL;
  }
  return a;
}

static gval* gmap_entry(gmap* g, uint64_t gh, int64_t batch, int64_t row, int32_t n_aggs) {
  int fresh;
  const int64_t s = swiss_entry(&g->map, gh, &fresh);
  gval* v = (gval*)swiss_val(&g->map, s);
  if (fresh) {
    if (g->n_order == g->cap_order) {
      g->cap_order = g->cap_order ? g->cap_order * 2 : 4;
      g->order = (uint64_t*)realloc(g->order, (size_t)g->cap_order * 8);
    }
    g->order[g->n_order++] = gh;
    v->batch = batch;
    v->row = row;
    v->accs = new_accs(n_aggs);
  }
  return v;
}

static void gmap_free(gmap* g) {
  for (int64_t i = 0; i < g->map.buckets; ++i)
    if (!(g->map.ctrl[i] & 0x80)) free(((gval*)swiss_val(&g->map, i))->accs);
  swiss_free(&g->map);
  free(g->order);
}

typedef struct {
  const abatch* b;
  int64_t nb, chunk;
  int32_t n_aggs;
  gmap* parts;
} agg_job;

static void agg_task(void* a, int64_t c) {
  agg_job* j = (agg_job*)a;
  gmap* g = &j->parts[c];
  swiss_init(&g->map, 0, (int32_t)sizeof(gval), 1);
  const int64_t lo = c * j->chunk, hi = lo + j->chunk < j->nb ? lo + j->chunk : j->nb;
  for (int64_t bi = lo; bi < hi; ++bi) {
    const abatch* b = &j->b[bi];
    for (int64_t r = 0; r < b->n; ++r) {
      gval* v = gmap_entry(g, group_hash(b, r), bi, r, j->n_aggs);
      for (int32_t k = 0; k < j->n_aggs; ++k) accumulate(&v->accs[k], b->vals[k][r]);
    }
  }
}

static double result_of(const acc_t* a, int32_t func) {
  switch (func) {
    case 0: return a->sum;
    case 1: return (double)a->count;
    case 2: return a->count ? a->sum / (double)a->count : 0.0;
    case 3: return a->count ? a->min : 0.0;
    default: return a->count ? a->max : 0.0;
  }
}

/* funcs: 0 SUM, 1 COUNT, 2 AVG, 3 MIN, 4 MAX.  Batches given are non-empty. */
static void aggregate(const abatch* b, int64_t nb, int32_t n_aggs, const int32_t* funcs, int32_t threads,
                      rc_groups* out, double* dense_count, double* dense_sum, int32_t n_attr) {
  int64_t chunk = nb / (threads > 0 ? threads : 1);
  if (chunk < 1) chunk = 1;
  const int64_t n_chunks = (nb + chunk - 1) / chunk;
  agg_job j = {b, nb, chunk, n_aggs, (gmap*)calloc((size_t)(n_chunks ? n_chunks : 1), sizeof(gmap))};
  run_tasks(threads, n_chunks, agg_task, &j);
  gmap fin;
  memset(&fin, 0, sizeof(fin));
  swiss_init(&fin.map, 0, (int32_t)sizeof(gval), 1);
  for (int64_t c = 0; c < n_chunks; ++c) { /* merge in chunk order (hash_aggregate.rs:276-296) */
    gmap* p = &j.parts[c];
    for (int64_t i = 0; i < p->n_order; ++i) {
      const uint64_t gh = p->order[i];
      const int64_t s = swiss_find(&p->map, map_hash(&p->map, gh), gh);
      const gval* pv = (const gval*)swiss_val(&p->map, s);
      gval* v = gmap_entry(&fin, gh, pv->batch, pv->row, n_aggs);
      for (int32_t k = 0; k < n_aggs; ++k) { /* GroupAccumulator::merge */
        acc_t* x = &v->accs[k];
        const acc_t* y = &pv->accs[k];
        x->sum += y->sum;
        x->count += y->count;
        if (y->min < x->min) x->min = y->min;
        if (y->max > x->max) x->max = y->max;
      }
    }
    gmap_free(p);
  }
  free(j.parts);
  if (out) {
    out->n_groups = fin.n_order;
    for (int64_t i = 0; i < fin.n_order && i < 1024; ++i) {
      const uint64_t gh = fin.order[i];
      const gval* v = (const gval*)swiss_val(&fin.map, swiss_find(&fin.map, map_hash(&fin.map, gh), gh));
      out->key_hash[i] = gh;
      out->first_row[i] = b[v->batch].row_base + v->row;
      for (int32_t k = 0; k < n_aggs && k < 4; ++k) out->agg[k][i] = result_of(&v->accs[k], funcs[k]);
    }
  }
  if (dense_count) { /* join-aggregate: the key is the attribute itself */
    for (int64_t i = 0; i < fin.n_order; ++i) {
      const uint64_t gh = fin.order[i];
      const gval* v = (const gval*)swiss_val(&fin.map, swiss_find(&fin.map, map_hash(&fin.map, gh), gh));
      const int64_t attr = (int64_t)gh;
      if (attr >= 0 && attr < n_attr) {
        dense_count[attr] = result_of(&v->accs[0], 0);
        dense_sum[attr] = result_of(&v->accs[1], 0);
      }
    }
  }
  gmap_free(&fin);
}

static double* fill_f64(double v, int64_t n) {
  double* d = (double*)malloc((size_t)(n ? n : 1) * 8);
  for (int64_t i = 0; i < n; ++i) d[i] = v;
  return d;
}

/* arrow cast Int64 -> Float64: a fresh array */
static double* cast_f64(const int64_t* v, int64_t n) {
  double* d = (double*)malloc((size_t)(n ? n : 1) * 8);
  for (int64_t i = 0; i < n; ++i) d[i] = (double)v[i];
  return d;
}

int rc_groupby_i64(const int64_t* key, const int64_t* value, int64_t n, int64_t batch_rows, int32_t threads,
                   rc_groups* out) {
  const int64_t nb = (n + batch_rows - 1) / batch_rows;
  abatch* b = (abatch*)calloc((size_t)(nb ? nb : 1), sizeof(abatch));
  const int32_t funcs[4] = {0, 1, 3, 4};
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t lo = i * batch_rows, m = lo + batch_rows < n ? batch_rows : n - lo;
    b[i].n = m;
    b[i].key_kind = KEY_I64;
    b[i].k64 = key + lo;
    b[i].row_base = lo;
    /* each aggregate casts its column on its own (hash_aggregate.rs:227-245) */
    for (int k = 0; k < 4; ++k) b[i].vals[k] = cast_f64(value + lo, m);
  }
  aggregate(b, nb, 4, funcs, threads, out, NULL, NULL, 0);
  for (int64_t i = 0; i < nb; ++i)
    for (int k = 0; k < 4; ++k) free((void*)b[i].vals[k]);
  free(b);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Filter -> Project -> Aggregate (executor.rs:91-187, filter_project.rs).    */
/* ------------------------------------------------------------------------- */
typedef struct {
  const int32_t* date;
  const int64_t *price, *disc;
  const uint8_t *fa, *fb;
  const int32_t *fao, *fbo;
  int64_t n, batch_rows;
  int32_t date_max, variant;
  /* per batch outputs */
  int64_t* m;
  int32_t** date_f;
  int64_t **price_f, **disc_f;
  uint8_t **fa_f, **fb_f;
  int32_t **fao_f, **fbo_f;
  double** rev;
} fpa_job;

static void filter_utf8(const uint8_t* s, const int32_t* o, const uint8_t* bits, int64_t n, int64_t m,
                        uint8_t** out_s, int32_t** out_o) {
  int32_t* no = (int32_t*)malloc((size_t)(m + 1) * 4);
  int64_t bytes = 0;
  for (int64_t r = 0; r < n; ++r)
    if ((bits[r >> 3] >> (r & 7)) & 1) bytes += o[r + 1] - o[r];
  uint8_t* ns = (uint8_t*)malloc((size_t)(bytes ? bytes : 1));
  int64_t k = 0, p = 0;
  no[0] = 0;
  for (int64_t r = 0; r < n; ++r)
    if ((bits[r >> 3] >> (r & 7)) & 1) {
      const int32_t len = o[r + 1] - o[r];
      memcpy(ns + p, s + o[r], (size_t)len);
      p += len;
      no[++k] = (int32_t)p;
    }
  *out_s = ns;
  *out_o = no;
}

static void filter_project_task(void* a, int64_t bi) {
  fpa_job* j = (fpa_job*)a;
  const int64_t lo = bi * j->batch_rows;
  const int64_t n = lo + j->batch_rows < j->n ? j->batch_rows : j->n - lo;
  /* evaluate_predicate: Le(date, Int32) -> BooleanArray (bit-packed) */
  uint8_t* bits = (uint8_t*)calloc((size_t)((n + 7) / 8 + 1), 1);
  int64_t m = 0;
  for (int64_t r = 0; r < n; ++r)
    if (j->date[lo + r] <= j->date_max) {
      bits[r >> 3] |= (uint8_t)(1u << (r & 7));
      ++m;
    }
  /* filter_batch: every column of the batch copied */
  int32_t* d = (int32_t*)malloc((size_t)(m ? m : 1) * 4);
  int64_t* pr = (int64_t*)malloc((size_t)(m ? m : 1) * 8);
  int64_t* ds = (int64_t*)malloc((size_t)(m ? m : 1) * 8);
  int64_t k = 0;
  for (int64_t r = 0; r < n; ++r)
    if ((bits[r >> 3] >> (r & 7)) & 1) {
      d[k] = j->date[lo + r];
      pr[k] = j->price[lo + r];
      ds[k] = j->disc[lo + r];
      ++k;
    }
  filter_utf8(j->fa, j->fao + lo, bits, n, m, &j->fa_f[bi], &j->fao_f[bi]);
  filter_utf8(j->fb, j->fbo + lo, bits, n, m, &j->fb_f[bi], &j->fbo_f[bi]);
  free(bits);
  double* mul;
  if (j->variant == 0) {
    /* project_batch_with_exprs: Mul(price, Sub(Literal(Int64 100), discount)) */
    int64_t* lit = (int64_t*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) lit[r] = 100;
    double* lit_f = cast_f64(lit, m);
    double* disc_f = cast_f64(ds, m);
    double* sub = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) sub[r] = lit_f[r] - disc_f[r];
    double* price_f = cast_f64(pr, m);
    mul = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) mul[r] = price_f[r] * sub[r];
    free(lit);
    free(lit_f);
    free(disc_f);
    free(sub);
    free(price_f);
  } else {
    /* Mul(Mul(price, Literal(0.01)), Sub(Literal(1.0), Mul(discount, Literal(0.01)))):
     * Float64 literals become arrays; Float64 operands are not cast again */
    double* price_f = cast_f64(pr, m);
    double* c1 = fill_f64(0.01, m);
    double* a = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) a[r] = price_f[r] * c1[r];
    double* disc_f = cast_f64(ds, m);
    double* c2 = fill_f64(0.01, m);
    double* dd = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) dd[r] = disc_f[r] * c2[r];
    double* one = fill_f64(1.0, m);
    double* sub = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) sub[r] = one[r] - dd[r];
    mul = (double*)malloc((size_t)(m ? m : 1) * 8);
    for (int64_t r = 0; r < m; ++r) mul[r] = a[r] * sub[r];
    free(price_f);
    free(c1);
    free(a);
    free(disc_f);
    free(c2);
    free(dd);
    free(one);
    free(sub);
  }
  j->m[bi] = m;
  j->date_f[bi] = d;
  j->price_f[bi] = pr;
  j->disc_f[bi] = ds;
  j->rev[bi] = mul;
}

int rc_filter_project_aggregate(const int32_t* date, const int64_t* price, const int64_t* discount,
                                const uint8_t* fa, const int32_t* fa_off, const uint8_t* fb, const int32_t* fb_off,
                                int64_t n, int32_t date_max, int32_t variant, int64_t batch_rows, int32_t threads,
                                rc_groups* out) {
  const int64_t nb = (n + batch_rows - 1) / batch_rows;
  fpa_job j;
  memset(&j, 0, sizeof(j));
  j.date = date;
  j.price = price;
  j.disc = discount;
  j.fa = fa;
  j.fb = fb;
  j.fao = fa_off;
  j.fbo = fb_off;
  j.n = n;
  j.batch_rows = batch_rows;
  j.date_max = date_max;
  j.variant = variant;
  const size_t z = (size_t)(nb ? nb : 1);
  j.m = (int64_t*)calloc(z, 8);
  j.date_f = (int32_t**)calloc(z, sizeof(void*));
  j.price_f = (int64_t**)calloc(z, sizeof(void*));
  j.disc_f = (int64_t**)calloc(z, sizeof(void*));
  j.fa_f = (uint8_t**)calloc(z, sizeof(void*));
  j.fb_f = (uint8_t**)calloc(z, sizeof(void*));
  j.fao_f = (int32_t**)calloc(z, sizeof(void*));
  j.fbo_f = (int32_t**)calloc(z, sizeof(void*));
  j.rev = (double**)calloc(z, sizeof(void*));
  run_tasks(threads, nb, filter_project_task, &j); /* Filter and Project, parallel over batches */
  abatch* b = (abatch*)calloc(z, sizeof(abatch));
  int64_t k = 0, base = 0;
  for (int64_t i = 0; i < nb; ++i) {
    if (j.m[i] == 0) continue; /* empty filter outputs are dropped */
    b[k].n = j.m[i];
    b[k].key_kind = KEY_UTF8x2;
    b[k].s0 = j.fa_f[i];
    b[k].o0 = j.fao_f[i];
    b[k].s1 = j.fb_f[i];
    b[k].o1 = j.fbo_f[i];
    for (int a = 0; a < 3; ++a) b[k].vals[a] = j.rev[i]; /* Float64 already: no cast */
    b[k].row_base = base;
    base += j.m[i];
    ++k;
  }
  const int32_t funcs[3] = {0, 1, 2};
  aggregate(b, k, 3, funcs, threads, out, NULL, NULL, 0);
  for (int64_t i = 0; i < nb; ++i) {
    free(j.date_f[i]);
    free(j.price_f[i]);
    free(j.disc_f[i]);
    free(j.fa_f[i]);
    free(j.fb_f[i]);
    free(j.fao_f[i]);
    free(j.fbo_f[i]);
    free(j.rev[i]);
  }
  free(j.m);
  free(j.date_f);
  free(j.price_f);
  free(j.disc_f);
  free(j.fa_f);
  free(j.fb_f);
  free(j.fao_f);
  free(j.fbo_f);
  free(j.rev);
  free(b);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Fused join -> aggregate as the executor runs it (executor.rs:333-458).     */
/* ------------------------------------------------------------------------- */
typedef struct {
  const rc_table* t;
  const int32_t* attr;
  const int64_t *key, *pay;
  int64_t n, batch_rows;
  int32_t threads;
  abatch* partial; /* per probe batch: its partial aggregate as a batch (attr, count, sum) */
  int32_t** p_attr;
  double **p_cnt, **p_sum;
  int64_t* pairs;
} ja_job;

static void ja_task(void* a, int64_t bi) {
  ja_job* j = (ja_job*)a;
  const int64_t lo = bi * j->batch_rows;
  const int64_t n = lo + j->batch_rows < j->n ? j->batch_rows : j->n - lo;
  vec_big pi, bv;
  probe_batch(j->t, j->key + lo, n, &pi, &bv);
  j->pairs[bi] = pi.len;
  j->partial[bi].n = 0;
  if (pi.len == 0) {
    free(pi.p);
    free(bv.p);
    return;
  }
  /* project_join_matches: take of the build attribute and the probe payload */
  const int64_t m = pi.len;
  int32_t* at = (int32_t*)malloc((size_t)m * 4);
  int64_t* py = (int64_t*)malloc((size_t)m * 8);
  for (int64_t i = 0; i < m; ++i) {
    at[i] = j->attr[bv.p[i]];
    py[i] = j->pay[lo + pi.p[i]];
  }
  free(pi.p);
  free(bv.p);
  /* hash_aggregate(projected, [COUNT(payload), SUM(payload)] by attr), one batch */
  abatch one;
  memset(&one, 0, sizeof(one));
  one.n = m;
  one.key_kind = KEY_I32;
  one.k32 = at;
  one.vals[0] = cast_f64(py, m);
  one.vals[1] = cast_f64(py, m);
  /* aggregate() over one batch is one chunk: run it on this thread */
  const int32_t funcs[2] = {1, 0}; /* COUNT, SUM */
  agg_job aj = {&one, 1, 1, 2, (gmap*)calloc(1, sizeof(gmap))};
  agg_task(&aj, 0);
  gmap* g = &aj.parts[0];
  const int64_t ng = g->n_order;
  int32_t* pa = (int32_t*)malloc((size_t)(ng ? ng : 1) * 4);
  double* pc = (double*)malloc((size_t)(ng ? ng : 1) * 8);
  double* ps = (double*)malloc((size_t)(ng ? ng : 1) * 8);
  for (int64_t i = 0; i < ng; ++i) {
    const uint64_t gh = g->order[i];
    const gval* v = (const gval*)swiss_val(&g->map, swiss_find(&g->map, map_hash(&g->map, gh), gh));
    pa[i] = at[v->row];
    pc[i] = result_of(&v->accs[0], funcs[0]);
    ps[i] = result_of(&v->accs[1], funcs[1]);
  }
  gmap_free(g);
  free(aj.parts);
  free((void*)one.vals[0]);
  free((void*)one.vals[1]);
  free(at);
  free(py);
  abatch* pb = &j->partial[bi];
  pb->n = ng;
  pb->key_kind = KEY_I32;
  pb->k32 = pa;
  pb->vals[0] = pc;
  pb->vals[1] = ps;
}

int64_t rc_join_aggregate(const rc_table* t, const int32_t* build_attr, const int64_t* probe_key,
                          const int64_t* probe_payload, int64_t n_probe, int64_t batch_rows, int32_t threads,
                          int32_t n_attr, double* count, double* sum) {
  const int64_t nb = (n_probe + batch_rows - 1) / batch_rows;
  const size_t z = (size_t)(nb ? nb : 1);
  ja_job j;
  memset(&j, 0, sizeof(j));
  j.t = t;
  j.attr = build_attr;
  j.key = probe_key;
  j.pay = probe_payload;
  j.n = n_probe;
  j.batch_rows = batch_rows;
  j.threads = threads;
  j.partial = (abatch*)calloc(z, sizeof(abatch));
  j.pairs = (int64_t*)calloc(z, 8);
  run_tasks(threads, nb, ja_task, &j); /* probe_matches_chunked + per-batch partial aggregates */
  /* merge of the partial batches: SUM of the partial COUNTs and SUMs by attr */
  abatch* b = (abatch*)calloc(z, sizeof(abatch));
  int64_t k = 0, pairs = 0;
  for (int64_t i = 0; i < nb; ++i) {
    pairs += j.pairs[i];
    if (j.partial[i].n) b[k++] = j.partial[i];
  }
  const int32_t funcs[2] = {0, 0};
  if (k) aggregate(b, k, 2, funcs, threads, NULL, count, sum, n_attr);
  for (int64_t i = 0; i < nb; ++i) {
    free((void*)j.partial[i].k32);
    free((void*)j.partial[i].vals[0]);
    free((void*)j.partial[i].vals[1]);
  }
  free(j.partial);
  free(j.pairs);
  free(b);
  return pairs;
}
