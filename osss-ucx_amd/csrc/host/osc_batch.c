/*
 * osc_batch.c - batched execution of many independent collective calls.
 *
 * Plan: calls are packed into waves (osc_batch_plan: skyline placement of
 * each call's destination rectangle, PE span x bytes, in call order), so
 * that within a wave no two calls write the same bytes.  Each call is cut
 * into tiles sized by the tuning table; tiles of one wave are grouped by
 * kernel family (reduce: element type, op, float mode, shape; data
 * movement: unit width) and each group becomes one persistent tile-loop
 * launch.  The whole launch sequence is captured once into a hipGraph and
 * replayed, so a batch of thousands of calls costs one graph launch.
 *
 * The reference has no counterpart: its team collectives serialise on
 * WORLD barriers (src/shcoll/src/reduce.c:1090-1100), so concurrent calls
 * of disjoint teams cannot even overlap there.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_internal.h"

int osc_tune_bucket(uint64_t bytes_per_pe);

typedef struct {
  int wave;
  int kind;  /* OSC_CALL_REDUCE, or -1 for data movement */
  int dtype, op, mode, shape, width;
  int maxlg, maxpes;
  uint32_t grid;
  uint64_t tile0, ntiles;
} osc_launch;

struct osc_batch {
  int ncalls, nwaves, nlaunches;
  osc_call *calls;
  oscd_bcall *h_calls;
  oscd_bcall *d_calls;
  oscd_tile *d_tiles;
  uint64_t ntiles;
  osc_launch *launches;
  uint64_t *d_digest;
  void *graph;
  void *graph_stream;
  double alg_bytes;
};

static uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

static uint64_t dest_bytes(const osc_call *c) {
  switch (c->kind) {
    case OSC_CALL_REDUCE: return c->count * osc_dtype_size(c->dtype);
    case OSC_CALL_FCOLLECT:
    case OSC_CALL_ALLTOALL: return c->count * (uint64_t)c->team_npes;
  }
  return c->count;
}

static uint64_t src_bytes(const osc_call *c) {
  switch (c->kind) {
    case OSC_CALL_REDUCE: return c->count * osc_dtype_size(c->dtype);
    case OSC_CALL_ALLTOALL: return c->count * (uint64_t)c->team_npes;
  }
  return c->count;
}

static double algorithmic_bytes(const osc_call *c) {
  const double P = c->team_npes, b = (double)c->count;
  switch (c->kind) {
    case OSC_CALL_REDUCE: return 2.0 * P * b * (double)osc_dtype_size(c->dtype);
    case OSC_CALL_FCOLLECT: return P * b + P * P * b;
    case OSC_CALL_ALLTOALL: return 2.0 * P * P * b;
  }
  return b + P * b;
}

int osc_batch_plan(osc_call *calls, int ncalls, uint64_t dest_base,
                   uint64_t dest_bytes_per_pe, int *nwaves) {
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  return osc_batch_plan_world(calls, ncalls, osc_g.npes, dest_base,
                              dest_bytes_per_pe, nwaves);
}

int osc_batch_plan_world(osc_call *calls, int ncalls, int world_npes,
                         uint64_t dest_base, uint64_t dest_bytes_per_pe,
                         int *nwaves) {
  uint64_t *top;
  int i, wave = 0;
  if (calls == NULL || ncalls < 0 || world_npes < 1)
    return osc_set_error(OSC_ERR_ARG, "no calls");
  top = (uint64_t *)calloc((size_t)world_npes, sizeof(uint64_t));
  if (top == NULL) return osc_set_error(OSC_ERR_NOMEM, "planner");
  for (i = 0; i < ncalls; ++i) {
    osc_call *c = &calls[i];
    const uint64_t need = align_up(dest_bytes(c), 256);
    const int lo = c->team_start;
    const int hi = c->team_start + (c->team_npes - 1) * c->team_stride;
    uint64_t off = 0;
    int p;
    if (c->team_npes < 1 || c->team_stride < 1 || lo < 0 || hi >= world_npes) {
      free(top);
      return osc_set_error(OSC_ERR_TEAM, "call %d: team outside the world", i);
    }
    if (need > dest_bytes_per_pe) {
      free(top);
      return osc_set_error(OSC_ERR_BUDGET, "call %d needs %llu bytes per PE > region %llu",
                           i, (unsigned long long)need,
                           (unsigned long long)dest_bytes_per_pe);
    }
    for (p = lo; p <= hi; ++p)
      if (top[p] > off) off = top[p];
    if (off + need > dest_bytes_per_pe) {
      ++wave;
      memset(top, 0, sizeof(uint64_t) * (size_t)world_npes);
      off = 0;
    }
    for (p = lo; p <= hi; ++p) top[p] = off + need;
    c->dest_off = dest_base + off;
    c->wave = wave;
  }
  free(top);
  if (nwaves) *nwaves = ncalls > 0 ? wave + 1 : 0;
  return OSC_OK;
}

static int unit_width_of(uint64_t a) {
  int w = 16;
  while (w > 1 && (a % (uint64_t)w) != 0) w >>= 1;
  return w;
}

static int call_class(const osc_call *c) {
  switch (c->kind) {
    case OSC_CALL_REDUCE: return osc_dtype_size(c->dtype) <= 4 ? 0 : 1;
    case OSC_CALL_FCOLLECT: return 2;
    case OSC_CALL_ALLTOALL: return 3;
  }
  return 4;
}

static int validate(const osc_call *c, int i) {
  const int hi = c->team_start + (c->team_npes - 1) * c->team_stride;
  if (c->kind < OSC_CALL_REDUCE || c->kind > OSC_CALL_BROADCAST)
    return osc_set_error(OSC_ERR_ARG, "call %d: kind %d", i, c->kind);
  if (c->team_npes < 1 || c->team_stride < 1 || c->team_start < 0 || hi >= osc_g.npes)
    return osc_set_error(OSC_ERR_TEAM, "call %d: team outside the world", i);
  if (c->src_off > osc_g.heap_bytes || src_bytes(c) > osc_g.heap_bytes - c->src_off ||
      c->dest_off > osc_g.heap_bytes || dest_bytes(c) > osc_g.heap_bytes - c->dest_off)
    return osc_set_error(OSC_ERR_PTR, "call %d: buffers outside the heap", i);
  if (c->kind == OSC_CALL_REDUCE) {
    if (c->op < 0 || c->op >= OSC_OP_COUNT || osc_dtype_size(c->dtype) == 0)
      return osc_set_error(OSC_ERR_ARG, "call %d: op %d type %d", i, c->op, c->dtype);
    if ((c->dtype >= OSC_T_F32 && c->op <= OSC_OP_XOR) ||
        (c->dtype >= OSC_T_C32 && (c->op == OSC_OP_MIN || c->op == OSC_OP_MAX)))
      return osc_set_error(OSC_ERR_UNSUPPORTED, "call %d: op %d on type %d", i, c->op, c->dtype);
    if ((c->src_off % 16) != 0 || (c->dest_off % 16) != 0)
      return osc_set_error(OSC_ERR_ARG, "call %d: batched reduce needs 16-byte aligned offsets", i);
  }
  if (c->kind == OSC_CALL_BROADCAST && (c->root < 0 || c->root >= c->team_npes))
    return osc_set_error(OSC_ERR_ARG, "call %d: root %d", i, c->root);
  {
    const uint64_t lo_d = c->dest_off, hi_d = c->dest_off + dest_bytes(c);
    const uint64_t lo_s = c->src_off, hi_s = c->src_off + src_bytes(c);
    const int same = c->kind == OSC_CALL_REDUCE && lo_d == lo_s;
    if (!same && lo_d < hi_s && lo_s < hi_d && dest_bytes(c) > 0)
      return osc_set_error(OSC_ERR_ARG, "call %d: dest and source overlap", i);
  }
  return OSC_OK;
}

/* family key order: wave, kind class, reduce (dtype, op, mode, shape) or
 * unit width */
typedef struct {
  int wave, fam_kind, dtype, op, mode, shape, width;
  int call;
  uint64_t begin, end;
  uint32_t aux;
} tile_rec;

static int fam_cmp(const tile_rec *a, const tile_rec *b) {
  if (a->wave != b->wave) return a->wave - b->wave;
  if (a->fam_kind != b->fam_kind) return a->fam_kind - b->fam_kind;
  if (a->fam_kind == 0) {
    if (a->dtype != b->dtype) return a->dtype - b->dtype;
    if (a->op != b->op) return a->op - b->op;
    if (a->mode != b->mode) return a->mode - b->mode;
    if (a->shape != b->shape) return a->shape - b->shape;
  } else if (a->width != b->width) {
    return a->width - b->width;
  }
  return 0;
}

static int tile_sort(const void *x, const void *y) {
  const tile_rec *a = (const tile_rec *)x, *b = (const tile_rec *)y;
  int c = fam_cmp(a, b);
  if (c != 0) return c;
  if (a->call != b->call) return a->call - b->call;
  return (a->begin > b->begin) - (a->begin < b->begin);
}

typedef struct {
  tile_rec *v;
  uint64_t n, cap;
} tile_vec;

static int push(tile_vec *tv, const tile_rec *t) {
  if (tv->n == tv->cap) {
    uint64_t nc = tv->cap ? tv->cap * 2 : 4096;
    tile_rec *nv = (tile_rec *)realloc(tv->v, nc * sizeof(tile_rec));
    if (nv == NULL) return OSC_ERR_NOMEM;
    tv->v = nv;
    tv->cap = nc;
  }
  tv->v[tv->n++] = *t;
  return OSC_OK;
}

static int make_tiles(const osc_call *c, int idx, tile_vec *tv) {
  int shape = 0, tile = 1, grid = 0;
  tile_rec t;
  const uint64_t db = dest_bytes(c);
  memset(&t, 0, sizeof(t));
  osc_tune_get(call_class(c), osc_tune_bucket(db), &shape, &tile, &grid);
  t.wave = c->wave;
  t.call = idx;
  if (c->kind == OSC_CALL_REDUCE) {
    const int vec = oscd_reduce_vec(c->dtype);
    const uint64_t nunits = (c->count + (uint64_t)vec - 1) / (uint64_t)vec;
    const int mode = osc_g.fpmode[c->op][c->dtype];
    const int span = oscd_reduce_span_units(c->dtype, c->op, mode, shape);
    const uint64_t tu = (uint64_t)span * (uint64_t)tile;
    uint64_t b;
    if (span <= 0) return OSC_ERR_ARG;
    t.fam_kind = 0;
    t.dtype = c->dtype;
    t.op = c->op;
    t.mode = mode;
    t.shape = shape;
    for (b = 0; b < nunits; b += tu) {
      t.begin = b;
      t.end = b + tu < nunits ? b + tu : nunits;
      if (push(tv, &t) != OSC_OK) return OSC_ERR_NOMEM;
    }
    return OSC_OK;
  }
  {
    const int w = unit_width_of(c->src_off | c->dest_off | c->count | osc_g.slice);
    const uint64_t nunits = c->count / (uint64_t)w;
    uint64_t tu = ((uint64_t)tile << 10) / (uint64_t)w;
    const int P = c->team_npes;
    uint64_t b;
    int i;
    if (tu < 256) tu = 256;
    t.fam_kind = 1;
    t.width = w;
    if (c->kind == OSC_CALL_FCOLLECT || c->kind == OSC_CALL_BROADCAST) {
      const int nsrc = c->kind == OSC_CALL_FCOLLECT ? P : 1;
      for (i = 0; i < nsrc; ++i)
        for (b = 0; b < nunits; b += tu) {
          t.aux = (uint32_t)i;
          t.begin = b;
          t.end = b + tu < nunits ? b + tu : nunits;
          if (push(tv, &t) != OSC_OK) return OSC_ERR_NOMEM;
        }
      return OSC_OK;
    }
    if (nunits >= 256) {
      uint64_t pair;
      for (pair = 0; pair < (uint64_t)P * (uint64_t)P; ++pair)
        for (b = 0; b < nunits; b += tu) {
          t.aux = (uint32_t)pair;
          t.begin = b;
          t.end = b + tu < nunits ? b + tu : nunits;
          if (push(tv, &t) != OSC_OK) return OSC_ERR_NOMEM;
        }
    } else if (nunits > 0) {
      const uint64_t npairs = (uint64_t)P * (uint64_t)P;
      const uint64_t ppt = tu / nunits > 0 ? tu / nunits : 1;
      for (b = 0; b < npairs; b += ppt) {
        t.aux = 0;
        t.begin = b;
        t.end = b + ppt < npairs ? b + ppt : npairs;
        if (push(tv, &t) != OSC_OK) return OSC_ERR_NOMEM;
      }
    }
  }
  return OSC_OK;
}

static int lg_of(int p) {
  int lg = 0;
  while ((2 << lg) <= p) ++lg;
  return lg;
}

void osc_tune_init_from_env(void);

int osc_batch_create(const osc_call *calls, int ncalls, osc_batch **out) {
  osc_batch *b;
  tile_vec tv = {NULL, 0, 0};
  oscd_tile *h_tiles = NULL;
  int i, rc = OSC_OK;
  uint64_t k;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (out == NULL || calls == NULL || ncalls <= 0 || ncalls > 65535)
    return osc_set_error(OSC_ERR_ARG, "ncalls %d (1..65535)", ncalls);
  for (i = 0; i < ncalls; ++i) {
    rc = validate(&calls[i], i);
    if (rc != OSC_OK) return rc;
  }
  b = (osc_batch *)calloc(1, sizeof(*b));
  if (b == NULL) return osc_set_error(OSC_ERR_NOMEM, "batch");
  b->ncalls = ncalls;
  b->calls = (osc_call *)malloc(sizeof(osc_call) * (size_t)ncalls);
  b->h_calls = (oscd_bcall *)calloc((size_t)ncalls, sizeof(oscd_bcall));
  if (b->calls == NULL || b->h_calls == NULL) {
    rc = osc_set_error(OSC_ERR_NOMEM, "batch");
    goto fail;
  }
  memcpy(b->calls, calls, sizeof(osc_call) * (size_t)ncalls);
  for (i = 0; i < ncalls; ++i) {
    const osc_call *c = &calls[i];
    oscd_bcall *d = &b->h_calls[i];
    d->src_off = c->src_off;
    d->dest_off = c->dest_off;
    d->count = c->count;
    d->start = c->team_start;
    d->stride = c->team_stride;
    d->npes = c->team_npes;
    d->lg = lg_of(c->team_npes);
    d->kind = c->kind;
    d->op = c->op;
    d->dtype = c->dtype;
    d->aux = c->kind == OSC_CALL_REDUCE ? osc_g.fpmode[c->op][c->dtype] : c->root;
    if (c->wave + 1 > b->nwaves) b->nwaves = c->wave + 1;
    b->alg_bytes += algorithmic_bytes(c);
    if (c->count == 0) continue;
    rc = make_tiles(c, i, &tv);
    if (rc != OSC_OK) {
      osc_set_error(rc, "tiling call %d", i);
      goto fail;
    }
  }
  qsort(tv.v, tv.n, sizeof(tile_rec), tile_sort);
  /* group into launches */
  b->launches = (osc_launch *)calloc(tv.n ? tv.n : 1, sizeof(osc_launch));
  h_tiles = (oscd_tile *)malloc(sizeof(oscd_tile) * (tv.n ? tv.n : 1));
  if (b->launches == NULL || h_tiles == NULL) {
    rc = osc_set_error(OSC_ERR_NOMEM, "batch tiles");
    goto fail;
  }
  for (k = 0; k < tv.n; ++k) {
    const tile_rec *t = &tv.v[k];
    const osc_call *c = &calls[t->call];
    osc_launch *L;
    int grid = 2048;
    h_tiles[k].call = (uint32_t)t->call;
    h_tiles[k].aux = t->aux;
    h_tiles[k].begin = t->begin;
    h_tiles[k].end = t->end;
    if (k == 0 || fam_cmp(t, &tv.v[k - 1]) != 0) {
      L = &b->launches[b->nlaunches++];
      L->wave = t->wave;
      L->kind = t->fam_kind == 0 ? OSC_CALL_REDUCE : -1;
      L->dtype = t->dtype;
      L->op = t->op;
      L->mode = t->mode;
      L->shape = t->shape;
      L->width = t->width;
      L->tile0 = k;
      L->grid = 0;
    }
    L = &b->launches[b->nlaunches - 1];
    L->ntiles++;
    if (lg_of(c->team_npes) > L->maxlg) L->maxlg = lg_of(c->team_npes);
    if (c->team_npes > L->maxpes) L->maxpes = c->team_npes;
    osc_tune_get(call_class(c), osc_tune_bucket(dest_bytes(c)), NULL, NULL, &grid);
    if ((uint32_t)grid > L->grid) L->grid = (uint32_t)grid;
  }
  for (i = 0; i < b->nlaunches; ++i)
    if (b->launches[i].grid > b->launches[i].ntiles)
      b->launches[i].grid = (uint32_t)b->launches[i].ntiles;
  b->ntiles = tv.n;
  rc = oscd_malloc((void **)&b->d_calls, sizeof(oscd_bcall) * (size_t)ncalls);
  if (rc == OSC_OK)
    rc = oscd_malloc((void **)&b->d_tiles, sizeof(oscd_tile) * (tv.n ? tv.n : 1));
  if (rc == OSC_OK)
    rc = oscd_malloc((void **)&b->d_digest, sizeof(uint64_t) * (size_t)ncalls);
  if (rc == OSC_OK)
    rc = oscd_memcpy_h2d(b->d_calls, b->h_calls, sizeof(oscd_bcall) * (size_t)ncalls);
  if (rc == OSC_OK && tv.n)
    rc = oscd_memcpy_h2d(b->d_tiles, h_tiles, sizeof(oscd_tile) * tv.n);
  if (rc != OSC_OK) {
    osc_hip_error(rc, "batch upload");
    goto fail;
  }
  free(tv.v);
  free(h_tiles);
  *out = b;
  return OSC_OK;
fail:
  free(tv.v);
  free(h_tiles);
  osc_batch_destroy(b);
  return rc;
}

static int enqueue_all(osc_batch *b, void *stream) {
  int i, rc = OSC_OK;
  for (i = 0; i < b->nlaunches && rc == OSC_OK; ++i) {
    const osc_launch *L = &b->launches[i];
    const oscd_tile *tiles = b->d_tiles + L->tile0;
    if (L->kind == OSC_CALL_REDUCE) {
      const size_t lds = (L->dtype == OSC_T_F32 || L->dtype == OSC_T_F64) &&
                                 (L->op == OSC_OP_MIN || L->op == OSC_OP_MAX)
                             ? (size_t)L->maxpes * osc_dtype_size(L->dtype)
                             : 0;
      rc = oscd_batch_reduce(L->dtype, L->op, L->mode, L->shape, L->maxlg,
                             b->d_calls, tiles, (uint32_t)L->ntiles, L->grid,
                             lds, osc_g.base, osc_g.slice, stream);
    } else {
      rc = oscd_batch_move(L->width, b->d_calls, tiles, (uint32_t)L->ntiles,
                           L->grid, osc_g.base, osc_g.slice, stream);
    }
  }
  return rc;
}

int osc_batch_launch(osc_batch *b, void *stream) {
  int rc;
  if (b == NULL) return osc_set_error(OSC_ERR_ARG, "null batch");
  if (stream == NULL) stream = osc_g.stream;
  if (b->graph != NULL && b->graph_stream != stream) {
    oscd_graph_destroy(b->graph);
    b->graph = NULL;
  }
  if (b->graph == NULL) {
    rc = oscd_graph_begin(stream);
    if (rc != OSC_OK) return osc_hip_error(rc, "hipStreamBeginCapture");
    rc = enqueue_all(b, stream);
    {
      void *exec = NULL;
      const int rc2 = oscd_graph_end(stream, &exec);
      if (rc != OSC_OK) {
        if (exec) oscd_graph_destroy(exec);
        return osc_hip_error(rc, "batch capture");
      }
      if (rc2 != OSC_OK) return osc_hip_error(rc2, "hipGraphInstantiate");
      b->graph = exec;
      b->graph_stream = stream;
    }
  }
  rc = oscd_graph_launch(b->graph, stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "hipGraphLaunch");
}

int osc_batch_launch_serial(osc_batch *b, void *stream) {
  int i, rc = OSC_OK;
  if (b == NULL) return osc_set_error(OSC_ERR_ARG, "null batch");
  if (stream == NULL) stream = osc_g.stream;
  for (i = 0; i < b->ncalls && rc == OSC_OK; ++i) {
    const osc_call *c = &b->calls[i];
    oscd_geom g;
    g.base = osc_g.base;
    g.slice = osc_g.slice;
    g.start = c->team_start;
    g.stride = c->team_stride;
    g.npes = c->team_npes;
    g.pad_ = 0;
    if (c->count == 0) continue;
    switch (c->kind) {
      case OSC_CALL_REDUCE:
        rc = oscd_reduce(&g, c->op, c->dtype, osc_g.fpmode[c->op][c->dtype],
                         osc_g.reduce_shape, c->dest_off, c->src_off, c->count, stream);
        break;
      case OSC_CALL_FCOLLECT:
        rc = oscd_fcollect(&g, c->dest_off, c->src_off, c->count, stream);
        break;
      case OSC_CALL_ALLTOALL:
        rc = oscd_alltoall(&g, c->dest_off, c->src_off, c->count, stream);
        break;
      default:
        rc = oscd_broadcast(&g, c->root, c->dest_off, c->src_off, c->count, stream);
    }
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "serial launch");
}

int osc_batch_digests(osc_batch *b, uint64_t *digests, void *stream) {
  int rc;
  if (b == NULL || digests == NULL) return osc_set_error(OSC_ERR_ARG, "null");
  if (stream == NULL) stream = osc_g.stream;
  rc = oscd_fill_u64(b->d_digest, 0, (uint64_t)b->ncalls, stream);
  if (rc == OSC_OK)
    rc = oscd_batch_digest(b->d_calls, (uint32_t)b->ncalls, b->d_digest, osc_g.base,
                           osc_g.slice, stream);
  if (rc == OSC_OK) rc = oscd_stream_sync(stream);
  if (rc == OSC_OK)
    rc = oscd_memcpy_d2h(digests, b->d_digest, sizeof(uint64_t) * (size_t)b->ncalls);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "digest");
}

int osc_batch_get_info(const osc_batch *b, osc_batch_info *info) {
  if (b == NULL || info == NULL) return OSC_ERR_ARG;
  info->ncalls = b->ncalls;
  info->nwaves = b->nwaves;
  info->nlaunches = b->nlaunches;
  info->ntiles = b->ntiles;
  info->algorithmic_bytes = b->alg_bytes;
  return OSC_OK;
}

int osc_batch_destroy(osc_batch *b) {
  if (b == NULL) return OSC_OK;
  if (osc_g.initialized && osc_g.stream) oscd_stream_sync(osc_g.stream);
  if (b->graph) oscd_graph_destroy(b->graph);
  if (b->d_calls) oscd_free(b->d_calls);
  if (b->d_tiles) oscd_free(b->d_tiles);
  if (b->d_digest) oscd_free(b->d_digest);
  free(b->calls);
  free(b->h_calls);
  free(b->launches);
  free(b);
  return OSC_OK;
}
