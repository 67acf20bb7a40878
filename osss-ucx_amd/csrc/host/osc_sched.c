/*
 * osc_sched.c - tile plans shared by the batched and served modes.
 *
 * Every call is cut into tiles (whole block iterations of the reduce
 * kernel, or runs of 16-byte units for data movement) sized by the tuning
 * table.  All tiles of a call belong to one kernel family, so a plan sorts
 * only the calls by (wave, family), counts each call's tiles in closed
 * form and writes them straight into the caller's tile array; each run of
 * one family becomes one launch of the persistent tile-loop kernel, with a
 * grid capped by the table.  Calls whose table row says tile 0 get a launch
 * of their own through the per-call kernel instead, ordered with the rest
 * of their wave.  Planning cost is O(calls log calls + tiles).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_sched.h"

uint64_t osc_call_dest_bytes(const osc_call *c) {
  switch (c->kind) {
    case OSC_CALL_REDUCE: return c->count * osc_dtype_size(c->dtype);
    case OSC_CALL_FCOLLECT:
    case OSC_CALL_ALLTOALL: return c->count * (uint64_t)c->team_npes;
  }
  return c->count;
}

uint64_t osc_call_src_bytes(const osc_call *c) {
  switch (c->kind) {
    case OSC_CALL_REDUCE: return c->count * osc_dtype_size(c->dtype);
    case OSC_CALL_ALLTOALL: return c->count * (uint64_t)c->team_npes;
  }
  return c->count;
}

double osc_call_alg_bytes(const osc_call *c) {
  const double P = c->team_npes, b = (double)c->count;
  switch (c->kind) {
    case OSC_CALL_REDUCE: return 2.0 * P * b * (double)osc_dtype_size(c->dtype);
    case OSC_CALL_FCOLLECT: return P * b + P * P * b;
    case OSC_CALL_ALLTOALL: return 2.0 * P * P * b;
  }
  return b + P * b;
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

int osc_call_team_fits(const osc_call *c, int world_npes) {
  /* in 64 bits: a large stride must not wrap the last member's PE */
  const long long hi =
      (long long)c->team_start + (long long)(c->team_npes - 1) * (long long)c->team_stride;
  return c->team_npes >= 1 && c->team_stride >= 1 && c->team_start >= 0 && hi < world_npes;
}

int osc_call_validate(const osc_call *c, int i) {
  if (c->kind < OSC_CALL_REDUCE || c->kind > OSC_CALL_BROADCAST)
    return osc_set_error(OSC_ERR_ARG, "call %d: kind %d", i, c->kind);
  if (!osc_call_team_fits(c, osc_g.npes))
    return osc_set_error(OSC_ERR_TEAM, "call %d: team outside the world", i);
  /* every kind's destination holds at least count bytes; bounding count
   * first keeps the byte counts below (count x size, count x P) exact */
  if (c->count > osc_g.heap_bytes)
    return osc_set_error(OSC_ERR_PTR, "call %d: count %llu larger than the heap", i,
                         (unsigned long long)c->count);
  if (c->src_off > osc_g.heap_bytes || osc_call_src_bytes(c) > osc_g.heap_bytes - c->src_off ||
      c->dest_off > osc_g.heap_bytes || osc_call_dest_bytes(c) > osc_g.heap_bytes - c->dest_off)
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
    const uint64_t lo_d = c->dest_off, hi_d = c->dest_off + osc_call_dest_bytes(c);
    const uint64_t lo_s = c->src_off, hi_s = c->src_off + osc_call_src_bytes(c);
    const int same = c->kind == OSC_CALL_REDUCE && lo_d == lo_s;
    if (!same && lo_d < hi_s && lo_s < hi_d && osc_call_dest_bytes(c) > 0)
      return osc_set_error(OSC_ERR_ARG, "call %d: dest and source overlap", i);
  }
  return OSC_OK;
}

/* ---- per-call tiling geometry ------------------------------------------------ */

/* The family a call's tiles run in and how they are cut. */
typedef struct {
  int wave, fam_kind; /* 0 reduce, 1 data movement, 2 direct (no tiles) */
  int dtype, op, mode, shape, width;
  int grid;
  uint64_t nunits;  /* units of the call (per block for data movement) */
  uint64_t tu;      /* units per tile (pairs per tile: small alltoall) */
  uint64_t ntiles;
  int npieces;      /* reduce: 1; fcollect: P sources; broadcast: 1;
                       alltoall: P*P pairs (one-pair tiles) or 1 */
  int small_a2a;    /* alltoall with < 256 units per block */
} call_geo;

static void call_geometry(const osc_call *c, call_geo *g) {
  int shape = 0, tile = 1, grid = 2048;
  memset(g, 0, sizeof(*g));
  g->wave = c->wave;
  osc_tune_get(call_class(c), osc_tune_bucket(osc_call_dest_bytes(c)), &shape, &tile, &grid);
  g->grid = grid;
  if (c->count == 0) return;
  if (tile == 0) { /* the call runs alone through the per-call kernel */
    g->fam_kind = 2;
    return;
  }
  if (c->kind == OSC_CALL_REDUCE) {
    const int vec = oscd_reduce_vec(c->dtype);
    const int mode = osc_g.fpmode[c->op][c->dtype];
    const int span = oscd_reduce_span_units(c->dtype, c->op, mode, shape);
    g->fam_kind = 0;
    g->dtype = c->dtype;
    g->op = c->op;
    g->mode = mode;
    g->shape = shape;
    g->nunits = (c->count + (uint64_t)vec - 1) / (uint64_t)vec;
    g->tu = (uint64_t)(span > 0 ? span : 256) * (uint64_t)tile;
    g->npieces = 1;
    g->ntiles = (g->nunits + g->tu - 1) / g->tu;
    return;
  }
  {
    const int w = unit_width_of(c->src_off | c->dest_off | c->count | osc_g.slice);
    const uint64_t P = (uint64_t)c->team_npes;
    uint64_t tu = ((uint64_t)tile << 10) / (uint64_t)w;
    if (tu < 256) tu = 256;
    g->fam_kind = 1;
    g->width = w;
    g->nunits = c->count / (uint64_t)w;
    g->tu = tu;
    if (c->kind == OSC_CALL_FCOLLECT || c->kind == OSC_CALL_BROADCAST) {
      g->npieces = c->kind == OSC_CALL_FCOLLECT ? (int)P : 1;
      g->ntiles = (uint64_t)g->npieces * ((g->nunits + tu - 1) / tu);
    } else if (g->nunits >= 256) {
      g->npieces = (int)(P * P);
      g->ntiles = P * P * ((g->nunits + tu - 1) / tu);
    } else if (g->nunits > 0) {
      g->small_a2a = 1;
      g->tu = tu / g->nunits > 0 ? tu / g->nunits : 1; /* pairs per tile */
      g->ntiles = (P * P + g->tu - 1) / g->tu;
    }
  }
}

/* write the call's g->ntiles tiles to out */
static void emit_tiles(const osc_call *c, const call_geo *g, uint32_t idx,
                       oscd_tile *out) {
  uint64_t k = 0, b;
  int piece;
  if (g->ntiles == 0) return;
  if (g->small_a2a) {
    const uint64_t npairs = (uint64_t)c->team_npes * (uint64_t)c->team_npes;
    for (b = 0; b < npairs; b += g->tu, ++k) {
      out[k].call = idx;
      out[k].aux = 0;
      out[k].begin = b;
      out[k].end = b + g->tu < npairs ? b + g->tu : npairs;
    }
    return;
  }
  for (piece = 0; piece < g->npieces; ++piece)
    for (b = 0; b < g->nunits; b += g->tu, ++k) {
      out[k].call = idx;
      out[k].aux = (uint32_t)piece;
      out[k].begin = b;
      out[k].end = b + g->tu < g->nunits ? b + g->tu : g->nunits;
    }
}

static int lg_of(int p) {
  int lg = 0;
  while ((2 << lg) <= p) ++lg;
  return lg;
}

void osc_call_to_device(const osc_call *c, oscd_bcall *d) {
  memset(d, 0, sizeof(*d));
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
}

uint64_t osc_call_tile_count(const osc_call *c) {
  call_geo g;
  call_geometry(c, &g);
  return g.ntiles;
}

void osc_plan_init(osc_plan *p) { memset(p, 0, sizeof(*p)); }

void osc_plan_free(osc_plan *p) {
  free(p->launches);
  free(p->tiles);
  memset(p, 0, sizeof(*p));
}

static int geo_cmp(const call_geo *a, const call_geo *b) {
  if (a->wave != b->wave) return a->wave - b->wave;
  if (a->fam_kind != b->fam_kind) return a->fam_kind - b->fam_kind;
  if (a->fam_kind == 0) {
    if (a->dtype != b->dtype) return a->dtype - b->dtype;
    if (a->op != b->op) return a->op - b->op;
    if (a->mode != b->mode) return a->mode - b->mode;
    return a->shape - b->shape;
  }
  return a->width - b->width;
}

static const call_geo *g_sort_geo; /* qsort context (planning is serial) */
static int order_cmp(const void *x, const void *y) {
  const int a = *(const int *)x, b = *(const int *)y;
  const int c = geo_cmp(&g_sort_geo[a], &g_sort_geo[b]);
  return c != 0 ? c : a - b;
}

int osc_plan_build_into(osc_plan *p, const osc_call *calls, int ncalls,
                        int use_waves, oscd_tile *out, uint64_t cap) {
  call_geo *geo = NULL;
  int *order = NULL;
  uint64_t total = 0, k = 0;
  int i, rc = OSC_OK;
  p->nlaunches = 0;
  p->ntiles = 0;
  geo = (call_geo *)calloc((size_t)(ncalls ? ncalls : 1), sizeof(call_geo));
  order = (int *)calloc((size_t)(ncalls ? ncalls : 1), sizeof(int));
  if (geo == NULL || order == NULL) {
    rc = osc_set_error(OSC_ERR_NOMEM, "plan");
    goto out;
  }
  for (i = 0; i < ncalls; ++i) {
    call_geometry(&calls[i], &geo[i]);
    if (!use_waves) geo[i].wave = 0;
    order[i] = i;
    total += geo[i].ntiles;
  }
  g_sort_geo = geo;
  qsort(order, (size_t)ncalls, sizeof(int), order_cmp);
  g_sort_geo = NULL;
  if (out == NULL) {
    if (total > p->cap_tiles) {
      oscd_tile *nt = (oscd_tile *)realloc(p->tiles, sizeof(oscd_tile) * (total ? total : 1));
      if (nt == NULL) {
        rc = osc_set_error(OSC_ERR_NOMEM, "plan tiles");
        goto out;
      }
      p->tiles = nt;
      p->cap_tiles = total;
    }
    out = p->tiles;
  } else if (total > cap) {
    rc = osc_set_error(OSC_ERR_BUDGET, "plan needs %llu tiles > %llu",
                       (unsigned long long)total, (unsigned long long)cap);
    goto out;
  }
  for (i = 0; i < ncalls; ++i) {
    const int ci = order[i];
    const call_geo *g = &geo[ci];
    const osc_call *c = &calls[ci];
    osc_launch *L;
    if (g->fam_kind == 2 && c->count > 0) {
      if (p->nlaunches == p->cap_launches) {
        const int nc = p->cap_launches ? 2 * p->cap_launches : 64;
        osc_launch *nl = (osc_launch *)realloc(p->launches, sizeof(osc_launch) * (size_t)nc);
        if (nl == NULL) {
          rc = osc_set_error(OSC_ERR_NOMEM, "plan launches");
          goto out;
        }
        p->launches = nl;
        p->cap_launches = nc;
      }
      L = &p->launches[p->nlaunches++];
      memset(L, 0, sizeof(*L));
      L->wave = g->wave;
      L->kind = OSC_LAUNCH_DIRECT;
      L->call = ci;
      L->tile0 = k;
      L->bytes = osc_call_alg_bytes(c);
      continue;
    }
    if (g->ntiles == 0) continue;
    if (p->nlaunches == 0 || geo_cmp(g, &geo[order[i - 1]]) != 0 ||
        p->launches[p->nlaunches - 1].kind == OSC_LAUNCH_DIRECT ||
        p->launches[p->nlaunches - 1].ntiles == 0) {
      if (p->nlaunches == p->cap_launches) {
        const int nc = p->cap_launches ? 2 * p->cap_launches : 64;
        osc_launch *nl = (osc_launch *)realloc(p->launches, sizeof(osc_launch) * (size_t)nc);
        if (nl == NULL) {
          rc = osc_set_error(OSC_ERR_NOMEM, "plan launches");
          goto out;
        }
        p->launches = nl;
        p->cap_launches = nc;
      }
      L = &p->launches[p->nlaunches++];
      memset(L, 0, sizeof(*L));
      L->wave = g->wave;
      L->kind = g->fam_kind == 0 ? OSC_CALL_REDUCE : OSC_LAUNCH_MOVE;
      L->dtype = g->dtype;
      L->op = g->op;
      L->mode = g->mode;
      L->shape = g->shape;
      L->width = g->width;
      L->tile0 = k;
    }
    L = &p->launches[p->nlaunches - 1];
    emit_tiles(c, g, (uint32_t)ci, out + k);
    k += g->ntiles;
    L->ntiles += g->ntiles;
    L->bytes += osc_call_alg_bytes(c);
    if (lg_of(c->team_npes) > L->maxlg) L->maxlg = lg_of(c->team_npes);
    if (c->team_npes > L->maxpes) L->maxpes = c->team_npes;
    if ((uint32_t)g->grid > L->grid) L->grid = (uint32_t)g->grid;
  }
  for (i = 0; i < p->nlaunches; ++i)
    if (p->launches[i].kind != OSC_LAUNCH_DIRECT &&
        p->launches[i].grid > p->launches[i].ntiles)
      p->launches[i].grid = (uint32_t)p->launches[i].ntiles;
  p->ntiles = k;
out:
  free(geo);
  free(order);
  return rc;
}

int osc_plan_build(osc_plan *p, const osc_call *calls, int ncalls,
                   int use_waves) {
  return osc_plan_build_into(p, calls, ncalls, use_waves, NULL, 0);
}

static int enqueue_one(const osc_launch *L, const osc_call *calls,
                       const oscd_bcall *d_calls, const oscd_tile *d_tiles,
                       void *stream) {
  const oscd_tile *tiles = d_tiles + L->tile0;
  if (L->kind == OSC_LAUNCH_DIRECT)
    return osc_call_launch_single(&calls[L->call], stream);
  if (L->kind == OSC_CALL_REDUCE) {
    const size_t lds = (L->dtype == OSC_T_F32 || L->dtype == OSC_T_F64) &&
                               (L->op == OSC_OP_MIN || L->op == OSC_OP_MAX)
                           ? (size_t)L->maxpes * osc_dtype_size(L->dtype)
                           : 0;
    return oscd_batch_reduce(L->dtype, L->op, L->mode, L->shape, L->maxlg,
                             d_calls, tiles, (uint32_t)L->ntiles, L->grid,
                             lds, osc_g.base, osc_g.slice, stream);
  }
  return oscd_batch_move(L->width, d_calls, tiles, (uint32_t)L->ntiles,
                         L->grid, osc_g.base, osc_g.slice, stream);
}

int osc_plan_enqueue(const osc_plan *p, const osc_call *calls,
                     const oscd_bcall *d_calls, const oscd_tile *d_tiles,
                     void *stream) {
  int i, rc = OSC_OK;
  for (i = 0; i < p->nlaunches && rc == OSC_OK; ++i)
    rc = enqueue_one(&p->launches[i], calls, d_calls, d_tiles, stream);
  return rc;
}

#define OSC_MAX_BRANCHES 8

static int is_reduce_launch(const osc_launch *L, const osc_call *calls) {
  return L->kind == OSC_CALL_REDUCE ||
         (L->kind == OSC_LAUNCH_DIRECT && calls[L->call].kind == OSC_CALL_REDUCE);
}

/* Which launches of a wave a pass issues: -1 all of each; 0 the
 * reductions (a split call's walk only); 1 the replicates of split
 * reductions (nothing for one that did not split) and the data movement. */
static int in_pass(const osc_launch *L, const osc_call *calls, int pass) {
  if (pass < 0) return 1;
  if (pass == 0) return is_reduce_launch(L, calls);
  return L->kind == OSC_LAUNCH_DIRECT || L->kind == OSC_LAUNCH_MOVE;
}

/* One pass over the launches [i0, i1) of a wave: forked onto nbranch
 * streams, largest first onto the least loaded branch, joined at the end. */
uint64_t osc_call_flag_bytes(const osc_call *c) {
  const uint64_t bytes = c->count * (uint64_t)osc_dtype_size(c->dtype);
  if (c->kind != OSC_CALL_REDUCE ||
      (c->dtype != OSC_T_F32 && c->dtype != OSC_T_F64) ||
      (c->op != OSC_OP_MIN && c->op != OSC_OP_MAX) || bytes < (1u << 20))
    return 0;
  return ((bytes + 15) / 16 + 255) / 256;
}

static int enqueue_pass(const osc_plan *p, int i0, int i1, int pass,
                        const osc_call *calls, const oscd_bcall *d_calls,
                        const oscd_tile *d_tiles, void *stream, int nbranch,
                        void *const *aux, void *ev_fork, void *const *ev_join,
                        uint8_t *d_flags, const uint64_t *flag_off) {
  double load[OSC_MAX_BRANCHES] = {0};
  int i, b, rc;
  char *done = (char *)calloc((size_t)(i1 - i0), 1);
  if (done == NULL) return osc_set_error(OSC_ERR_NOMEM, "branches");
  rc = oscd_event_record(ev_fork, stream);
  for (b = 1; b < nbranch && rc == OSC_OK; ++b)
    rc = oscd_stream_wait_event(aux[b - 1], ev_fork);
  for (i = i0; i < i1; ++i)
    if (!in_pass(&p->launches[i], calls, pass)) done[i - i0] = 1;
  /* launches of a wave are adjacent and few hundred at most: a selection
   * pass per launch */
  for (i = i0; i < i1 && rc == OSC_OK; ++i) {
    int j, best = -1, bb = 0;
    double w;
    for (j = i0; j < i1; ++j)
      if (!done[j - i0] &&
          (best < 0 || p->launches[j].bytes > p->launches[best].bytes))
        best = j;
    if (best < 0) break;
    for (b = 1; b < nbranch; ++b)
      if (load[b] < load[bb]) bb = b;
    done[best - i0] = 1;
    /* a split reduction's walk reads, its replicate writes: half each */
    w = p->launches[best].bytes;
    if (pass >= 0 && p->launches[best].kind == OSC_LAUNCH_DIRECT &&
        calls[p->launches[best].call].kind == OSC_CALL_REDUCE)
      w *= 0.5;
    load[bb] += w;
    if (pass >= 0 && is_reduce_launch(&p->launches[best], calls)) {
      const osc_launch *L = &p->launches[best];
      oscd_set_split_phase(pass);
      if (L->kind == OSC_LAUNCH_DIRECT && d_flags != NULL &&
          flag_off != NULL && flag_off[L->call] != UINT64_MAX)
        oscd_set_split_flags(d_flags + flag_off[L->call]);
    }
    rc = enqueue_one(&p->launches[best], calls, d_calls, d_tiles,
                     bb == 0 ? stream : aux[bb - 1]);
    oscd_set_split_phase(-1);
    oscd_set_split_flags(NULL);
  }
  free(done);
  for (b = 1; b < nbranch && rc == OSC_OK; ++b) {
    rc = oscd_event_record(ev_join[b - 1], aux[b - 1]);
    if (rc == OSC_OK) rc = oscd_stream_wait_event(stream, ev_join[b - 1]);
  }
  return rc;
}

int osc_plan_enqueue_branches(const osc_plan *p, const osc_call *calls,
                              const oscd_bcall *d_calls,
                              const oscd_tile *d_tiles, void *stream,
                              int nbranch, void *const *aux, void *ev_fork,
                              void *const *ev_join, uint8_t *d_flags,
                              const uint64_t *flag_off) {
  int i0 = 0, rc = OSC_OK;
  if (nbranch > OSC_MAX_BRANCHES) nbranch = OSC_MAX_BRANCHES;
  if (nbranch <= 1) return osc_plan_enqueue(p, calls, d_calls, d_tiles, stream);
  while (i0 < p->nlaunches && rc == OSC_OK) {
    int i1 = i0;
    while (i1 < p->nlaunches && p->launches[i1].wave == p->launches[i0].wave)
      ++i1;
    if (i1 - i0 == 1) { /* nothing to overlap */
      rc = enqueue_one(&p->launches[i0], calls, d_calls, d_tiles, stream);
    } else if (osc_g.batch_phases) {
      /* the wave in two phases: reads of the split reductions (and the
       * reductions that stay whole), then their replicates with the data
       * movement, so each phase's traffic runs mostly one way.  Measured on
       * MI355X (the 4096-call mix on 128 members, alternated in one
       * process): 107.8-107.9 ms against 110.1-110.2 ms in one phase; the
       * large alltoalls in the first phase instead measured the same. */
      rc = enqueue_pass(p, i0, i1, 0, calls, d_calls, d_tiles, stream,
                        nbranch, aux, ev_fork, ev_join, d_flags, flag_off);
      if (rc == OSC_OK)
        rc = enqueue_pass(p, i0, i1, 1, calls, d_calls, d_tiles, stream,
                          nbranch, aux, ev_fork, ev_join, d_flags, flag_off);
    } else {
      rc = enqueue_pass(p, i0, i1, -1, calls, d_calls, d_tiles, stream,
                        nbranch, aux, ev_fork, ev_join, NULL, NULL);
    }
    i0 = i1;
  }
  return rc;
}

int osc_call_launch_single(const osc_call *c, void *stream) {
  oscd_geom g;
  g.base = osc_g.base;
  g.slice = osc_g.slice;
  g.start = c->team_start;
  g.stride = c->team_stride;
  g.npes = c->team_npes;
  g.pad_ = 0;
  if (c->count == 0) return OSC_OK;
  switch (c->kind) {
    case OSC_CALL_REDUCE:
      return oscd_reduce(&g, c->op, c->dtype, osc_g.fpmode[c->op][c->dtype],
                         osc_g.reduce_shape, c->dest_off, c->src_off, c->count, stream);
    case OSC_CALL_FCOLLECT:
      return oscd_fcollect(&g, c->dest_off, c->src_off, c->count, stream);
    case OSC_CALL_ALLTOALL:
      return oscd_alltoall(&g, c->dest_off, c->src_off, c->count, stream);
  }
  return oscd_broadcast(&g, c->root, c->dest_off, c->src_off, c->count, 0,
                        stream);
}
