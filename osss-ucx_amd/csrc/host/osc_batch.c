/*
 * osc_batch.c - batched execution of many independent collective calls.
 *
 * Plan: calls are packed into waves (osc_batch_plan: skyline placement of
 * each call's destination rectangle, PE span x bytes, in call order), so
 * that within a wave no two calls write the same bytes.  Each call is cut
 * into tiles sized by the tuning table; tiles of one wave are grouped by
 * kernel family (reduce: element type, op, float mode, shape; data
 * movement: unit width) and each group becomes one persistent tile-loop
 * launch (osc_sched.c).  The whole launch sequence is captured once into a
 * hipGraph and replayed, so a batch of thousands of calls costs one graph
 * launch; inside the graph the launches of a wave form parallel branches
 * (osc_set_batch_branches, OSC_BATCH_BRANCHES), joined before the next
 * wave reuses the destinations.
 *
 * The reference has no counterpart: its team collectives serialise on
 * WORLD barriers (src/shcoll/src/reduce.c:1090-1100), so concurrent calls
 * of disjoint teams cannot even overlap there.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_sched.h"

#define OSC_BATCH_MAX_BRANCHES 8

struct osc_batch {
  int ncalls, nwaves;
  osc_call *calls;
  oscd_bcall *d_calls;
  oscd_tile *d_tiles;
  osc_plan plan;
  uint64_t *d_digest;
  void *graph;
  void *graph_stream;
  double alg_bytes;
  /* parallel branches of the captured graph: streams, fork/join events */
  int nbranch;
  int phases; /* osc_g.batch_phases when the graph was captured */
  /* span flags of the float MIN/MAX reductions that may split in a
   * two-phase wave: call i's at d_flags + flag_off[i] (UINT64_MAX: none) */
  uint8_t *d_flags;
  uint64_t *flag_off;
  void *aux[OSC_BATCH_MAX_BRANCHES - 1];
  void *ev_fork;
  void *ev_join[OSC_BATCH_MAX_BRANCHES - 1];
};

/* streams and events for nbranch parallel branches (created once) */
static int batch_branches(osc_batch *b, int nbranch) {
  int i, rc = OSC_OK;
  if (nbranch < 1) nbranch = 1;
  if (nbranch > OSC_BATCH_MAX_BRANCHES) nbranch = OSC_BATCH_MAX_BRANCHES;
  if (nbranch > 1 && b->ev_fork == NULL) rc = oscd_event_create(&b->ev_fork);
  for (i = 0; i < nbranch - 1 && rc == OSC_OK; ++i) {
    if (b->aux[i] == NULL) rc = oscd_stream_create(&b->aux[i]);
    if (rc == OSC_OK && b->ev_join[i] == NULL)
      rc = oscd_event_create(&b->ev_join[i]);
  }
  if (rc == OSC_OK) b->nbranch = nbranch;
  return rc;
}

static uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

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
    uint64_t need, off = 0;
    int lo, hi, p;
    if (!osc_call_team_fits(c, world_npes)) {
      free(top);
      return osc_set_error(OSC_ERR_TEAM, "call %d: team outside the world", i);
    }
    lo = c->team_start;
    hi = c->team_start + (c->team_npes - 1) * c->team_stride;
    /* a destination holds at least count bytes: this bound keeps the
     * byte count below from wrapping */
    if (c->count > dest_bytes_per_pe) {
      free(top);
      return osc_set_error(OSC_ERR_BUDGET, "call %d: count %llu > region %llu", i,
                           (unsigned long long)c->count,
                           (unsigned long long)dest_bytes_per_pe);
    }
    need = align_up(osc_call_dest_bytes(c), 256);
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

int osc_batch_create(const osc_call *calls, int ncalls, osc_batch **out) {
  osc_batch *b;
  oscd_bcall *h_calls = NULL;
  int i, rc = OSC_OK;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (out == NULL || calls == NULL || ncalls <= 0 || ncalls > 65535)
    return osc_set_error(OSC_ERR_ARG, "ncalls %d (1..65535)", ncalls);
  for (i = 0; i < ncalls; ++i) {
    rc = osc_call_validate(&calls[i], i);
    if (rc != OSC_OK) return rc;
  }
  b = (osc_batch *)calloc(1, sizeof(*b));
  if (b == NULL) return osc_set_error(OSC_ERR_NOMEM, "batch");
  osc_plan_init(&b->plan);
  b->ncalls = ncalls;
  b->calls = (osc_call *)malloc(sizeof(osc_call) * (size_t)ncalls);
  h_calls = (oscd_bcall *)calloc((size_t)ncalls, sizeof(oscd_bcall));
  if (b->calls == NULL || h_calls == NULL) {
    rc = osc_set_error(OSC_ERR_NOMEM, "batch");
    goto fail;
  }
  memcpy(b->calls, calls, sizeof(osc_call) * (size_t)ncalls);
  for (i = 0; i < ncalls; ++i) {
    osc_call_to_device(&calls[i], &h_calls[i]);
    if (calls[i].wave + 1 > b->nwaves) b->nwaves = calls[i].wave + 1;
    b->alg_bytes += osc_call_alg_bytes(&calls[i]);
  }
  rc = osc_plan_build(&b->plan, calls, ncalls, 1);
  if (rc != OSC_OK) goto fail;
  b->flag_off = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ncalls);
  if (b->flag_off == NULL) {
    rc = osc_set_error(OSC_ERR_NOMEM, "batch");
    goto fail;
  }
  {
    uint64_t total = 0;
    for (i = 0; i < ncalls; ++i) {
      const uint64_t fb = osc_call_flag_bytes(&calls[i]);
      b->flag_off[i] = fb ? total : UINT64_MAX;
      total += fb;
    }
    if (total > 0) {
      rc = oscd_malloc((void **)&b->d_flags, total);
      if (rc != OSC_OK) {
        osc_hip_error(rc, "batch span flags");
        goto fail;
      }
    }
  }
  rc = oscd_malloc((void **)&b->d_calls, sizeof(oscd_bcall) * (size_t)ncalls);
  if (rc == OSC_OK)
    rc = oscd_malloc((void **)&b->d_tiles,
                     sizeof(oscd_tile) * (b->plan.ntiles ? b->plan.ntiles : 1));
  if (rc == OSC_OK)
    rc = oscd_malloc((void **)&b->d_digest, sizeof(uint64_t) * (size_t)ncalls);
  if (rc == OSC_OK)
    rc = oscd_memcpy_h2d(b->d_calls, h_calls, sizeof(oscd_bcall) * (size_t)ncalls);
  if (rc == OSC_OK && b->plan.ntiles)
    rc = oscd_memcpy_h2d(b->d_tiles, b->plan.tiles, sizeof(oscd_tile) * b->plan.ntiles);
  if (rc != OSC_OK) {
    osc_hip_error(rc, "batch upload");
    goto fail;
  }
  free(h_calls);
  *out = b;
  return OSC_OK;
fail:
  free(h_calls);
  osc_batch_destroy(b);
  return rc;
}

int osc_batch_launch(osc_batch *b, void *stream) {
  int rc;
  if (b == NULL) return osc_set_error(OSC_ERR_ARG, "null batch");
  if (stream == NULL) stream = osc_g.stream;
  if (b->graph != NULL &&
      (b->graph_stream != stream || b->nbranch != osc_g.batch_branches ||
       b->phases != osc_g.batch_phases)) {
    oscd_graph_destroy(b->graph);
    b->graph = NULL;
  }
  if (b->graph == NULL) {
    void *exec = NULL;
    int rc2;
    rc = batch_branches(b, osc_g.batch_branches);
    if (rc != OSC_OK) return osc_hip_error(rc, "batch branch streams");
    rc = oscd_graph_begin(stream);
    if (rc != OSC_OK) return osc_hip_error(rc, "hipStreamBeginCapture");
    rc = osc_plan_enqueue_branches(&b->plan, b->calls, b->d_calls, b->d_tiles,
                                   stream, b->nbranch, b->aux, b->ev_fork,
                                   b->ev_join, b->d_flags, b->flag_off);
    rc2 = oscd_graph_end(stream, &exec);
    if (rc != OSC_OK) {
      if (exec) oscd_graph_destroy(exec);
      return osc_hip_error(rc, "batch capture");
    }
    if (rc2 != OSC_OK) return osc_hip_error(rc2, "hipGraphInstantiate");
    b->graph = exec;
    b->graph_stream = stream;
    b->phases = osc_g.batch_phases;
  }
  rc = oscd_graph_launch(b->graph, stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "hipGraphLaunch");
}

int osc_batch_launch_serial(osc_batch *b, void *stream) {
  int i, rc = OSC_OK;
  if (b == NULL) return osc_set_error(OSC_ERR_ARG, "null batch");
  if (stream == NULL) stream = osc_g.stream;
  for (i = 0; i < b->ncalls && rc == OSC_OK; ++i)
    rc = osc_call_launch_single(&b->calls[i], stream);
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
  info->nlaunches = b->plan.nlaunches;
  info->ntiles = b->plan.ntiles;
  info->algorithmic_bytes = b->alg_bytes;
  return OSC_OK;
}

int osc_batch_destroy(osc_batch *b) {
  if (b == NULL) return OSC_OK;
  if (osc_g.initialized && osc_g.stream) oscd_stream_sync(osc_g.stream);
  if (b->graph) oscd_graph_destroy(b->graph);
  {
    int i;
    for (i = 0; i < OSC_BATCH_MAX_BRANCHES - 1; ++i) {
      if (b->aux[i]) oscd_stream_destroy(b->aux[i]);
      if (b->ev_join[i]) oscd_event_destroy(b->ev_join[i]);
    }
    if (b->ev_fork) oscd_event_destroy(b->ev_fork);
  }
  if (b->d_calls) oscd_free(b->d_calls);
  if (b->d_tiles) oscd_free(b->d_tiles);
  if (b->d_digest) oscd_free(b->d_digest);
  if (b->d_flags) oscd_free(b->d_flags);
  free(b->flag_off);
  osc_plan_free(&b->plan);
  free(b->calls);
  free(b);
  return OSC_OK;
}
