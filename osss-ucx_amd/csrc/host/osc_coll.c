/*
 * osc_coll.c - argument checking and launch of the team collectives.
 *
 * Mirrors the checks of the reference's team wrappers
 * (SHCOLL_REDUCE_DEFINITION src/shcoll/src/reduce.c:1078-1102,
 * SHCOLL_FCOLLECT_TYPE_DEFINITION fcollect.c:554-584,
 * SHCOLL_ALLTOALL_TYPE_DEFINITION alltoall.c:303-326, broadcast.c:511-539),
 * which the reference only performs in debug builds, and turns each
 * failure into a return code instead of an abort.  The pWrk allocation
 * (two WORLD barriers per reduce, reduce.c:1090-1100) and the pSync
 * resets have no counterpart: the device stream orders everything.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

#include "osc_internal.h"

static int overlap(size_t a, size_t alen, size_t b, size_t blen) {
  return a < b + blen && b < a + alen;
}

/* ---- per-call records (OSC_LOG=1) ------------------------------------------
 * The reference logs each collective as a COLLECTIVES event with its
 * arguments (src/shmemu/shmemu.h:135-151, logger.c:270).  Here a record
 * also carries what the call cost: the launch is bracketed by two HIP
 * events on the call's stream, the call waits for the second, and one
 * JSON line goes to OSC_LOG_FILE (appended) or stderr:
 *   {"event": "COLLECTIVES", "call": "reduce", "op": "sum", "type": "f32",
 *    "algo": "hip", "team": [start, stride, npes], "count": n,
 *    "bytes": algorithmic bytes, "us": ..., "GBps": ..., "pct_peak": ...}
 * Algorithmic bytes: reduce 2*P*n*s, fcollect P*m + P*P*m, alltoall
 * 2*P*P*b, broadcast m + P*m, collect sum(n_i) * (P + 1), alltoalls
 * 2*P*P*n*s.  Recording serialises the calls that take it (one event pair,
 * one lock): it is an observability mode. */
#define OSC_PEAK_GBPS 8000.0
static pthread_mutex_t rec_mu = PTHREAD_MUTEX_INITIALIZER;

typedef struct {
  int on;
  void *s;
} call_rec;

static void rec_begin(call_rec *r, void *stream) {
  r->on = 0;
  if (!osc_g.log_calls) return;
  pthread_mutex_lock(&rec_mu);
  if ((osc_g.log_ev0 == NULL && oscd_event_create(&osc_g.log_ev0) != OSC_OK) ||
      (osc_g.log_ev1 == NULL && oscd_event_create(&osc_g.log_ev1) != OSC_OK) ||
      oscd_event_record(osc_g.log_ev0, stream) != OSC_OK) {
    pthread_mutex_unlock(&rec_mu);
    return;
  }
  r->on = 1;
  r->s = stream;
}

static const char *const k_op_names[OSC_OP_COUNT] = {"and", "or",  "xor", "min",
                                                     "max", "sum", "prod"};
static const char *const k_type_names[OSC_T_COUNT] = {
    "u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64", "f32", "f64", "c32", "c64"};
static const char *const k_algo_names[] = {"hip", "rec_dbl", "linear", "binomial",
                                           "rabenseifner"};

static void rec_end(call_rec *r, const char *what, const osc_team *t, int op,
                    int dtype, int fpmode, uint64_t count, double bytes, int rc) {
  float ms = 0.0f;
  FILE *f;
  if (!r->on) return;
  if (rc == OSC_OK && oscd_event_record(osc_g.log_ev1, r->s) == OSC_OK &&
      oscd_event_sync(osc_g.log_ev1) == OSC_OK &&
      oscd_event_elapsed_ms(osc_g.log_ev0, osc_g.log_ev1, &ms) == OSC_OK) {
    const double us = 1e3 * (double)ms;
    const double gbps = us > 0.0 ? bytes / (us * 1e3) : 0.0;
    const char *path = getenv("OSC_LOG_FILE");
    f = (path != NULL && path[0] != '\0') ? fopen(path, "a") : stderr;
    if (f != NULL) {
      fprintf(f, "{\"event\": \"COLLECTIVES\", \"call\": \"%s\"", what);
      if (op >= 0 && op < OSC_OP_COUNT)
        fprintf(f, ", \"op\": \"%s\"", k_op_names[op]);
      if (dtype >= 0 && dtype < OSC_T_COUNT)
        fprintf(f, ", \"type\": \"%s\"", k_type_names[dtype]);
      if (fpmode >= 0 && fpmode <= OSC_FP_RABENSEIFNER)
        fprintf(f, ", \"algo\": \"%s\"", k_algo_names[fpmode]);
      fprintf(f,
              ", \"team\": [%d, %d, %d], \"count\": %llu, \"bytes\": %.0f, "
              "\"us\": %.3f, \"GBps\": %.1f, \"pct_peak\": %.2f}\n",
              t->start, t->stride, t->npes, (unsigned long long)count, bytes, us,
              gbps, 100.0 * gbps / OSC_PEAK_GBPS);
      if (f != stderr) fclose(f);
    }
  }
  r->on = 0;
  pthread_mutex_unlock(&rec_mu);
}

int osc_reduce(void *team, int op, int dtype, void *dest, const void *src,
               size_t nreduce, void *stream) {
  return osc_reduce_mode(team, op, dtype, dest, src, nreduce, -1, stream);
}

/* fpmode < 0: the mode the registry resolved for (op, dtype) */
int osc_reduce_mode(void *team, int op, int dtype, void *dest, const void *src,
                    size_t nreduce, int fpmode, void *stream) {
  const osc_team *t = osc_team_get(team);
  const size_t es = osc_dtype_size(dtype);
  size_t doff, soff;
  oscd_geom g;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (es == 0 || op < 0 || op >= OSC_OP_COUNT)
    return osc_set_error(OSC_ERR_ARG, "bad op %d / type %d", op, dtype);
  if (dtype >= OSC_T_F32 && op <= OSC_OP_XOR)
    return osc_set_error(OSC_ERR_UNSUPPORTED, "bitwise op on a float type");
  if (dtype >= OSC_T_C32 && (op == OSC_OP_MIN || op == OSC_OP_MAX))
    return osc_set_error(OSC_ERR_UNSUPPORTED, "min/max on a complex type");
  if (nreduce > ((size_t)-1) / es) return osc_set_error(OSC_ERR_ARG, "nreduce overflow");
  rc = osc_sym_range(dest, nreduce * es, &doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, nreduce * es, &soff);
  if (rc != OSC_OK) return rc;
  /* dest == source is allowed (OpenSHMEM 1.5 9.9.10); partial overlap not */
  if (doff != soff && overlap(doff, nreduce * es, soff, nreduce * es))
    return osc_set_error(OSC_ERR_ARG, "dest and source partially overlap");
  if ((size_t)t->npes > (size_t)osc_g.npes) return osc_set_error(OSC_ERR_TEAM, "team too large");
  if (nreduce == 0) return OSC_OK;
  osc_team_geom(t, &g);
  if (fpmode < 0) fpmode = osc_g.fpmode[op][dtype];
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_reduce(&g, op, dtype, fpmode, osc_g.reduce_shape, doff, soff,
                     nreduce, s);
    rec_end(&r, "reduce", t, op, dtype, fpmode, nreduce,
            2.0 * t->npes * (double)nreduce * (double)es, rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "reduce launch");
}

int osc_fcollect(void *team, void *dest, const void *src, size_t nbytes,
                 void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff;
  oscd_geom g;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (nbytes > ((size_t)-1) / (size_t)t->npes) return osc_set_error(OSC_ERR_ARG, "size overflow");
  rc = osc_sym_range(dest, nbytes * (size_t)t->npes, &doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, nbytes, &soff);
  if (rc != OSC_OK) return rc;
  if (nbytes > 0 && overlap(doff, nbytes * (size_t)t->npes, soff, nbytes))
    return osc_set_error(OSC_ERR_ARG, "dest and source overlap");
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_fcollect(&g, doff, soff, nbytes, s);
    rec_end(&r, "fcollect", t, -1, -1, -1, nbytes,
            (double)nbytes * t->npes * (1.0 + t->npes), rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "fcollect launch");
}

int osc_alltoall(void *team, void *dest, const void *src, size_t nbytes,
                 void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff, tot;
  oscd_geom g;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (nbytes > ((size_t)-1) / (size_t)t->npes) return osc_set_error(OSC_ERR_ARG, "size overflow");
  tot = nbytes * (size_t)t->npes;
  rc = osc_sym_range(dest, tot, &doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, tot, &soff);
  if (rc != OSC_OK) return rc;
  if (tot > 0 && overlap(doff, tot, soff, tot))
    return osc_set_error(OSC_ERR_ARG, "dest and source overlap");
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_alltoall(&g, doff, soff, nbytes, s);
    rec_end(&r, "alltoall", t, -1, -1, -1, nbytes,
            2.0 * (double)nbytes * t->npes * t->npes, rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "alltoall launch");
}

int osc_broadcast_impl(void *team, void *dest, const void *src, size_t nbytes,
                       int root, int skip_root, void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff;
  oscd_geom g;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (root < 0 || root >= t->npes) return osc_set_error(OSC_ERR_ARG, "root %d", root);
  rc = osc_sym_range(dest, nbytes, &doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, nbytes, &soff);
  if (rc != OSC_OK) return rc;
  if (doff != soff && overlap(doff, nbytes, soff, nbytes))
    return osc_set_error(OSC_ERR_ARG, "dest and source partially overlap");
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_broadcast(&g, root, doff, soff, nbytes, skip_root, s);
    rec_end(&r, "broadcast", t, -1, -1, -1, nbytes,
            (double)nbytes * (1.0 + t->npes - (skip_root ? 1 : 0)), rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "broadcast launch");
}

int osc_broadcast(void *team, void *dest, const void *src, size_t nbytes,
                  int root, void *stream) {
  return osc_broadcast_impl(team, dest, src, nbytes, root, 0, stream);
}

/* The (offset, size) table of the variable-size collect lives in one
 * pinned staging buffer and one device buffer, reused once the previous
 * collect's kernel has read them; tab_mu serialises host threads. */
static pthread_mutex_t tab_mu = PTHREAD_MUTEX_INITIALIZER;

/* Widest unit (16, 8, 4, 2 or 1 bytes) dividing every offset and size of
 * a collect, the arena base and the slice. */
static int collect_unit(size_t doff, size_t soff, const size_t *nbytes,
                        int npes) {
  uint64_t align = (uint64_t)doff | (uint64_t)soff | (uint64_t)osc_g.slice |
                   (uint64_t)(uintptr_t)osc_g.base;
  size_t m = 0;
  int i, W;
  for (i = 0; i < npes; ++i) {
    align |= (uint64_t)m | (uint64_t)nbytes[i];
    m += nbytes[i];
  }
  for (W = 16; W > 1 && (align % (uint64_t)W) != 0; W >>= 1) {
  }
  return W;
}

static int collect_launch(const osc_team *t, size_t doff, size_t soff,
                          const size_t *nbytes, size_t maxn, void *s) {
  size_t m;
  oscd_geom g;
  int i, rc, W;
  if (osc_g.tab_host == NULL || osc_g.tab_dev == NULL ||
      osc_g.tab_event == NULL) {
    const size_t tb = (size_t)2 * OSC_MAX_NPES * sizeof(uint64_t);
    rc = osc_g.tab_host ? OSC_OK : oscd_host_alloc((void **)&osc_g.tab_host, tb);
    if (rc == OSC_OK && osc_g.tab_dev == NULL)
      rc = oscd_malloc((void **)&osc_g.tab_dev, tb);
    if (rc == OSC_OK && osc_g.tab_event == NULL)
      rc = oscd_event_create(&osc_g.tab_event);
    if (rc != OSC_OK) return osc_hip_error(rc, "collect table");
    osc_g.tab_used = 0;
  }
  if (osc_g.tab_used) {
    rc = oscd_event_sync(osc_g.tab_event);
    if (rc != OSC_OK) return osc_hip_error(rc, "collect table wait");
  }
  for (i = 0, m = 0; i < t->npes; ++i) {
    osc_g.tab_host[2 * i] = m;
    osc_g.tab_host[2 * i + 1] = nbytes[i];
    m += nbytes[i];
  }
  W = collect_unit(doff, soff, nbytes, t->npes);
  rc = oscd_memcpy_h2d_async(osc_g.tab_dev, osc_g.tab_host,
                             (size_t)2 * t->npes * sizeof(uint64_t), s);
  if (rc != OSC_OK) return osc_hip_error(rc, "collect table copy");
  osc_team_geom(t, &g);
  rc = oscd_collect(&g, doff, soff, osc_g.tab_dev, maxn, W, s);
  if (rc != OSC_OK) return osc_hip_error(rc, "collect launch");
  rc = oscd_event_record(osc_g.tab_event, s);
  if (rc != OSC_OK) return osc_hip_error(rc, "collect table event");
  osc_g.tab_used = 1;
  return OSC_OK;
}

/* Argument checks shared by both collect entry points; *doff, *soff and
 * *maxn describe the call (maxn == 0: nothing to move). */
static int collect_check(const osc_team *t, void *dest, const void *src,
                         const size_t *nbytes, size_t *doff, size_t *soff,
                         size_t *maxn) {
  size_t total = 0;
  int i, rc;
  *maxn = 0;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (nbytes == NULL) return osc_set_error(OSC_ERR_ARG, "NULL size table");
  for (i = 0; i < t->npes; ++i) {
    if (nbytes[i] > ((size_t)-1) - total)
      return osc_set_error(OSC_ERR_ARG, "size overflow");
    total += nbytes[i];
    if (nbytes[i] > *maxn) *maxn = nbytes[i];
  }
  rc = osc_sym_range(dest, total, doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, *maxn, soff);
  if (rc != OSC_OK) return rc;
  if (total > 0 && *maxn > 0 && overlap(*doff, total, *soff, *maxn))
    return osc_set_error(OSC_ERR_ARG, "dest and source overlap");
  return OSC_OK;
}

/* sum of a collect's per-member byte counts (checked not to wrap) */
static double collect_total(const osc_team *t, const size_t *nbytes) {
  double tot = 0.0;
  int i;
  for (i = 0; i < t->npes; ++i) tot += (double)nbytes[i];
  return tot;
}

/* Variable-size collect (collect.c:406-492 closed form): the per-member
 * byte counts are known to the one host caller, so the exclusive prefix
 * sum the reference computes with a scan over the PEs (util/scan.c:10-60)
 * is a host loop here, and the (offset, size) table goes to the device
 * through a pinned staging buffer on the call's stream.  The staging is
 * reused only after the previous collect's kernel has consumed it. */
int osc_collect(void *team, void *dest, const void *src, const size_t *nbytes,
                void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff, maxn;
  void *s = stream ? stream : osc_g.stream;
  int rc = collect_check(t, dest, src, nbytes, &doff, &soff, &maxn);
  call_rec r;
  if (rc != OSC_OK || maxn == 0) return rc;
  pthread_mutex_lock(&tab_mu);
  rec_begin(&r, s);
  rc = collect_launch(t, doff, soff, nbytes, maxn, s);
  rec_end(&r, "collect", t, -1, -1, -1, (uint64_t)collect_total(t, nbytes),
          collect_total(t, nbytes) * (1.0 + t->npes), rc);
  pthread_mutex_unlock(&tab_mu);
  return rc;
}

/* The same collect without any allocation: the table goes into the
 * library's static device array (OSCD_STATIC_TAB_MEMBERS members) through
 * kernel arguments on the call's stream, so the call neither allocates
 * nor copies from host memory.  Every call on the device shares that
 * array: calls that overlap in time on one GPU (different streams or host
 * threads without ordering) race on it. */
int osc_collect_noalloc(void *team, void *dest, const void *src,
                        const size_t *nbytes, void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff, maxn;
  oscd_geom g;
  int rc = collect_check(t, dest, src, nbytes, &doff, &soff, &maxn);
  if (rc != OSC_OK) return rc;
  if (t->npes > OSCD_STATIC_TAB_MEMBERS)
    return osc_set_error(OSC_ERR_ARG,
                         "team of %d members: the allocation-free collect takes"
                         " at most %d", t->npes, OSCD_STATIC_TAB_MEMBERS);
  if (maxn == 0) return OSC_OK;
  osc_team_geom(t, &g);
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_collect_static(&g, doff, soff, (const uint64_t *)nbytes, maxn,
                             collect_unit(doff, soff, nbytes, t->npes),
                             osc_g.device, s);
    rec_end(&r, "collect_noalloc", t, -1, -1, -1,
            (uint64_t)collect_total(t, nbytes),
            collect_total(t, nbytes) * (1.0 + t->npes), rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "collect launch");
}


/* Strided alltoall (alltoalls.c:54-90 closed form); dst/sst in elements
 * of elem_bytes, both >= 1 as the reference's checks require
 * (alltoalls.c:316-318). */
int osc_alltoalls(void *team, void *dest, const void *src, ptrdiff_t dst,
                  ptrdiff_t sst, size_t elem_bytes, size_t nelems,
                  void *stream) {
  const osc_team *t = osc_team_get(team);
  size_t doff, soff, need_d, need_s, span;
  oscd_geom g;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (dst < 1 || sst < 1 || elem_bytes == 0)
    return osc_set_error(OSC_ERR_ARG, "strides %td/%td, element %zu bytes",
                         dst, sst, elem_bytes);
  if (nelems == 0) return OSC_OK;
  /* the last element touched is ((P*n - 1)*stride)*es + es */
  span = (size_t)t->npes * nelems;
  if (nelems > ((size_t)-1) / (size_t)t->npes ||
      span - 1 > (((size_t)-1) - elem_bytes) / elem_bytes / (size_t)dst ||
      span - 1 > (((size_t)-1) - elem_bytes) / elem_bytes / (size_t)sst)
    return osc_set_error(OSC_ERR_ARG, "size overflow");
  need_d = (span - 1) * (size_t)dst * elem_bytes + elem_bytes;
  need_s = (span - 1) * (size_t)sst * elem_bytes + elem_bytes;
  rc = osc_sym_range(dest, need_d, &doff);
  if (rc != OSC_OK) return rc;
  rc = osc_sym_range(src, need_s, &soff);
  if (rc != OSC_OK) return rc;
  if (overlap(doff, need_d, soff, need_s))
    return osc_set_error(OSC_ERR_ARG, "dest and source overlap");
  osc_team_geom(t, &g);
  {
    void *s = stream ? stream : osc_g.stream;
    call_rec r;
    rec_begin(&r, s);
    rc = oscd_alltoalls(&g, doff, soff, elem_bytes, (uint64_t)dst,
                        (uint64_t)sst, nelems, s);
    rec_end(&r, "alltoalls", t, -1, -1, -1, nelems,
            2.0 * (double)nelems * (double)elem_bytes * t->npes * t->npes, rc);
  }
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "alltoalls launch");
}

/* Active set checks of the reference's sized wrappers
 * (SHMEMU_CHECK_ACTIVE_SET_RANGE, e.g. reduce.c:936-940). */
int osc_active_set(int PE_start, int logPE_stride, int PE_size, osc_team *t) {
  long stride, last;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (PE_size < 1 || PE_start < 0 || logPE_stride < 0 || logPE_stride > 30)
    return osc_set_error(OSC_ERR_TEAM, "active set (%d, %d, %d)", PE_start,
                         logPE_stride, PE_size);
  stride = 1L << logPE_stride;
  last = (long)PE_start + (long)(PE_size - 1) * stride;
  if (last >= osc_g.npes)
    return osc_set_error(OSC_ERR_TEAM, "active set (%d, %d, %d) past PE %d",
                         PE_start, logPE_stride, PE_size, osc_g.npes - 1);
  t->magic = OSC_TEAM_MAGIC;
  t->start = PE_start;
  t->stride = (int)stride;
  t->npes = PE_size;
  t->predefined = 0;
  t->num_contexts = 0;
  t->parent = NULL;
  return OSC_OK;
}
