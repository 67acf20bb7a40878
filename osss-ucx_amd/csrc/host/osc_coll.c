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
#include <stdio.h>

#include "osc_internal.h"

static int overlap(size_t a, size_t alen, size_t b, size_t blen) {
  return a < b + blen && b < a + alen;
}

static void log_call(const char *what, const osc_team *t, int op, int dtype,
                     size_t n) {
  if (osc_g.log_calls)
    fprintf(stderr, "[osc] %s team(start=%d,stride=%d,npes=%d) op=%d type=%d n=%zu\n",
            what, t->start, t->stride, t->npes, op, dtype, n);
}

int osc_reduce(void *team, int op, int dtype, void *dest, const void *src,
               size_t nreduce, void *stream) {
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
  log_call("reduce", t, op, dtype, nreduce);
  if (nreduce == 0) return OSC_OK;
  osc_team_geom(t, &g);
  rc = oscd_reduce(&g, op, dtype, osc_g.fpmode[op][dtype], osc_g.reduce_shape,
                   doff, soff, nreduce, stream ? stream : osc_g.stream);
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
  log_call("fcollect", t, -1, -1, nbytes);
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  rc = oscd_fcollect(&g, doff, soff, nbytes, stream ? stream : osc_g.stream);
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
  log_call("alltoall", t, -1, -1, nbytes);
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  rc = oscd_alltoall(&g, doff, soff, nbytes, stream ? stream : osc_g.stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "alltoall launch");
}

int osc_broadcast(void *team, void *dest, const void *src, size_t nbytes,
                  int root, void *stream) {
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
  log_call("broadcast", t, -1, -1, nbytes);
  if (nbytes == 0) return OSC_OK;
  osc_team_geom(t, &g);
  rc = oscd_broadcast(&g, root, doff, soff, nbytes, stream ? stream : osc_g.stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "broadcast launch");
}
