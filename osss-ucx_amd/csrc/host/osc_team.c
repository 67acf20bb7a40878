/*
 * osc_team.c - team geometry.
 *
 * A team is an affine set of global PEs: member i is global PE
 * start + i*stride (the reference's struct, src/shmemc/ucx/pe.h:95-132,
 * keeps the same geometry plus hash maps).  Splits compose the parent's
 * geometry, so start/stride are always global; the reference stores them
 * parent-relative (src/shmemc/teams.c:510-511) while its collectives treat
 * them as global (src/shcoll/src/reduce.c:230-253), which is only right for
 * teams split from WORLD.  Any stride >= 1 is accepted; the reference's
 * debug build requires a power of two (it passes log2(stride) down).
 * pSync arrays do not exist here: stream order replaces them.
 */
#include <stdlib.h>

#include "../../../include/shmem.h"
#include "osc_internal.h"

osc_team osc_team_world = {0};
osc_team osc_team_shared = {0};

shmem_team_t SHMEM_TEAM_WORLD = (shmem_team_t)&osc_team_world;
shmem_team_t SHMEM_TEAM_SHARED = (shmem_team_t)&osc_team_shared;
shmem_team_t SHMEM_TEAM_INVALID = NULL;

const osc_team *osc_team_get(void *team) {
  const osc_team *t = (const osc_team *)team;
  if (t == NULL || t->magic != OSC_TEAM_MAGIC || t->npes < 1) return NULL;
  return t;
}

void osc_team_geom(const osc_team *t, oscd_geom *g) {
  g->base = osc_g.base;
  g->slice = osc_g.slice;
  g->start = t->start;
  g->stride = t->stride;
  g->npes = t->npes;
  g->pad_ = 0;
}

/* rank of global PE `pe` in team t, or -1 */
int osc_team_rank(const osc_team *t, int pe) {
  int d = pe - t->start;
  if (d < 0 || d % t->stride != 0) return -1;
  d /= t->stride;
  return d < t->npes ? d : -1;
}
#define rank_of osc_team_rank

int shmem_team_my_pe(shmem_team_t team) {
  const osc_team *t = osc_team_get(team);
  if (t == NULL) return -1;
  return rank_of(t, osc_my_pe_get());
}

int shmem_team_n_pes(shmem_team_t team) {
  const osc_team *t = osc_team_get(team);
  return t == NULL ? -1 : t->npes;
}

int shmem_team_get_config(shmem_team_t team, long config_mask,
                          shmem_team_config_t *config) {
  const osc_team *t = osc_team_get(team);
  if (t == NULL || config == NULL) return -1;
  if (config_mask & SHMEM_TEAM_NUM_CONTEXTS) config->num_contexts = t->num_contexts;
  return 0;
}

int shmem_team_translate_pe(shmem_team_t src_team, int src_pe,
                            shmem_team_t dest_team) {
  const osc_team *s = osc_team_get(src_team);
  const osc_team *d = osc_team_get(dest_team);
  if (s == NULL || d == NULL || src_pe < 0 || src_pe >= s->npes) return -1;
  return rank_of(d, s->start + src_pe * s->stride);
}

static osc_team *new_team(const osc_team *parent, int start, int stride,
                          int size, const shmem_team_config_t *config,
                          long config_mask) {
  osc_team *t = (osc_team *)calloc(1, sizeof(*t));
  if (t == NULL) return NULL;
  t->magic = OSC_TEAM_MAGIC;
  t->start = parent->start + start * parent->stride;
  /* a one-member team accepts any stride (its last member is its first);
   * the stride then means nothing, and 1 keeps the product below in range */
  t->stride = size == 1 ? 1 : stride * parent->stride;
  t->npes = size;
  t->parent = (osc_team *)parent;
  t->num_contexts =
      (config != NULL && (config_mask & SHMEM_TEAM_NUM_CONTEXTS)) ? config->num_contexts : 0;
  return t;
}

/* Team-call semantics: the host creates the team once for all of its
 * members and gets a valid handle.  In the SPMD mode every member of the
 * parent calls, the team is created once, and PEs outside it get
 * SHMEM_TEAM_INVALID (reference src/shmemc/teams.c:489-555). */
int shmem_team_split_strided(shmem_team_t parent_team, int start, int stride,
                             int size, const shmem_team_config_t *config,
                             long config_mask, shmem_team_t *new_team_out) {
  if (osc_g.spmd)
    return osc_spmd_split_strided(parent_team, start, stride, size, config,
                                  config_mask, new_team_out);
  return osc_team_split_strided_tc(parent_team, start, stride, size, config,
                                   config_mask, new_team_out);
}

int osc_team_split_strided_tc(void *parent_team, int start, int stride,
                              int size, const void *config_v,
                              long config_mask, void **new_team_out) {
  const shmem_team_config_t *config = (const shmem_team_config_t *)config_v;
  const osc_team *p = osc_team_get(parent_team);
  osc_team *t;
  if (new_team_out == NULL) return -1;
  *new_team_out = SHMEM_TEAM_INVALID;
  if (p == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid parent team");
  if (size < 1 || start < 0 || stride < 1 || start >= p->npes ||
      (long)start + (long)(size - 1) * (long)stride >= (long)p->npes)
    return osc_set_error(OSC_ERR_ARG,
                         "split (start %d, stride %d, size %d) outside a %d-PE parent",
                         start, stride, size, p->npes);
  t = new_team(p, start, stride, size, config, config_mask);
  if (t == NULL) return osc_set_error(OSC_ERR_NOMEM, "team allocation");
  *new_team_out = (shmem_team_t)t;
  return 0;
}

/* The x team holds the calling PE's row of xrange consecutive ranks, the y
 * team its column (stride xrange), as OpenSHMEM 1.5 section 9.4.8. */
int shmem_team_split_2d(shmem_team_t parent_team, int xrange,
                        const shmem_team_config_t *xaxis_config,
                        long xaxis_mask, shmem_team_t *xaxis_team,
                        const shmem_team_config_t *yaxis_config,
                        long yaxis_mask, shmem_team_t *yaxis_team) {
  const osc_team *p = osc_team_get(parent_team);
  int me, row, col, xs, ys;
  osc_team *x, *y;
  if (osc_g.spmd)
    return osc_spmd_split_2d(parent_team, xrange, xaxis_config, xaxis_mask,
                             xaxis_team, yaxis_config, yaxis_mask,
                             yaxis_team);
  if (xaxis_team == NULL || yaxis_team == NULL) return -1;
  *xaxis_team = *yaxis_team = SHMEM_TEAM_INVALID;
  if (p == NULL || xrange < 1) return osc_set_error(OSC_ERR_ARG, "bad 2d split");
  if (xrange > p->npes) xrange = p->npes;
  me = rank_of(p, osc_my_pe_get());
  if (me < 0) return osc_set_error(OSC_ERR_TEAM, "calling PE not in parent");
  row = me / xrange;
  col = me % xrange;
  xs = (row + 1) * xrange <= p->npes ? xrange : p->npes - row * xrange;
  ys = (p->npes - col + xrange - 1) / xrange;
  x = new_team(p, row * xrange, 1, xs, xaxis_config, xaxis_mask);
  y = new_team(p, col, xrange, ys, yaxis_config, yaxis_mask);
  if (x == NULL || y == NULL) {
    free(x);
    free(y);
    return osc_set_error(OSC_ERR_NOMEM, "team allocation");
  }
  *xaxis_team = (shmem_team_t)x;
  *yaxis_team = (shmem_team_t)y;
  return 0;
}

void shmem_team_destroy(shmem_team_t team) {
  osc_team *t = (osc_team *)team;
  if (osc_team_get(team) == NULL || t->predefined) return;
  if (osc_g.spmd) {
    osc_spmd_team_destroy(team);
    return;
  }
  t->magic = 0;
  free(t);
}

/* Every row team (PEs row*xrange .. of the parent) and every column team
 * (col, col + xrange, ...) of a 2-d split, for the SPMD leader: rows has
 * ceil(P / xrange) entries, cols min(xrange, P). */
int osc_team_split_2d_all(const osc_team *p, int xrange, const void *xc,
                          long xm, const void *yc, long ym, osc_team **rows,
                          osc_team **cols) {
  const int P = p->npes;
  const int nrows = (P + xrange - 1) / xrange;
  const int ncols = xrange < P ? xrange : P;
  int r, c;
  for (r = 0; r < nrows; ++r) rows[r] = NULL;
  for (c = 0; c < ncols; ++c) cols[c] = NULL;
  for (r = 0; r < nrows; ++r) {
    const int xs = (r + 1) * xrange <= P ? xrange : P - r * xrange;
    rows[r] = new_team(p, r * xrange, 1, xs, (const shmem_team_config_t *)xc, xm);
    if (rows[r] == NULL) goto fail;
  }
  for (c = 0; c < ncols; ++c) {
    const int ys = (P - c + xrange - 1) / xrange;
    cols[c] = new_team(p, c, xrange, ys, (const shmem_team_config_t *)yc, ym);
    if (cols[c] == NULL) goto fail;
  }
  return OSC_OK;
fail:
  for (r = 0; r < nrows; ++r) free(rows[r]);
  for (c = 0; c < ncols; ++c) free(cols[c]);
  return osc_set_error(OSC_ERR_NOMEM, "team allocation");
}
