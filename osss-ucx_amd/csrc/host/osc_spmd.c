/*
 * osc_spmd.c - SPMD calling mode: one host thread per emulated PE.
 *
 * An SPMD OpenSHMEM program calls every collective on every member PE, each
 * with its own copy of the symmetric arguments (the reference runs the
 * helper on every PE, e.g. src/shcoll/src/reduce.c:182-293).
 * osc_spmd_run starts one host thread per PE.  A collective call of a
 * member thread deposits its arguments at the member set's meeting point;
 * the last member to arrive runs the team-call path once for the whole
 * team (one launch, synchronised) and every member returns its result.
 * Members must agree on the arguments (same symmetric offsets, counts, op,
 * root), as OpenSHMEM requires; a disagreement fails the call on every
 * member instead of launching.  The per-member size of collect is the one
 * argument allowed to differ (collect.c:406-492: each PE brings its own
 * nelems).  Heap calls, barriers and team splits meet the same way.
 *
 * Meeting points are keyed by the member set (start, stride, size), so a
 * team handle and an active set with the same members meet together, as
 * the member threads are in the same call sequence either way.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/shmem.h"
#include "osc_internal.h"

typedef struct meet {
  struct meet *next;
  int start, stride, npes;
  pthread_cond_t cv;
  int arrived;
  unsigned long gen;
  int rc;
  char msg[256];
  unsigned char key[96];
  size_t key_len;
  int mismatch;
  size_t *dep;    /* per-member deposit (collect: its byte count) */
  uintptr_t *res; /* per-member results, two each (2-d split) */
  uintptr_t result;
} meet;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static meet *g_meets;
/* per PE: its thread's main has returned (set under g_mu).  A member that
 * returns without joining a call fails that call on the members waiting
 * for it, where an SPMD job over processes would hang. */
static char *g_exited;

static int member_exited(const meet *m) {
  int k;
  if (g_exited == NULL) return 0;
  for (k = 0; k < m->npes; ++k)
    if (g_exited[m->start + k * m->stride]) return 1;
  return 0;
}

static meet *meet_get(int start, int stride, int npes) {
  meet *m;
  for (m = g_meets; m != NULL; m = m->next)
    if (m->start == start && m->stride == stride && m->npes == npes) return m;
  m = (meet *)calloc(1, sizeof(*m));
  if (m == NULL) return NULL;
  m->dep = (size_t *)calloc((size_t)npes, sizeof(size_t));
  m->res = (uintptr_t *)calloc(2 * (size_t)npes, sizeof(uintptr_t));
  if (m->dep == NULL || m->res == NULL) {
    free(m->dep);
    free(m->res);
    free(m);
    return NULL;
  }
  m->start = start;
  m->stride = stride;
  m->npes = npes;
  pthread_cond_init(&m->cv, NULL);
  m->next = g_meets;
  g_meets = m;
  return m;
}

static void meets_free(void) {
  meet *m, *n;
  pthread_mutex_lock(&g_mu);
  for (m = g_meets; m != NULL; m = n) {
    n = m->next;
    pthread_cond_destroy(&m->cv);
    free(m->dep);
    free(m->res);
    free(m);
  }
  g_meets = NULL;
  pthread_mutex_unlock(&g_mu);
}

typedef int (*lead_fn)(meet *m, void *ctx);

/* Arrive at t's meeting point as member `rank` with the call's key; the
 * last arrival runs fn.  Every member returns fn's code (and message). */
static int meet_run(const osc_team *t, int rank, const void *key,
                    size_t key_len, size_t deposit, lead_fn fn, void *ctx,
                    meet **out) {
  meet *m;
  unsigned long gen;
  int rc;
  if (key_len > sizeof(m->key)) return osc_set_error(OSC_ERR_ARG, "SPMD key");
  pthread_mutex_lock(&g_mu);
  m = meet_get(t->start, t->stride, t->npes);
  if (m == NULL) {
    pthread_mutex_unlock(&g_mu);
    return osc_set_error(OSC_ERR_NOMEM, "SPMD meeting point");
  }
  gen = m->gen;
  if (m->arrived == 0) {
    memcpy(m->key, key, key_len);
    m->key_len = key_len;
    m->mismatch = 0;
  } else if (key_len != m->key_len || memcmp(m->key, key, key_len) != 0) {
    m->mismatch = 1;
  }
  m->dep[rank] = deposit;
  if (++m->arrived == t->npes) {
    osc_clear_error();
    rc = m->mismatch
             ? osc_set_error(OSC_ERR_ARG,
                             "SPMD members disagree on the call's arguments")
             : fn(m, ctx);
    m->rc = rc;
    snprintf(m->msg, sizeof(m->msg), "%s", rc != OSC_OK ? osc_last_error() : "");
    m->arrived = 0;
    m->gen++;
    pthread_cond_broadcast(&m->cv);
  } else {
    while (m->gen == gen) {
      if (member_exited(m)) { /* end the call for every waiting member */
        m->rc = osc_set_error(OSC_ERR_STATE,
                              "a member PE returned without joining the call");
        snprintf(m->msg, sizeof(m->msg), "%s", osc_last_error());
        m->arrived = 0;
        m->gen++;
        pthread_cond_broadcast(&m->cv);
        break;
      }
      pthread_cond_wait(&m->cv, &g_mu);
    }
  }
  rc = m->rc;
  if (rc != OSC_OK) osc_set_error(rc, "%s", m->msg);
  if (out != NULL) *out = m;
  pthread_mutex_unlock(&g_mu);
  return rc;
}

/* the calling PE's rank in team, or -1 (error recorded) */
static int my_rank(void *team, const osc_team **t) {
  int r;
  *t = osc_team_get(team);
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised"), -1;
  if (*t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team"), -1;
  r = osc_team_rank(*t, osc_my_pe_get());
  if (r < 0)
    osc_set_error(OSC_ERR_TEAM, "PE %d is not a member of the team",
                  osc_my_pe_get());
  return r;
}

static size_t off_of(const void *p) {
  size_t o;
  return (p != NULL && osc_sym_offset(p, NULL, &o) == OSC_OK) ? o : (size_t)-1;
}

enum { K_REDUCE = 1, K_MOVE, K_COLLECT, K_ALLTOALLS, K_SYNC, K_HEAP, K_SPLIT,
       K_SPLIT2D, K_DESTROY };

typedef struct {
  int kind, op, dtype, root, skip, pad_;
  size_t doff, soff, n, es;
  ptrdiff_t dst, sst;
} call_key;

typedef struct {
  void *team;
  int op, dtype, root, skip, move, fpmode;
  void *dest;
  const void *src;
  size_t n, es;
  ptrdiff_t dst, sst;
} call_ctx;

static int synced(int rc) {
  if (rc != OSC_OK) return rc;
  rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "stream wait");
}

static int lead_reduce(meet *m, void *c) {
  call_ctx *x = (call_ctx *)c;
  (void)m;
  if (x->dtype < 0)
    return osc_set_error(OSC_ERR_UNSUPPORTED, "long double reduction (no 80-bit type on gfx950)");
  return synced(osc_reduce_mode(x->team, x->op, x->dtype, x->dest, x->src, x->n,
                                x->fpmode, NULL));
}

int osc_spmd_reduce(void *team, int op, int dtype, void *dest,
                    const void *src, size_t n, int fpmode) {
  const osc_team *t;
  call_key k;
  call_ctx x;
  const int r = my_rank(team, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_REDUCE;
  k.op = op;
  k.dtype = dtype;
  k.doff = off_of(dest);
  k.soff = off_of(src);
  k.n = n;
  memset(&x, 0, sizeof(x));
  x.team = team;
  x.op = op;
  x.dtype = dtype;
  x.dest = dest;
  x.src = src;
  x.n = n;
  x.fpmode = fpmode;
  return meet_run(t, r, &k, sizeof(k), 0, lead_reduce, &x, NULL);
}

static int lead_move(meet *m, void *c) {
  call_ctx *x = (call_ctx *)c;
  (void)m;
  switch (x->move) {
    case OSC_SPMD_FCOLLECT:
      return synced(osc_fcollect(x->team, x->dest, x->src, x->n, NULL));
    case OSC_SPMD_ALLTOALL:
      return synced(osc_alltoall(x->team, x->dest, x->src, x->n, NULL));
    default:
      return synced(osc_broadcast_impl(x->team, x->dest, x->src, x->n, x->root,
                                       x->skip, NULL));
  }
}

int osc_spmd_move(int kind, void *team, void *dest, const void *src,
                  size_t nbytes, int root, int skip_root) {
  const osc_team *t;
  call_key k;
  call_ctx x;
  const int r = my_rank(team, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_MOVE;
  k.op = kind;
  k.root = root;
  k.skip = skip_root;
  k.doff = off_of(dest);
  k.soff = off_of(src);
  k.n = nbytes;
  memset(&x, 0, sizeof(x));
  x.team = team;
  x.move = kind;
  x.root = root;
  x.skip = skip_root;
  x.dest = dest;
  x.src = src;
  x.n = nbytes;
  return meet_run(t, r, &k, sizeof(k), 0, lead_move, &x, NULL);
}

static int lead_collect(meet *m, void *c) {
  call_ctx *x = (call_ctx *)c;
  return synced(osc_collect(x->team, x->dest, x->src, m->dep, NULL));
}

/* each member brings its own byte count; the leader passes all of them to
 * the variable-size collect */
int osc_spmd_collect(void *team, void *dest, const void *src,
                     size_t my_nbytes) {
  const osc_team *t;
  call_key k;
  call_ctx x;
  const int r = my_rank(team, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_COLLECT;
  k.doff = off_of(dest);
  k.soff = off_of(src);
  memset(&x, 0, sizeof(x));
  x.team = team;
  x.dest = dest;
  x.src = src;
  return meet_run(t, r, &k, sizeof(k), my_nbytes, lead_collect, &x, NULL);
}

static int lead_alltoalls(meet *m, void *c) {
  call_ctx *x = (call_ctx *)c;
  (void)m;
  return synced(osc_alltoalls(x->team, x->dest, x->src, x->dst, x->sst, x->es,
                              x->n, NULL));
}

int osc_spmd_alltoalls(void *team, void *dest, const void *src, ptrdiff_t dst,
                       ptrdiff_t sst, size_t elem_bytes, size_t nelems) {
  const osc_team *t;
  call_key k;
  call_ctx x;
  const int r = my_rank(team, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_ALLTOALLS;
  k.doff = off_of(dest);
  k.soff = off_of(src);
  k.dst = dst;
  k.sst = sst;
  k.es = elem_bytes;
  k.n = nelems;
  memset(&x, 0, sizeof(x));
  x.team = team;
  x.dest = dest;
  x.src = src;
  x.dst = dst;
  x.sst = sst;
  x.es = elem_bytes;
  x.n = nelems;
  return meet_run(t, r, &k, sizeof(k), 0, lead_alltoalls, &x, NULL);
}

static int lead_sync(meet *m, void *c) {
  int rc;
  (void)m;
  (void)c;
  rc = oscd_device_sync();
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "barrier");
}

/* barrier: the members meet and the device work enqueued so far completes */
int osc_spmd_sync(void *team) {
  const osc_team *t;
  call_key k;
  const int r = my_rank(team, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_SYNC;
  return meet_run(t, r, &k, sizeof(k), 0, lead_sync, NULL, NULL);
}

typedef struct {
  int kind;
  size_t off, size, align;
} heap_ctx;

static int lead_heap(meet *m, void *c) {
  heap_ctx *h = (heap_ctx *)c;
  size_t off = (size_t)-1;
  int rc = OSC_OK;
  switch (h->kind) {
    case OSC_SPMD_MALLOC:
    case OSC_SPMD_ALIGN:
      rc = osc_heap_call_alloc(h->size, h->align, 0, &off);
      break;
    case OSC_SPMD_CALLOC:
      rc = osc_heap_call_alloc(h->size, 0, 1, &off);
      break;
    case OSC_SPMD_FREE:
      if (h->off != (size_t)-1) rc = osc_heap_call_free(h->off);
      break;
    default: /* realloc */
      if (h->off == (size_t)-1)
        rc = osc_heap_call_alloc(h->size, 0, 0, &off);
      else if (h->size == 0)
        rc = osc_heap_call_free(h->off);
      else
        rc = osc_heap_call_realloc(h->off, h->size, &off);
      break;
  }
  m->result = (uintptr_t)off;
  return rc;
}

/* Heap calls meet on the world: the block is allocated once and every PE
 * gets its own copy's address. */
void *osc_spmd_heap(int kind, void *ptr, size_t a, size_t b) {
  const osc_team *t;
  call_key k;
  heap_ctx h;
  meet *m = NULL;
  size_t off;
  const int r = my_rank(&osc_team_world, &t);
  if (r < 0) return NULL;
  memset(&k, 0, sizeof(k));
  memset(&h, 0, sizeof(h));
  h.kind = kind;
  h.off = (kind == OSC_SPMD_FREE || kind == OSC_SPMD_REALLOC) ? off_of(ptr)
                                                              : (size_t)-1;
  h.size = a;
  h.align = b;
  k.kind = K_HEAP;
  k.op = kind;
  k.doff = h.off;
  k.n = a;
  k.es = b;
  if (meet_run(t, r, &k, sizeof(k), 0, lead_heap, &h, &m) != OSC_OK)
    return NULL;
  off = (size_t)m->result;
  if (off == (size_t)-1) return NULL;
  return osc_g.base + (size_t)osc_my_pe_get() * osc_g.slice + off;
}

typedef struct {
  void *parent;
  int start, stride, size, xrange;
  const void *config, *yconfig;
  long mask, ymask;
} split_ctx;

static int lead_split(meet *m, void *c) {
  split_ctx *s = (split_ctx *)c;
  void *nt = NULL;
  const int rc = osc_team_split_strided_tc(s->parent, s->start, s->stride,
                                           s->size, s->config, s->mask, &nt);
  m->result = (uintptr_t)nt;
  return rc;
}

int osc_spmd_split_strided(void *parent, int start, int stride, int size,
                           const void *config, long config_mask,
                           void **new_team) {
  const osc_team *t, *nt;
  call_key k;
  split_ctx s;
  meet *m = NULL;
  int r, rc;
  if (new_team == NULL) return -1;
  *new_team = SHMEM_TEAM_INVALID;
  r = my_rank(parent, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_SPLIT;
  k.op = start;
  k.dtype = stride;
  k.root = size;
  memset(&s, 0, sizeof(s));
  s.parent = parent;
  s.start = start;
  s.stride = stride;
  s.size = size;
  s.config = config;
  s.mask = config_mask;
  rc = meet_run(t, r, &k, sizeof(k), 0, lead_split, &s, &m);
  if (rc != OSC_OK) return rc;
  nt = osc_team_get((void *)m->result);
  if (nt != NULL && osc_team_rank(nt, osc_my_pe_get()) >= 0)
    *new_team = (void *)m->result;
  return 0;
}

static int lead_split_2d(meet *m, void *c) {
  split_ctx *s = (split_ctx *)c;
  const osc_team *p = osc_team_get(s->parent);
  int P, xr, nrows, ncols;
  osc_team **rows, **cols;
  int rank, rc;
  if (p == NULL || s->xrange < 1) return osc_set_error(OSC_ERR_ARG, "bad 2d split");
  P = p->npes;
  xr = s->xrange > P ? P : s->xrange;
  nrows = (P + xr - 1) / xr;
  ncols = xr;
  rows = (osc_team **)calloc((size_t)nrows, sizeof(*rows));
  cols = (osc_team **)calloc((size_t)ncols, sizeof(*cols));
  if (rows == NULL || cols == NULL) {
    free(rows);
    free(cols);
    return osc_set_error(OSC_ERR_NOMEM, "team allocation");
  }
  rc = osc_team_split_2d_all(p, xr, s->config, s->mask, s->yconfig, s->ymask,
                             rows, cols);
  if (rc == OSC_OK)
    for (rank = 0; rank < P; ++rank) {
      m->res[2 * rank] = (uintptr_t)rows[rank / xr];
      m->res[2 * rank + 1] = (uintptr_t)cols[rank % xr];
    }
  free(rows);
  free(cols);
  return rc;
}

int osc_spmd_split_2d(void *parent, int xrange, const void *xconfig,
                      long xmask, void **xteam, const void *yconfig,
                      long ymask, void **yteam) {
  const osc_team *t;
  call_key k;
  split_ctx s;
  meet *m = NULL;
  int r, rc;
  if (xteam == NULL || yteam == NULL) return -1;
  *xteam = *yteam = SHMEM_TEAM_INVALID;
  r = my_rank(parent, &t);
  if (r < 0) return osc_last_error_code();
  memset(&k, 0, sizeof(k));
  k.kind = K_SPLIT2D;
  k.op = xrange;
  memset(&s, 0, sizeof(s));
  s.parent = parent;
  s.xrange = xrange;
  s.config = xconfig;
  s.mask = xmask;
  s.yconfig = yconfig;
  s.ymask = ymask;
  rc = meet_run(t, r, &k, sizeof(k), 0, lead_split_2d, &s, &m);
  if (rc != OSC_OK) return rc;
  *xteam = (void *)m->res[2 * r];
  *yteam = (void *)m->res[2 * r + 1];
  return 0;
}

static int lead_destroy(meet *m, void *c) {
  osc_team *t = (osc_team *)c;
  (void)m;
  t->magic = 0;
  free(t);
  return OSC_OK;
}

void osc_spmd_team_destroy(void *team) {
  const osc_team *t;
  call_key k;
  const int r = my_rank(team, &t);
  if (r < 0) return;
  memset(&k, 0, sizeof(k));
  k.kind = K_DESTROY;
  k.n = (size_t)(uintptr_t)team;
  meet_run(t, r, &k, sizeof(k), 0, lead_destroy, team, NULL);
}

/* ---- launcher ------------------------------------------------------------------ */

typedef struct {
  int pe, rc;
  int (*fn)(int, void *);
  void *arg;
} pe_job;

static pthread_mutex_t g_gate_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_gate_cv = PTHREAD_COND_INITIALIZER;
static int g_gate; /* 0 closed, 1 run, -1 abort */

static void *pe_thread(void *a) {
  pe_job *j = (pe_job *)a;
  int go;
  pthread_mutex_lock(&g_gate_mu);
  while (g_gate == 0) pthread_cond_wait(&g_gate_cv, &g_gate_mu);
  go = g_gate > 0;
  pthread_mutex_unlock(&g_gate_mu);
  if (!go) return NULL;
  osc_set_my_pe(j->pe);
  j->rc = j->fn(j->pe, j->arg);
  pthread_mutex_lock(&g_mu);
  g_exited[j->pe] = 1;
  {
    meet *m;
    for (m = g_meets; m != NULL; m = m->next) pthread_cond_broadcast(&m->cv);
  }
  pthread_mutex_unlock(&g_mu);
  return NULL;
}

int osc_spmd_run(int npes, size_t heap_bytes_per_pe, int device,
                 int (*pe_main)(int pe, void *arg), void *arg) {
  pthread_t *th;
  pe_job *jobs;
  int i, made = 0, rc;
  if (pe_main == NULL) return osc_set_error(OSC_ERR_ARG, "NULL PE function");
  rc = osc_init(npes, heap_bytes_per_pe, device);
  if (rc != OSC_OK) return rc;
  th = (pthread_t *)calloc((size_t)npes, sizeof(*th));
  jobs = (pe_job *)calloc((size_t)npes, sizeof(*jobs));
  g_exited = (char *)calloc((size_t)npes, 1);
  if (th == NULL || jobs == NULL || g_exited == NULL) {
    free(th);
    free(jobs);
    free(g_exited);
    g_exited = NULL;
    osc_finalize();
    return osc_set_error(OSC_ERR_NOMEM, "PE threads");
  }
  osc_g.spmd = 1;
  g_gate = 0;
  /* every PE thread waits at a gate until all exist, so a failed thread
   * creation cannot leave the others waiting for it in a collective */
  for (i = 0; i < npes; ++i) {
    jobs[i].pe = i;
    jobs[i].fn = pe_main;
    jobs[i].arg = arg;
    if (pthread_create(&th[i], NULL, pe_thread, &jobs[i]) != 0) break;
    ++made;
  }
  pthread_mutex_lock(&g_gate_mu);
  g_gate = made == npes ? 1 : -1;
  pthread_cond_broadcast(&g_gate_cv);
  pthread_mutex_unlock(&g_gate_mu);
  for (i = 0; i < made; ++i) pthread_join(th[i], NULL);
  rc = OSC_OK;
  if (made != npes)
    rc = osc_set_error(OSC_ERR_NOMEM, "created %d of %d PE threads", made, npes);
  else
    for (i = 0; i < npes; ++i)
      if (jobs[i].rc != 0) {
        rc = jobs[i].rc;
        break;
      }
  osc_g.spmd = 0;
  free(th);
  free(jobs);
  meets_free();
  pthread_mutex_lock(&g_mu);
  free(g_exited);
  g_exited = NULL;
  pthread_mutex_unlock(&g_mu);
  osc_finalize();
  return rc;
}

int osc_spmd_active(void) { return osc_g.spmd; }
