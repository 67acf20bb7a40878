/*
 * shmem_api.c - the OpenSHMEM entry points (the drop-in boundary).
 *
 * Every typed collective is a one-line adapter onto the osc_* layer,
 * generated from the type tables in <shmem.h> the way the reference
 * generates its shim (src/api/collectives/shcoll-shim.c:185-197 alltoall,
 * :618-651 fcollect, :1025-1370 reductions): same symbol set, same
 * signatures, same return convention (0 = success).  Calls are blocking:
 * the library stream is synchronised before returning, so dest is final
 * on every member when the call returns, as in the reference.
 */
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../../include/shmem.h"
#include "osc_internal.h"

_Static_assert(sizeof(long) == 8 && sizeof(long long) == 8 &&
                   sizeof(size_t) == 8 && sizeof(ptrdiff_t) == 8,
               "LP64 expected");
_Static_assert((char)-1 < 0, "char is signed on this ABI");

int osc_parse_size(const char *s, size_t *out);

/* typename -> kernel element type (long double: not on gfx950) */
#define OSC_DT_char OSC_T_I8
#define OSC_DT_schar OSC_T_I8
#define OSC_DT_short OSC_T_I16
#define OSC_DT_int OSC_T_I32
#define OSC_DT_long OSC_T_I64
#define OSC_DT_longlong OSC_T_I64
#define OSC_DT_ptrdiff OSC_T_I64
#define OSC_DT_uchar OSC_T_U8
#define OSC_DT_ushort OSC_T_U16
#define OSC_DT_uint OSC_T_U32
#define OSC_DT_ulong OSC_T_U64
#define OSC_DT_ulonglong OSC_T_U64
#define OSC_DT_int8 OSC_T_I8
#define OSC_DT_int16 OSC_T_I16
#define OSC_DT_int32 OSC_T_I32
#define OSC_DT_int64 OSC_T_I64
#define OSC_DT_uint8 OSC_T_U8
#define OSC_DT_uint16 OSC_T_U16
#define OSC_DT_uint32 OSC_T_U32
#define OSC_DT_uint64 OSC_T_U64
#define OSC_DT_size OSC_T_U64
#define OSC_DT_float OSC_T_F32
#define OSC_DT_double OSC_T_F64
#define OSC_DT_longdouble (-1)
#define OSC_DT_complexf OSC_T_C32
#define OSC_DT_complexd OSC_T_C64

/* ---- setup ----------------------------------------------------------------- */

/* shmem_finalize was called: by the process (team-call mode) or by this
 * PE's thread (SPMD mode); the same for shmem_init in the SPMD mode */
static int g_finalized;
static __thread int t_pe_inited, t_pe_finalized;

static int start_from_env(void) {
  const char *np = getenv("OSC_NPES");
  const char *hs = getenv("SHMEM_SYMMETRIC_SIZE");
  const char *dv = getenv("OSC_DEVICE");
  size_t heap = (size_t)32 << 20; /* reference default, configure.ac:23 */
  int npes = np ? atoi(np) : 8;
  int rc;
  if (osc_g.spmd) { /* the launcher initialised; every PE meets here */
    t_pe_inited = 1;
    t_pe_finalized = 0;
    return osc_spmd_sync(SHMEM_TEAM_WORLD);
  }
  if (osc_g.initialized) return OSC_OK;
  if (hs != NULL && osc_parse_size(hs, &heap) != OSC_OK)
    return osc_set_error(OSC_ERR_ARG, "bad SHMEM_SYMMETRIC_SIZE=%s", hs);
  rc = osc_init(npes, heap, dv ? atoi(dv) : 0);
  if (rc == OSC_OK) g_finalized = 0;
  return rc;
}

void shmem_init(void) {
  const int rc = start_from_env();
  if (rc != OSC_OK) { /* fatal, as the reference's init */
    fprintf(stderr, "shmem_init: %s (%s)\n", osc_strerror(rc), osc_last_error());
    abort();
  }
}

static int thread_level(int requested) {
  const int cap = osc_g.spmd ? SHMEM_THREAD_FUNNELED : SHMEM_THREAD_MULTIPLE;
  return requested < cap ? requested : cap;
}

int shmem_init_thread(int requested, int *provided) {
  if (requested < SHMEM_THREAD_SINGLE || requested > SHMEM_THREAD_MULTIPLE)
    return osc_set_error(OSC_ERR_ARG, "thread level %d", requested), -1;
  if (start_from_env() != OSC_OK) return -1;
  if (provided != NULL) *provided = thread_level(requested);
  return 0;
}

void shmem_query_thread(int *provided) {
  if (provided != NULL) *provided = thread_level(SHMEM_THREAD_MULTIPLE);
}

void shmem_finalize(void) {
  if (osc_g.spmd) {
    osc_spmd_sync(SHMEM_TEAM_WORLD);
    t_pe_finalized = 1;
    return;
  }
  if (osc_g.initialized) g_finalized = 1;
  osc_finalize();
}

int shmem_initialized(void) {
  if (osc_g.spmd) return t_pe_inited && !t_pe_finalized;
  return osc_g.initialized;
}

int shmem_finalized(void) { return osc_g.spmd ? t_pe_finalized : g_finalized; }

void shmem_global_exit(int status) {
  fflush(NULL);
  _exit(status);
}

void shmem_pcontrol(const int level, ...) { (void)level; }

/* names deprecated since OpenSHMEM 1.2 (reference src/api/deprecated/
 * init.c:21-28, ranks.c:18-32): the PE count argument of start_pes is
 * ignored, the library starts as shmem_init does */
void start_pes(int npes) {
  (void)npes;
  shmem_init();
}
int _my_pe(void) { return shmem_my_pe(); }
int _num_pes(void) { return shmem_n_pes(); }

int shmem_pe_accessible(int pe) {
  return osc_g.initialized && pe >= 0 && pe < osc_g.npes;
}

int shmem_my_pe(void) { return osc_g.initialized ? osc_my_pe_get() : -1; }
int shmem_n_pes(void) { return osc_g.initialized ? osc_g.npes : -1; }

/* One host thread acts for all PEs, so a barrier is completion of all
 * enqueued device work; in the SPMD mode the members also meet. */
void shmem_barrier_all(void) {
  if (osc_g.spmd)
    osc_spmd_sync(SHMEM_TEAM_WORLD);
  else if (osc_g.initialized)
    oscd_device_sync();
}
int shmem_team_sync(shmem_team_t team) {
  if (osc_team_get(team) == NULL) return -1;
  if (osc_g.spmd) return osc_spmd_sync(team) == OSC_OK ? 0 : -1;
  return oscd_device_sync() == OSC_OK ? 0 : -1;
}
void shmem_quiet(void) {
  if (osc_g.initialized) oscd_device_sync();
}
void shmem_fence(void) { /* stream order already orders puts */ }

void shmem_info_get_version(int *major, int *minor) {
  if (major) *major = 1;
  if (minor) *minor = 5;
}
void shmem_info_get_version_number(int *version) {
  if (version) *version = 105;
}
void shmem_info_get_vendor_version(int *major, int *minor, int *patch) {
  if (major) *major = 1;
  if (minor) *minor = 0;
  if (patch) *patch = 0;
}
void shmem_info_get_vendor_version_number(int *version) {
  if (version) *version = 10000;
}
void shmem_info_get_name(char *name) {
  if (name) strcpy(name, "osc-mi355x (OpenSHMEM collectives on gfx950)");
}

/* ---- symmetric heap ---------------------------------------------------------- */
/* The offset-level operations act for every PE at once; the entry points
 * turn the offset into the calling PE's address.  In the SPMD mode every
 * PE calls and one of them runs the operation (osc_spmd_heap). */

int osc_heap_call_alloc(size_t size, size_t align, int zero, size_t *off) {
  int rc, pe;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  pthread_mutex_lock(&osc_g.lock);
  rc = osc_heap_alloc(size, align, off);
  pthread_mutex_unlock(&osc_g.lock);
  if (rc != OSC_OK)
    return osc_set_error(rc, "symmetric heap exhausted (%zu bytes requested)", size);
  if (zero && size > 0) {
    for (pe = 0; pe < osc_g.npes; ++pe)
      oscd_memset_async(osc_g.base + (size_t)pe * osc_g.slice + *off, 0, size,
                        osc_g.stream);
    rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
    if (rc != OSC_OK) return osc_hip_error(rc, "calloc clear");
  }
  return OSC_OK;
}

int osc_heap_call_free(size_t off) {
  int rc;
  /* outstanding work of any context may still use the object */
  oscd_device_sync();
  pthread_mutex_lock(&osc_g.lock);
  rc = osc_heap_free(off);
  pthread_mutex_unlock(&osc_g.lock);
  return rc;
}

int osc_heap_call_realloc(size_t off, size_t size, size_t *new_off) {
  size_t old, keep;
  int pe, rc;
  pthread_mutex_lock(&osc_g.lock);
  rc = osc_heap_block_size(off, &old);
  pthread_mutex_unlock(&osc_g.lock);
  if (rc != OSC_OK) return osc_set_error(rc, "realloc of a non-heap block");
  rc = osc_heap_call_alloc(size, 0, 0, new_off);
  if (rc != OSC_OK) return rc;
  keep = old < size ? old : size;
  for (pe = 0; pe < osc_g.npes && keep > 0; ++pe)
    oscd_copy(osc_g.base + (size_t)pe * osc_g.slice + *new_off,
              osc_g.base + (size_t)pe * osc_g.slice + off, keep, osc_g.stream);
  return osc_heap_call_free(off);
}

static void *my_addr(size_t off) {
  return osc_g.base + (size_t)osc_my_pe_get() * osc_g.slice + off;
}

void *shmem_malloc(size_t size) {
  size_t off;
  if (osc_g.spmd) return osc_spmd_heap(OSC_SPMD_MALLOC, NULL, size, 0);
  return osc_heap_call_alloc(size, 0, 0, &off) == OSC_OK ? my_addr(off) : NULL;
}

void *shmem_align(size_t alignment, size_t size) {
  size_t off;
  if (alignment == 0 || (alignment & (alignment - 1)) != 0) return NULL;
  if (osc_g.spmd) return osc_spmd_heap(OSC_SPMD_ALIGN, NULL, size, alignment);
  return osc_heap_call_alloc(size, alignment, 0, &off) == OSC_OK ? my_addr(off)
                                                                : NULL;
}

void *shmem_calloc(size_t count, size_t size) {
  size_t off;
  if (size != 0 && count > ((size_t)-1) / size) return NULL;
  if (osc_g.spmd) return osc_spmd_heap(OSC_SPMD_CALLOC, NULL, count * size, 0);
  return osc_heap_call_alloc(count * size, 0, 1, &off) == OSC_OK ? my_addr(off)
                                                                 : NULL;
}

void shmem_free(void *ptr) {
  size_t off;
  if (osc_g.spmd) {
    osc_spmd_heap(OSC_SPMD_FREE, ptr, 0, 0);
    return;
  }
  if (ptr == NULL || osc_sym_offset(ptr, NULL, &off) != OSC_OK) return;
  osc_heap_call_free(off);
}

void *shmem_realloc(void *ptr, size_t size) {
  size_t off, noff;
  if (osc_g.spmd) return osc_spmd_heap(OSC_SPMD_REALLOC, ptr, size, 0);
  if (ptr == NULL) return shmem_malloc(size);
  if (size == 0) {
    shmem_free(ptr);
    return NULL;
  }
  if (osc_sym_offset(ptr, NULL, &off) != OSC_OK) return NULL;
  return osc_heap_call_realloc(off, size, &noff) == OSC_OK ? my_addr(noff)
                                                           : NULL;
}

void *shmem_ptr(const void *dest, int pe) { return osc_pe_ptr(dest, pe); }

void *shmem_team_ptr(shmem_team_t team, const void *dest, int pe) {
  const osc_team *t = osc_team_get(team);
  if (t == NULL || pe < 0 || pe >= t->npes) return NULL;
  return osc_pe_ptr(dest, t->start + pe * t->stride);
}

void *shmem_malloc_with_hints(size_t size, long hints) {
  (void)hints;
  return shmem_malloc(size);
}

/* names deprecated since OpenSHMEM 1.2 (reference src/api/deprecated/
 * shmalloc.c:25-48) */
void *shmalloc(size_t size) { return shmem_malloc(size); }
void shfree(void *ptr) { shmem_free(ptr); }
void *shrealloc(void *ptr, size_t size) { return shmem_realloc(ptr, size); }
void *shmemalign(size_t alignment, size_t size) {
  return shmem_align(alignment, size);
}

int shmem_addr_accessible(const void *addr, int pe) {
  return osc_pe_ptr(addr, pe) != NULL;
}

/* *out = nelems * es; OSC_ERR_ARG (recorded) if that overflows */
static int sized_bytes(size_t nelems, size_t es, size_t *out) {
  if (nelems > ((size_t)-1) / es) return osc_set_error(OSC_ERR_ARG, "size overflow");
  *out = nelems * es;
  return OSC_OK;
}

/* ---- RMA between emulated PEs ------------------------------------------------ */
/* The local side may be a symmetric (device) address or host memory; the
 * remote side is the symmetric object on PE pe (reference putget.h:403-600:
 * typed and sized forms are byte counts over the same copy). */

static int is_sym(const void *p) {
  const char *c = (const char *)p;
  return osc_g.initialized && c >= osc_g.base &&
         c < osc_g.base + (size_t)osc_g.npes * osc_g.slice;
}

/* device address of [sym + lo, sym + lo + bytes) on PE pe, range-checked */
static char *remote(const void *sym, ptrdiff_t lo, size_t bytes, int pe) {
  size_t off;
  if (!osc_g.initialized) {
    osc_set_error(OSC_ERR_NOTINIT, "not initialised");
    return NULL;
  }
  if (pe < 0 || pe >= osc_g.npes) {
    osc_set_error(OSC_ERR_ARG, "PE %d outside [0, %d)", pe, osc_g.npes);
    return NULL;
  }
  if (osc_sym_range((const char *)sym + lo, bytes, &off) != OSC_OK) return NULL;
  return osc_g.base + (size_t)pe * osc_g.slice + off - lo;
}

/* ---- communication contexts --------------------------------------------------
 * A context is a device stream (and, from shmem_team_create_ctx, a team
 * whose indices its PE numbers are).  SHMEM_CTX_DEFAULT stands for the
 * library stream.  Contexts not destroyed by the program are released by
 * osc_finalize (osc_ctx_release_all). */
#define OSC_CTX_MAGIC 0x6f736378u
typedef struct osc_ctx {
  uint32_t magic;
  void *stream;
  shmem_team_t team; /* NULL: PE numbers are global */
  long options;
  struct osc_ctx *next;
} osc_ctx;
static osc_ctx g_ctx_default = {OSC_CTX_MAGIC, NULL, NULL, 0, NULL};
shmem_ctx_t SHMEM_CTX_DEFAULT = &g_ctx_default;
static pthread_mutex_t g_ctx_mu = PTHREAD_MUTEX_INITIALIZER;
static osc_ctx *g_ctxs; /* created and not destroyed */

static const osc_ctx *ctx_get(shmem_ctx_t ctx) {
  const osc_ctx *c = (const osc_ctx *)ctx;
  if (c == NULL || c->magic != OSC_CTX_MAGIC) {
    osc_set_error(OSC_ERR_ARG, "invalid context");
    return NULL;
  }
  if (!osc_g.initialized) {
    osc_set_error(OSC_ERR_NOTINIT, "not initialised");
    return NULL;
  }
  return c;
}

/* The context's stream and the global PE of its PE number pe; -1 (error
 * recorded) if either is invalid. */
int osc_ctx_target(void *ctx, int pe, void **stream) {
  const osc_ctx *c = ctx_get(ctx);
  const osc_team *t;
  if (c == NULL) return -1;
  *stream = c == &g_ctx_default ? osc_g.stream : c->stream;
  if (c->team == NULL) {
    if (pe < 0 || pe >= osc_g.npes)
      return osc_set_error(OSC_ERR_ARG, "PE %d outside 0..%d", pe, osc_g.npes - 1), -1;
    return pe;
  }
  t = osc_team_get(c->team);
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "the context's team is gone"), -1;
  if (pe < 0 || pe >= t->npes)
    return osc_set_error(OSC_ERR_ARG, "PE %d outside the context's team of %d", pe,
                         t->npes), -1;
  return t->start + pe * t->stride;
}

static int ctx_new(shmem_team_t team, long options, shmem_ctx_t *out) {
  osc_ctx *c;
  int rc;
  if (out == NULL) return osc_set_error(OSC_ERR_ARG, "NULL context address"), -1;
  *out = SHMEM_CTX_INVALID;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised"), -1;
  c = (osc_ctx *)calloc(1, sizeof(*c));
  if (c == NULL) return osc_set_error(OSC_ERR_NOMEM, "context"), -1;
  rc = oscd_stream_create(&c->stream);
  if (rc != OSC_OK) {
    free(c);
    return osc_hip_error(rc, "context stream"), -1;
  }
  c->magic = OSC_CTX_MAGIC;
  c->team = team;
  c->options = options;
  pthread_mutex_lock(&g_ctx_mu);
  c->next = g_ctxs;
  g_ctxs = c;
  pthread_mutex_unlock(&g_ctx_mu);
  *out = c;
  return 0;
}

int shmem_ctx_create(long options, shmem_ctx_t *ctx) {
  return ctx_new(NULL, options, ctx);
}

int shmem_team_create_ctx(shmem_team_t team, long options, shmem_ctx_t *ctx) {
  if (ctx != NULL) *ctx = SHMEM_CTX_INVALID;
  if (osc_team_get(team) == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team"), -1;
  return ctx_new(team, options, ctx);
}

static void ctx_free(osc_ctx *c) {
  oscd_stream_sync(c->stream);
  oscd_stream_destroy(c->stream);
  c->magic = 0;
  free(c);
}

void shmem_ctx_destroy(shmem_ctx_t ctx) {
  osc_ctx *c = (osc_ctx *)ctx, **pp;
  if (ctx_get(ctx) == NULL) return;
  if (c == &g_ctx_default) {
    osc_set_error(OSC_ERR_ARG, "the default context cannot be destroyed");
    return;
  }
  pthread_mutex_lock(&g_ctx_mu);
  for (pp = &g_ctxs; *pp != NULL; pp = &(*pp)->next)
    if (*pp == c) {
      *pp = c->next;
      break;
    }
  pthread_mutex_unlock(&g_ctx_mu);
  ctx_free(c);
}

void osc_ctx_release_all(void) {
  pthread_mutex_lock(&g_ctx_mu);
  while (g_ctxs != NULL) {
    osc_ctx *c = g_ctxs;
    g_ctxs = c->next;
    ctx_free(c);
  }
  pthread_mutex_unlock(&g_ctx_mu);
}

int shmem_ctx_get_team(shmem_ctx_t ctx, shmem_team_t *team) {
  const osc_ctx *c;
  if (team == NULL) return osc_set_error(OSC_ERR_ARG, "NULL team address"), -1;
  *team = SHMEM_TEAM_INVALID;
  c = ctx_get(ctx);
  if (c == NULL) return -1;
  *team = c->team != NULL ? c->team : SHMEM_TEAM_WORLD;
  return 0;
}

/* operations of one context are ordered by its stream already */
void shmem_ctx_fence(shmem_ctx_t ctx) { (void)ctx_get(ctx); }

void shmem_ctx_quiet(shmem_ctx_t ctx) {
  const osc_ctx *c = ctx_get(ctx);
  int rc;
  if (c == NULL) return;
  rc = oscd_stream_wait(c == &g_ctx_default ? osc_g.stream : c->stream,
                        osc_g.wait_spin_us);
  if (rc != OSC_OK) osc_hip_error(rc, "context quiet");
}

/* A copy on stream st; the host side is copied once the stream's earlier
 * work is done. */
static void xfer(void *st, void *dst, const void *src, size_t nbytes, int blocking) {
  size_t off;
  int rc;
  if (nbytes == 0) return;
  /* a symmetric local side must lie inside one PE's heap as a whole, or
   * the copy would run into the next PE's slice (or past the arena) */
  if ((is_sym(dst) && osc_sym_range(dst, nbytes, &off) != OSC_OK) ||
      (is_sym(src) && osc_sym_range(src, nbytes, &off) != OSC_OK))
    return;
  if (is_sym(dst) && is_sym(src)) {
    rc = oscd_copy(dst, src, nbytes, st);
    if (rc == OSC_OK && blocking) rc = oscd_stream_wait(st, osc_g.wait_spin_us);
  } else if (is_sym(dst)) {
    rc = oscd_stream_wait(st, osc_g.wait_spin_us);
    if (rc == OSC_OK) rc = oscd_memcpy_h2d(dst, src, nbytes);
  } else {
    rc = oscd_stream_wait(st, osc_g.wait_spin_us);
    if (rc == OSC_OK) rc = oscd_memcpy_d2h(dst, src, nbytes);
  }
  if (rc != OSC_OK) osc_hip_error(rc, "RMA copy");
}

/* nelems elements of es bytes (a byte count overflowing size_t is an
 * error, not a shorter copy) */
static void rma_mem(shmem_ctx_t ctx, void *dest, const void *source,
                    size_t nelems, size_t es, int pe, int is_put, int blocking) {
  void *st = NULL;
  size_t nbytes = 0;
  int g;
  if (sized_bytes(nelems, es, &nbytes) != OSC_OK) return;
  g = osc_ctx_target(ctx, pe, &st);
  if (g < 0) return;
  if (is_put) {
    void *d = remote(dest, 0, nbytes, g);
    if (d != NULL) xfer(st, d, source, nbytes, blocking);
  } else {
    const void *s = remote(source, 0, nbytes, g);
    if (s != NULL) xfer(st, dest, s, nbytes, blocking);
  }
}

void shmem_ctx_putmem(shmem_ctx_t ctx, void *dest, const void *source,
                      size_t nbytes, int pe) {
  rma_mem(ctx, dest, source, nbytes, 1, pe, 1, 1);
}
void shmem_ctx_putmem_nbi(shmem_ctx_t ctx, void *dest, const void *source,
                          size_t nbytes, int pe) {
  rma_mem(ctx, dest, source, nbytes, 1, pe, 1, 0);
}
void shmem_ctx_getmem(shmem_ctx_t ctx, void *dest, const void *source,
                      size_t nbytes, int pe) {
  rma_mem(ctx, dest, source, nbytes, 1, pe, 0, 1);
}
void shmem_ctx_getmem_nbi(shmem_ctx_t ctx, void *dest, const void *source,
                          size_t nbytes, int pe) {
  rma_mem(ctx, dest, source, nbytes, 1, pe, 0, 0);
}
void shmem_putmem(void *dest, const void *source, size_t nbytes, int pe) {
  rma_mem(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, pe, 1, 1);
}
void shmem_putmem_nbi(void *dest, const void *source, size_t nbytes, int pe) {
  rma_mem(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, pe, 1, 0);
}
void shmem_getmem(void *dest, const void *source, size_t nbytes, int pe) {
  rma_mem(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, pe, 0, 1);
}
void shmem_getmem_nbi(void *dest, const void *source, size_t nbytes, int pe) {
  rma_mem(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, pe, 0, 0);
}

/* Strided put (is_put) or get of n elements of es bytes; strides in
 * elements, any sign.  Sequential semantics (putget.h:130-186: one put per
 * element in index order): with a zero stride on the written side only
 * the last element survives, so only it is moved.  A host-side array is
 * packed into (or unpacked from) a contiguous device staging buffer. */
static void strided(shmem_ctx_t ctx, void *dest, const void *src, size_t es,
                    ptrdiff_t tst, ptrdiff_t sst, size_t n, int pe, int is_put) {
  const ptrdiff_t rst = is_put ? tst : sst; /* remote stride */
  const ptrdiff_t lst = is_put ? sst : tst; /* local stride  */
  void *rsym = is_put ? dest : (void *)src;
  void *lp = is_put ? (void *)src : dest;
  ptrdiff_t lo, hi;
  char *r, *stage = NULL;
  void *st = NULL;
  int rc = OSC_OK;
  if (n == 0) return;
  pe = osc_ctx_target(ctx, pe, &st);
  if (pe < 0) return;
  {
    /* every byte offset below, (n-1)*|stride|*es + es, must fit ptrdiff_t */
    const size_t at = (size_t)(rst < 0 ? -rst : rst), al = (size_t)(lst < 0 ? -lst : lst);
    const size_t amax = (at > al ? at : al) > 0 ? (at > al ? at : al) : 1;
    if (es > (size_t)PTRDIFF_MAX / 2 || n - 1 > ((size_t)PTRDIFF_MAX / 2) / es / amax) {
      osc_set_error(OSC_ERR_ARG, "strided RMA of %zu elements overflows", n);
      return;
    }
  }
  if (tst == 0) { /* every element lands on one target: the last wins */
    if (is_put)
      lp = (char *)lp + (ptrdiff_t)(n - 1) * lst * (ptrdiff_t)es;
    else
      rsym = (char *)rsym + (ptrdiff_t)(n - 1) * rst * (ptrdiff_t)es;
    n = 1;
  }
  lo = rst < 0 ? (ptrdiff_t)(n - 1) * rst * (ptrdiff_t)es : 0;
  hi = (rst > 0 ? (ptrdiff_t)(n - 1) * rst * (ptrdiff_t)es : 0) + (ptrdiff_t)es;
  r = remote(rsym, lo, (size_t)(hi - lo), pe);
  if (r == NULL) return;
  if (is_sym(lp)) {
    /* the local side's extent, as the remote one's above */
    const ptrdiff_t llo = lst < 0 ? (ptrdiff_t)(n - 1) * lst * (ptrdiff_t)es : 0;
    const ptrdiff_t lhi =
        (lst > 0 ? (ptrdiff_t)(n - 1) * lst * (ptrdiff_t)es : 0) + (ptrdiff_t)es;
    size_t loff;
    if (osc_sym_range((char *)lp + llo, (size_t)(lhi - llo), &loff) != OSC_OK)
      return;
    rc = is_put ? oscd_copy_strided(r, lp, es, (uint64_t)(rst * (ptrdiff_t)es),
                                    (uint64_t)(lst * (ptrdiff_t)es), n,
                                    st)
                : oscd_copy_strided(lp, r, es, (uint64_t)(lst * (ptrdiff_t)es),
                                    (uint64_t)(rst * (ptrdiff_t)es), n,
                                    st);
    if (rc == OSC_OK) rc = oscd_stream_wait(st, osc_g.wait_spin_us);
    if (rc != OSC_OK) osc_hip_error(rc, "strided RMA");
    return;
  }
  {
    char *host = (char *)malloc(n * es);
    size_t t;
    if (host == NULL) {
      osc_set_error(OSC_ERR_NOMEM, "strided RMA staging");
      return;
    }
    rc = oscd_malloc((void **)&stage, n * es);
    if (rc == OSC_OK && is_put) {
      for (t = 0; t < n; ++t)
        memcpy(host + t * es, (char *)lp + (ptrdiff_t)t * lst * (ptrdiff_t)es, es);
      rc = oscd_stream_wait(st, osc_g.wait_spin_us);
      if (rc == OSC_OK) rc = oscd_memcpy_h2d(stage, host, n * es);
      if (rc == OSC_OK)
        rc = oscd_copy_strided(r, stage, es, (uint64_t)(rst * (ptrdiff_t)es), es,
                               n, st);
      if (rc == OSC_OK) rc = oscd_stream_wait(st, osc_g.wait_spin_us);
    } else if (rc == OSC_OK) {
      rc = oscd_copy_strided(stage, r, es, es, (uint64_t)(rst * (ptrdiff_t)es), n,
                             st);
      if (rc == OSC_OK) rc = oscd_stream_wait(st, osc_g.wait_spin_us);
      if (rc == OSC_OK) rc = oscd_memcpy_d2h(host, stage, n * es);
      if (rc == OSC_OK)
        for (t = 0; t < n; ++t)
          memcpy((char *)lp + (ptrdiff_t)t * lst * (ptrdiff_t)es, host + t * es, es);
    }
    if (stage != NULL) oscd_free(stage);
    free(host);
    if (rc != OSC_OK) osc_hip_error(rc, "strided RMA");
  }
}

/* put then signal update on the same stream, so the data is in place
 * before the signal changes (put_signal.h:28-50, SHMEM_SIGNAL_SET / ADD) */
static void put_signal(shmem_ctx_t ctx, void *dest, const void *src,
                       size_t nelems, size_t es, uint64_t *sig_addr,
                       uint64_t signal, int sig_op, int pe, int blocking) {
  void *st = NULL;
  char *d;
  uint64_t *s;
  size_t nbytes = 0;
  int rc;
  if (sized_bytes(nelems, es, &nbytes) != OSC_OK) return;
  pe = osc_ctx_target(ctx, pe, &st);
  if (pe < 0) return;
  d = remote(dest, 0, nbytes, pe);
  if (d == NULL) return;
  s = (uint64_t *)remote(sig_addr, 0, sizeof(uint64_t), pe);
  if (s == NULL) return;
  if (sig_op != SHMEM_SIGNAL_SET && sig_op != SHMEM_SIGNAL_ADD) {
    osc_set_error(OSC_ERR_ARG, "signal op %d", sig_op);
    return;
  }
  xfer(st, d, src, nbytes, 0);
  rc = oscd_signal(s, signal, sig_op, st);
  if (rc == OSC_OK && blocking) rc = oscd_stream_wait(st, osc_g.wait_spin_us);
  if (rc != OSC_OK) osc_hip_error(rc, "signal update");
}

void shmem_putmem_signal(void *dest, const void *source, size_t nbytes,
                         uint64_t *sig_addr, uint64_t signal, int sig_op,
                         int pe) {
  put_signal(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, sig_addr, signal,
             sig_op, pe, 1);
}
void shmem_putmem_signal_nbi(void *dest, const void *source, size_t nbytes,
                             uint64_t *sig_addr, uint64_t signal, int sig_op,
                             int pe) {
  put_signal(SHMEM_CTX_DEFAULT, dest, source, nbytes, 1, sig_addr, signal,
             sig_op, pe, 0);
}
void shmem_ctx_putmem_signal(shmem_ctx_t ctx, void *dest, const void *source,
                             size_t nbytes, uint64_t *sig_addr,
                             uint64_t signal, int sig_op, int pe) {
  put_signal(ctx, dest, source, nbytes, 1, sig_addr, signal, sig_op, pe, 1);
}
void shmem_ctx_putmem_signal_nbi(shmem_ctx_t ctx, void *dest,
                                 const void *source, size_t nbytes,
                                 uint64_t *sig_addr, uint64_t signal,
                                 int sig_op, int pe) {
  put_signal(ctx, dest, source, nbytes, 1, sig_addr, signal, sig_op, pe, 0);
}

static int read_signal(const uint64_t *sig_addr, uint64_t *v) {
  size_t off;
  int rc;
  *v = 0;
  if (!is_sym(sig_addr) || ((uintptr_t)sig_addr & 7u) != 0 ||
      osc_sym_range(sig_addr, sizeof(uint64_t), &off) != OSC_OK)
    return osc_set_error(OSC_ERR_PTR, "signal address %p",
                         (const void *)sig_addr);
  rc = oscd_memcpy_d2h(v, sig_addr, sizeof(*v));
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "signal read");
}

uint64_t shmem_signal_fetch(const uint64_t *sig_addr) {
  uint64_t v;
  read_signal(sig_addr, &v);
  return v;
}

static int cmp_ok(uint64_t v, int cmp, uint64_t x) {
  switch (cmp) {
    case SHMEM_CMP_EQ: return v == x;
    case SHMEM_CMP_NE: return v != x;
    case SHMEM_CMP_GT: return v > x;
    case SHMEM_CMP_LE: return v <= x;
    case SHMEM_CMP_LT: return v < x;
    default: return v >= x; /* SHMEM_CMP_GE */
  }
}

/* Polls the signal word; work already enqueued by this process is let
 * finish first.  Like the reference, it blocks until another thread's
 * put_signal satisfies the comparison if it is not satisfied yet; a bad
 * address or comparison returns at once with the error recorded. */
uint64_t shmem_signal_wait_until(uint64_t *sig_addr, int cmp,
                                 uint64_t cmp_value) {
  uint64_t v;
  if (cmp < SHMEM_CMP_EQ || cmp > SHMEM_CMP_GE) {
    osc_set_error(OSC_ERR_ARG, "comparison %d", cmp);
    return 0;
  }
  oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  for (;;) {
    if (read_signal(sig_addr, &v) != OSC_OK || cmp_ok(v, cmp, cmp_value))
      return v;
    sched_yield();
  }
}

/* Every routine is its context form on SHMEM_CTX_DEFAULT. */
#define OSC_DEF_RMA(_type, _typename)                                          \
  void shmem_ctx_##_typename##_put(shmem_ctx_t ctx, _type *dest,               \
                                   const _type *src, size_t nelems, int pe) {  \
    rma_mem(ctx, dest, src, nelems, sizeof(_type), pe, 1, 1);                  \
  }                                                                            \
  void shmem_ctx_##_typename##_put_nbi(shmem_ctx_t ctx, _type *dest,           \
                                       const _type *src, size_t nelems,        \
                                       int pe) {                               \
    rma_mem(ctx, dest, src, nelems, sizeof(_type), pe, 1, 0);                  \
  }                                                                            \
  void shmem_ctx_##_typename##_get(shmem_ctx_t ctx, _type *dest,               \
                                   const _type *src, size_t nelems, int pe) {  \
    rma_mem(ctx, dest, src, nelems, sizeof(_type), pe, 0, 1);                  \
  }                                                                            \
  void shmem_ctx_##_typename##_get_nbi(shmem_ctx_t ctx, _type *dest,           \
                                       const _type *src, size_t nelems,        \
                                       int pe) {                               \
    rma_mem(ctx, dest, src, nelems, sizeof(_type), pe, 0, 0);                  \
  }                                                                            \
  void shmem_ctx_##_typename##_p(shmem_ctx_t ctx, _type *dest, _type value,    \
                                 int pe) {                                     \
    rma_mem(ctx, dest, &value, 1, sizeof(_type), pe, 1, 1);                    \
  }                                                                            \
  _type shmem_ctx_##_typename##_g(shmem_ctx_t ctx, const _type *src, int pe) { \
    _type v;                                                                   \
    memset(&v, 0, sizeof(v));                                                  \
    rma_mem(ctx, &v, src, 1, sizeof(_type), pe, 0, 1);                         \
    return v;                                                                  \
  }                                                                            \
  void shmem_ctx_##_typename##_iput(shmem_ctx_t ctx, _type *dest,              \
                                    const _type *src, ptrdiff_t tst,           \
                                    ptrdiff_t sst, size_t nelems, int pe) {    \
    strided(ctx, dest, src, sizeof(_type), tst, sst, nelems, pe, 1);           \
  }                                                                            \
  void shmem_ctx_##_typename##_iget(shmem_ctx_t ctx, _type *dest,              \
                                    const _type *src, ptrdiff_t tst,           \
                                    ptrdiff_t sst, size_t nelems, int pe) {    \
    strided(ctx, dest, src, sizeof(_type), tst, sst, nelems, pe, 0);           \
  }                                                                            \
  void shmem_ctx_##_typename##_put_signal(                                     \
      shmem_ctx_t ctx, _type *dest, const _type *src, size_t nelems,           \
      uint64_t *sig_addr, uint64_t signal, int sig_op, int pe) {               \
    put_signal(ctx, dest, src, nelems, sizeof(_type), sig_addr, signal,        \
               sig_op, pe, 1);                                                 \
  }                                                                            \
  void shmem_ctx_##_typename##_put_signal_nbi(                                 \
      shmem_ctx_t ctx, _type *dest, const _type *src, size_t nelems,           \
      uint64_t *sig_addr, uint64_t signal, int sig_op, int pe) {               \
    put_signal(ctx, dest, src, nelems, sizeof(_type), sig_addr, signal,        \
               sig_op, pe, 0);                                                 \
  }                                                                            \
  void shmem_##_typename##_put(_type *dest, const _type *src, size_t nelems,   \
                               int pe) {                                       \
    shmem_ctx_##_typename##_put(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);     \
  }                                                                            \
  void shmem_##_typename##_put_nbi(_type *dest, const _type *src,              \
                                   size_t nelems, int pe) {                    \
    shmem_ctx_##_typename##_put_nbi(SHMEM_CTX_DEFAULT, dest, src, nelems, pe); \
  }                                                                            \
  void shmem_##_typename##_get(_type *dest, const _type *src, size_t nelems,   \
                               int pe) {                                       \
    shmem_ctx_##_typename##_get(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);     \
  }                                                                            \
  void shmem_##_typename##_get_nbi(_type *dest, const _type *src,              \
                                   size_t nelems, int pe) {                    \
    shmem_ctx_##_typename##_get_nbi(SHMEM_CTX_DEFAULT, dest, src, nelems, pe); \
  }                                                                            \
  void shmem_##_typename##_p(_type *dest, _type value, int pe) {               \
    shmem_ctx_##_typename##_p(SHMEM_CTX_DEFAULT, dest, value, pe);             \
  }                                                                            \
  _type shmem_##_typename##_g(const _type *src, int pe) {                      \
    return shmem_ctx_##_typename##_g(SHMEM_CTX_DEFAULT, src, pe);              \
  }                                                                            \
  void shmem_##_typename##_iput(_type *dest, const _type *src, ptrdiff_t tst,  \
                                ptrdiff_t sst, size_t nelems, int pe) {        \
    strided(SHMEM_CTX_DEFAULT, dest, src, sizeof(_type), tst, sst, nelems, pe, \
            1);                                                                \
  }                                                                            \
  void shmem_##_typename##_iget(_type *dest, const _type *src, ptrdiff_t tst,  \
                                ptrdiff_t sst, size_t nelems, int pe) {        \
    strided(SHMEM_CTX_DEFAULT, dest, src, sizeof(_type), tst, sst, nelems, pe, \
            0);                                                                \
  }                                                                            \
  void shmem_##_typename##_put_signal(_type *dest, const _type *src,           \
                                      size_t nelems, uint64_t *sig_addr,       \
                                      uint64_t signal, int sig_op, int pe) {   \
    put_signal(SHMEM_CTX_DEFAULT, dest, src, nelems, sizeof(_type), sig_addr,  \
               signal, sig_op, pe, 1);                                         \
  }                                                                            \
  void shmem_##_typename##_put_signal_nbi(_type *dest, const _type *src,       \
                                          size_t nelems, uint64_t *sig_addr,   \
                                          uint64_t signal, int sig_op,         \
                                          int pe) {                            \
    put_signal(SHMEM_CTX_DEFAULT, dest, src, nelems, sizeof(_type), sig_addr,  \
               signal, sig_op, pe, 0);                                         \
  }
OSC_RMA_TYPES(OSC_DEF_RMA)

#define OSC_DEF_SIZED_RMA(_size)                                               \
  void shmem_ctx_put##_size(shmem_ctx_t ctx, void *dest, const void *src,      \
                            size_t nelems, int pe) {                           \
    rma_mem(ctx, dest, src, nelems, (_size) / 8, pe, 1, 1);                    \
  }                                                                            \
  void shmem_ctx_put##_size##_nbi(shmem_ctx_t ctx, void *dest,                 \
                                  const void *src, size_t nelems, int pe) {    \
    rma_mem(ctx, dest, src, nelems, (_size) / 8, pe, 1, 0);                    \
  }                                                                            \
  void shmem_ctx_get##_size(shmem_ctx_t ctx, void *dest, const void *src,      \
                            size_t nelems, int pe) {                           \
    rma_mem(ctx, dest, src, nelems, (_size) / 8, pe, 0, 1);                    \
  }                                                                            \
  void shmem_ctx_get##_size##_nbi(shmem_ctx_t ctx, void *dest,                 \
                                  const void *src, size_t nelems, int pe) {    \
    rma_mem(ctx, dest, src, nelems, (_size) / 8, pe, 0, 0);                    \
  }                                                                            \
  void shmem_ctx_iput##_size(shmem_ctx_t ctx, void *dest, const void *src,     \
                             ptrdiff_t tst, ptrdiff_t sst, size_t nelems,      \
                             int pe) {                                         \
    strided(ctx, dest, src, (_size) / 8, tst, sst, nelems, pe, 1);             \
  }                                                                            \
  void shmem_ctx_iget##_size(shmem_ctx_t ctx, void *dest, const void *src,     \
                             ptrdiff_t tst, ptrdiff_t sst, size_t nelems,      \
                             int pe) {                                         \
    strided(ctx, dest, src, (_size) / 8, tst, sst, nelems, pe, 0);             \
  }                                                                            \
  void shmem_ctx_put##_size##_signal(shmem_ctx_t ctx, void *dest,              \
                                     const void *src, size_t nelems,           \
                                     uint64_t *sig_addr, uint64_t signal,      \
                                     int sig_op, int pe) {                     \
    put_signal(ctx, dest, src, nelems, (_size) / 8, sig_addr, signal,          \
               sig_op, pe, 1);                                                 \
  }                                                                            \
  void shmem_ctx_put##_size##_signal_nbi(shmem_ctx_t ctx, void *dest,          \
                                         const void *src, size_t nelems,       \
                                         uint64_t *sig_addr, uint64_t signal,  \
                                         int sig_op, int pe) {                 \
    put_signal(ctx, dest, src, nelems, (_size) / 8, sig_addr, signal,          \
               sig_op, pe, 0);                                                 \
  }                                                                            \
  void shmem_put##_size(void *dest, const void *src, size_t nelems, int pe) {  \
    shmem_ctx_put##_size(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);            \
  }                                                                            \
  void shmem_put##_size##_nbi(void *dest, const void *src, size_t nelems,      \
                              int pe) {                                        \
    shmem_ctx_put##_size##_nbi(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);      \
  }                                                                            \
  void shmem_get##_size(void *dest, const void *src, size_t nelems, int pe) {  \
    shmem_ctx_get##_size(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);            \
  }                                                                            \
  void shmem_get##_size##_nbi(void *dest, const void *src, size_t nelems,      \
                              int pe) {                                        \
    shmem_ctx_get##_size##_nbi(SHMEM_CTX_DEFAULT, dest, src, nelems, pe);      \
  }                                                                            \
  void shmem_iput##_size(void *dest, const void *src, ptrdiff_t tst,           \
                         ptrdiff_t sst, size_t nelems, int pe) {               \
    shmem_ctx_iput##_size(SHMEM_CTX_DEFAULT, dest, src, tst, sst, nelems, pe); \
  }                                                                            \
  void shmem_iget##_size(void *dest, const void *src, ptrdiff_t tst,           \
                         ptrdiff_t sst, size_t nelems, int pe) {               \
    shmem_ctx_iget##_size(SHMEM_CTX_DEFAULT, dest, src, tst, sst, nelems, pe); \
  }                                                                            \
  void shmem_put##_size##_signal(void *dest, const void *src, size_t nelems,   \
                                 uint64_t *sig_addr, uint64_t signal,          \
                                 int sig_op, int pe) {                         \
    shmem_ctx_put##_size##_signal(SHMEM_CTX_DEFAULT, dest, src, nelems,        \
                                  sig_addr, signal, sig_op, pe);               \
  }                                                                            \
  void shmem_put##_size##_signal_nbi(void *dest, const void *src,              \
                                     size_t nelems, uint64_t *sig_addr,        \
                                     uint64_t signal, int sig_op, int pe) {    \
    shmem_ctx_put##_size##_signal_nbi(SHMEM_CTX_DEFAULT, dest, src, nelems,    \
                                      sig_addr, signal, sig_op, pe);           \
  }
OSC_SIZED_RMA(OSC_DEF_SIZED_RMA)

/* ---- collectives ---------------------------------------------------------------- */

static int finish(int rc) {
  if (rc != OSC_OK) return rc;
  rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  return rc == OSC_OK ? 0 : osc_hip_error(rc, "stream wait");
}

static int blocking_reduce(shmem_team_t team, int op, int dtype, void *dest,
                           const void *src, size_t n) {
  if (osc_g.spmd) /* every member arrives, dtype < 0 included */
    return osc_spmd_reduce(team, op, dtype, dest, src, n, -1);
  if (dtype < 0)
    return osc_set_error(OSC_ERR_UNSUPPORTED, "long double reduction (no 80-bit type on gfx950)");
  return finish(osc_reduce(team, op, dtype, dest, src, n, NULL));
}

/* blocking data movement, team-call or SPMD */
static int do_fcollect(void *team, void *dest, const void *src, size_t nb) {
  if (osc_g.spmd)
    return osc_spmd_move(OSC_SPMD_FCOLLECT, team, dest, src, nb, 0, 0);
  return finish(osc_fcollect(team, dest, src, nb, NULL));
}
static int do_alltoall(void *team, void *dest, const void *src, size_t nb) {
  if (osc_g.spmd)
    return osc_spmd_move(OSC_SPMD_ALLTOALL, team, dest, src, nb, 0, 0);
  return finish(osc_alltoall(team, dest, src, nb, NULL));
}
static int do_broadcast(void *team, void *dest, const void *src, size_t nb,
                        int root, int skip_root) {
  if (osc_g.spmd)
    return osc_spmd_move(OSC_SPMD_BROADCAST, team, dest, src, nb, root,
                         skip_root);
  return finish(osc_broadcast_impl(team, dest, src, nb, root, skip_root, NULL));
}
static int do_alltoalls(void *team, void *dest, const void *src, ptrdiff_t dst,
                        ptrdiff_t sst, size_t es, size_t n) {
  if (osc_g.spmd) return osc_spmd_alltoalls(team, dest, src, dst, sst, es, n);
  return finish(osc_alltoalls(team, dest, src, dst, sst, es, n, NULL));
}

#define OSC_DEF_REDUCE(_type, _typename, _op, _OPC)                            \
  int shmem_##_typename##_##_op##_reduce(shmem_team_t team, _type *dest,       \
                                         const _type *source, size_t nreduce) { \
    return blocking_reduce(team, _OPC, OSC_DT_##_typename, dest, source,       \
                           nreduce);                                           \
  }
#define OSC_DEF_AND(_t, _n) OSC_DEF_REDUCE(_t, _n, and, OSC_OP_AND)
#define OSC_DEF_OR(_t, _n) OSC_DEF_REDUCE(_t, _n, or, OSC_OP_OR)
#define OSC_DEF_XOR(_t, _n) OSC_DEF_REDUCE(_t, _n, xor, OSC_OP_XOR)
#define OSC_DEF_MIN(_t, _n) OSC_DEF_REDUCE(_t, _n, min, OSC_OP_MIN)
#define OSC_DEF_MAX(_t, _n) OSC_DEF_REDUCE(_t, _n, max, OSC_OP_MAX)
#define OSC_DEF_SUM(_t, _n) OSC_DEF_REDUCE(_t, _n, sum, OSC_OP_SUM)
#define OSC_DEF_PROD(_t, _n) OSC_DEF_REDUCE(_t, _n, prod, OSC_OP_PROD)
OSC_BITWISE_TYPES(OSC_DEF_AND)
OSC_BITWISE_TYPES(OSC_DEF_OR)
OSC_BITWISE_TYPES(OSC_DEF_XOR)
OSC_MINMAX_TYPES(OSC_DEF_MIN)
OSC_MINMAX_TYPES(OSC_DEF_MAX)
OSC_ARITH_TYPES(OSC_DEF_SUM)
OSC_ARITH_TYPES(OSC_DEF_PROD)

#define OSC_DEF_MOVE(_type, _typename)                                         \
  int shmem_##_typename##_fcollect(shmem_team_t team, _type *dest,             \
                                   const _type *source, size_t nelems) {       \
    size_t nb = 0;                                                             \
    if (sized_bytes(nelems, sizeof(_type), &nb) != OSC_OK) return OSC_ERR_ARG; \
    return do_fcollect(team, dest, source, nb);                                \
  }                                                                            \
  int shmem_##_typename##_alltoall(shmem_team_t team, _type *dest,             \
                                   const _type *source, size_t nelems) {       \
    size_t nb = 0;                                                             \
    if (sized_bytes(nelems, sizeof(_type), &nb) != OSC_OK) return OSC_ERR_ARG; \
    return do_alltoall(team, dest, source, nb);                                \
  }                                                                            \
  int shmem_##_typename##_broadcast(shmem_team_t team, _type *dest,            \
                                    const _type *source, size_t nelems,        \
                                    int PE_root) {                             \
    size_t nb = 0;                                                             \
    if (sized_bytes(nelems, sizeof(_type), &nb) != OSC_OK) return OSC_ERR_ARG; \
    return do_broadcast(team, dest, source, nb, PE_root, 0);                   \
  }
OSC_RMA_TYPES(OSC_DEF_MOVE)

int shmem_fcollectmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes) {
  return do_fcollect(team, dest, source, nbytes);
}
int shmem_alltoallmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes) {
  return do_alltoall(team, dest, source, nbytes);
}
int shmem_broadcastmem(shmem_team_t team, void *dest, const void *source,
                       size_t nbytes, int PE_root) {
  return do_broadcast(team, dest, source, nbytes, PE_root, 0);
}

/* collect: in team-call mode nelems is every member's count; the
 * variable-size case is osc_collect() with one size per member. */
static int uniform_collect(shmem_team_t team, void *dest, const void *source,
                           size_t nbytes) {
  const osc_team *t = osc_team_get(team);
  size_t *sizes;
  int i, rc;
  /* SPMD: nbytes is this member's own contribution (the variable case) */
  if (osc_g.spmd) return osc_spmd_collect(team, dest, source, nbytes);
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  sizes = (size_t *)malloc(sizeof(size_t) * (size_t)t->npes);
  if (sizes == NULL) return osc_set_error(OSC_ERR_NOMEM, "collect sizes");
  for (i = 0; i < t->npes; ++i) sizes[i] = nbytes;
  rc = osc_collect(team, dest, source, sizes, NULL);
  free(sizes);
  return finish(rc);
}

#define OSC_DEF_MOVE2(_type, _typename)                                        \
  int shmem_##_typename##_collect(shmem_team_t team, _type *dest,              \
                                  const _type *source, size_t nelems) {        \
    if (nelems > ((size_t)-1) / sizeof(_type))                                 \
      return osc_set_error(OSC_ERR_ARG, "size overflow");                      \
    return uniform_collect(team, dest, source, nelems * sizeof(_type));        \
  }                                                                            \
  int shmem_##_typename##_alltoalls(shmem_team_t team, _type *dest,            \
                                    const _type *source, ptrdiff_t dst,        \
                                    ptrdiff_t sst, size_t nelems) {            \
    return do_alltoalls(team, dest, source, dst, sst, sizeof(_type), nelems);  \
  }
OSC_RMA_TYPES(OSC_DEF_MOVE2)

int shmem_collectmem(shmem_team_t team, void *dest, const void *source,
                     size_t nelems) {
  return uniform_collect(team, dest, source, nelems);
}
/* one element of nelems bytes per member pair (alltoalls.c:391-428) */
int shmem_alltoallsmem(shmem_team_t team, void *dest, const void *source,
                       ptrdiff_t dst, ptrdiff_t sst, size_t nelems) {
  return do_alltoalls(team, dest, source, dst, sst, nelems, 1);
}

/* ---- active-set forms (deprecated in 1.5) ------------------------------------- */
/* The shim's legacy entry points (shcoll-shim.c:238-266 alltoall,
 * :373-420 alltoalls, :525-566 collect, :672-714 fcollect, :824-866
 * broadcast, :867-1014 to_all, :1442-1477 sync/barrier) map the active set
 * to a transient team and run the same kernels.  They return void. */

static void legacy_reduce(int op, int dtype, void *target, const void *source,
                          int nreduce, int PE_start, int logPE_stride,
                          int PE_size) {
  osc_team t;
  if (osc_g.spmd) {
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK) return;
    osc_spmd_reduce(&t, op, dtype, target, source,
                    nreduce < 0 ? (size_t)-1 : (size_t)nreduce,
                    dtype < 0 ? -1 : osc_g.fpmode_to_all[op][dtype]);
    return;
  }
  if (dtype < 0) {
    osc_set_error(OSC_ERR_UNSUPPORTED, "long double reduction (no 80-bit type on gfx950)");
    return;
  }
  if (nreduce < 0) {
    osc_set_error(OSC_ERR_ARG, "nreduce %d", nreduce);
    return;
  }
  if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK) return;
  finish(osc_reduce_mode(&t, op, dtype, target, source, (size_t)nreduce,
                         osc_g.fpmode_to_all[op][dtype], NULL));
}

#define OSC_DEF_TO_ALL(_type, _typename, _op, _OPC)                            \
  void shmem_##_typename##_##_op##_to_all(                                     \
      _type *target, const _type *source, int nreduce, int PE_start,           \
      int logPE_stride, int PE_size, _type *pWrk, long *pSync) {               \
    (void)pWrk;                                                                \
    (void)pSync;                                                               \
    legacy_reduce(_OPC, OSC_DT_##_typename, target, source, nreduce,           \
                  PE_start, logPE_stride, PE_size);                            \
  }
#define OSC_DEF_AND_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, and, OSC_OP_AND)
#define OSC_DEF_OR_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, or, OSC_OP_OR)
#define OSC_DEF_XOR_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, xor, OSC_OP_XOR)
#define OSC_DEF_MIN_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, min, OSC_OP_MIN)
#define OSC_DEF_MAX_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, max, OSC_OP_MAX)
#define OSC_DEF_SUM_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, sum, OSC_OP_SUM)
#define OSC_DEF_PROD_TO_ALL(_t, _n) OSC_DEF_TO_ALL(_t, _n, prod, OSC_OP_PROD)
OSC_TO_ALL_BITWISE_TYPES(OSC_DEF_AND_TO_ALL)
OSC_TO_ALL_BITWISE_TYPES(OSC_DEF_OR_TO_ALL)
OSC_TO_ALL_BITWISE_TYPES(OSC_DEF_XOR_TO_ALL)
OSC_TO_ALL_MINMAX_TYPES(OSC_DEF_MIN_TO_ALL)
OSC_TO_ALL_MINMAX_TYPES(OSC_DEF_MAX_TO_ALL)
OSC_TO_ALL_ARITH_TYPES(OSC_DEF_SUM_TO_ALL)
OSC_TO_ALL_ARITH_TYPES(OSC_DEF_PROD_TO_ALL)

#define OSC_DEF_SIZED_COLL(_size)                                              \
  void shmem_fcollect##_size(void *target, const void *source, size_t nelems, \
                             int PE_start, int logPE_stride, int PE_size,     \
                             long *pSync) {                                    \
    osc_team t;                                                                \
    size_t nb = 0;                                                             \
    (void)pSync;                                                               \
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK ||       \
        sized_bytes(nelems, (_size) / 8, &nb) != OSC_OK)                       \
      return;                                                                  \
    do_fcollect(&t, target, source, nb);                                       \
  }                                                                            \
  void shmem_collect##_size(void *target, const void *source, size_t nelems,  \
                            int PE_start, int logPE_stride, int PE_size,      \
                            long *pSync) {                                     \
    osc_team t;                                                                \
    size_t nb = 0;                                                             \
    (void)pSync;                                                               \
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK ||       \
        sized_bytes(nelems, (_size) / 8, &nb) != OSC_OK)                       \
      return;                                                                  \
    uniform_collect(&t, target, source, nb);                                   \
  }                                                                            \
  void shmem_alltoall##_size(void *target, const void *source, size_t nelems, \
                             int PE_start, int logPE_stride, int PE_size,     \
                             long *pSync) {                                    \
    osc_team t;                                                                \
    size_t nb = 0;                                                             \
    (void)pSync;                                                               \
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK ||       \
        sized_bytes(nelems, (_size) / 8, &nb) != OSC_OK)                       \
      return;                                                                  \
    do_alltoall(&t, target, source, nb);                                       \
  }                                                                            \
  void shmem_alltoalls##_size(void *target, const void *source,               \
                              ptrdiff_t dst, ptrdiff_t sst, size_t nelems,    \
                              int PE_start, int logPE_stride, int PE_size,    \
                              long *pSync) {                                   \
    osc_team t;                                                                \
    (void)pSync;                                                               \
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK) return; \
    do_alltoalls(&t, target, source, dst, sst, (_size) / 8, nelems);           \
  }                                                                            \
  void shmem_broadcast##_size(void *target, const void *source,               \
                              size_t nelems, int PE_root, int PE_start,       \
                              int logPE_stride, int PE_size, long *pSync) {   \
    osc_team t;                                                                \
    size_t nb = 0;                                                             \
    (void)pSync;                                                               \
    if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK ||       \
        sized_bytes(nelems, (_size) / 8, &nb) != OSC_OK)                       \
      return;                                                                  \
    do_broadcast(&t, target, source, nb, PE_root, 1);                          \
  }
OSC_DEF_SIZED_COLL(32)
OSC_DEF_SIZED_COLL(64)

/* With one host thread standing for every PE, a barrier or sync over an
 * active set is completion of the device work enqueued so far. */
void shmem_barrier(int PE_start, int logPE_stride, int PE_size, long *pSync) {
  osc_team t;
  int rc;
  (void)pSync;
  if (osc_active_set(PE_start, logPE_stride, PE_size, &t) != OSC_OK) return;
  if (osc_g.spmd) {
    osc_spmd_sync(&t);
    return;
  }
  rc = oscd_device_sync();
  if (rc != OSC_OK) osc_hip_error(rc, "barrier");
}
void shmem_sync_deprecated(int PE_start, int logPE_stride, int PE_size,
                           long *pSync) {
  shmem_barrier(PE_start, logPE_stride, PE_size, pSync);
}
void shmem_sync_all(void) {
  if (osc_g.spmd)
    osc_spmd_sync(SHMEM_TEAM_WORLD);
  else if (osc_g.initialized)
    oscd_device_sync();
}
