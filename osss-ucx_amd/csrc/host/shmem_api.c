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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

void shmem_init(void) {
  const char *np = getenv("OSC_NPES");
  const char *hs = getenv("SHMEM_SYMMETRIC_SIZE");
  const char *dv = getenv("OSC_DEVICE");
  size_t heap = (size_t)32 << 20; /* reference default, configure.ac:23 */
  int npes = np ? atoi(np) : 8;
  int rc;
  if (osc_g.initialized) return;
  if (hs != NULL && osc_parse_size(hs, &heap) != OSC_OK) {
    fprintf(stderr, "shmem_init: bad SHMEM_SYMMETRIC_SIZE=%s\n", hs);
    abort();
  }
  rc = osc_init(npes, heap, dv ? atoi(dv) : 0);
  if (rc != OSC_OK) {
    fprintf(stderr, "shmem_init: %s (%s)\n", osc_strerror(rc), osc_last_error());
    abort();
  }
}

void shmem_finalize(void) { osc_finalize(); }

int shmem_my_pe(void) { return osc_g.initialized ? osc_my_pe_get() : -1; }
int shmem_n_pes(void) { return osc_g.initialized ? osc_g.npes : -1; }

/* One host thread acts for all PEs, so a barrier is completion of all
 * enqueued device work. */
void shmem_barrier_all(void) {
  if (osc_g.initialized) oscd_device_sync();
}
int shmem_team_sync(shmem_team_t team) {
  if (osc_team_get(team) == NULL) return -1;
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
void shmem_info_get_name(char *name) {
  if (name) strcpy(name, "osc-mi355x (OpenSHMEM collectives on gfx950)");
}

/* ---- symmetric heap ---------------------------------------------------------- */

static void *heap_alloc(size_t size, size_t align) {
  size_t off;
  int rc;
  if (!osc_g.initialized) return NULL;
  pthread_mutex_lock(&osc_g.lock);
  rc = osc_heap_alloc(size, align, &off);
  pthread_mutex_unlock(&osc_g.lock);
  if (rc != OSC_OK) {
    osc_set_error(rc, "symmetric heap exhausted (%zu bytes requested)", size);
    return NULL;
  }
  return osc_g.base + (size_t)osc_my_pe_get() * osc_g.slice + off;
}

void *shmem_malloc(size_t size) { return heap_alloc(size, 0); }

void *shmem_align(size_t alignment, size_t size) {
  if (alignment == 0 || (alignment & (alignment - 1)) != 0) return NULL;
  return heap_alloc(size, alignment);
}

void *shmem_calloc(size_t count, size_t size) {
  char *p;
  size_t off, bytes;
  int pe;
  if (size != 0 && count > ((size_t)-1) / size) return NULL;
  bytes = count * size;
  p = (char *)heap_alloc(bytes, 0);
  if (p == NULL) return NULL;
  osc_sym_offset(p, NULL, &off);
  for (pe = 0; pe < osc_g.npes; ++pe)
    oscd_memset_async(osc_g.base + (size_t)pe * osc_g.slice + off, 0, bytes,
                      osc_g.stream);
  oscd_stream_sync(osc_g.stream);
  return p;
}

void shmem_free(void *ptr) {
  size_t off;
  if (ptr == NULL || osc_sym_offset(ptr, NULL, &off) != OSC_OK) return;
  /* outstanding kernels may still use the object */
  oscd_stream_sync(osc_g.stream);
  pthread_mutex_lock(&osc_g.lock);
  osc_heap_free(off);
  pthread_mutex_unlock(&osc_g.lock);
}

void *shmem_realloc(void *ptr, size_t size) {
  size_t off, old;
  char *np;
  int pe, rc;
  if (ptr == NULL) return shmem_malloc(size);
  if (size == 0) {
    shmem_free(ptr);
    return NULL;
  }
  if (osc_sym_offset(ptr, NULL, &off) != OSC_OK) return NULL;
  pthread_mutex_lock(&osc_g.lock);
  rc = osc_heap_block_size(off, &old);
  pthread_mutex_unlock(&osc_g.lock);
  if (rc != OSC_OK) return NULL;
  np = (char *)shmem_malloc(size);
  if (np == NULL) return NULL;
  {
    size_t noff, keep = old < size ? old : size;
    osc_sym_offset(np, NULL, &noff);
    for (pe = 0; pe < osc_g.npes; ++pe)
      oscd_copy(osc_g.base + (size_t)pe * osc_g.slice + noff,
                osc_g.base + (size_t)pe * osc_g.slice + off, keep, osc_g.stream);
  }
  shmem_free(ptr);
  return np;
}

void *shmem_ptr(const void *dest, int pe) { return osc_pe_ptr(dest, pe); }

int shmem_addr_accessible(const void *addr, int pe) {
  return osc_pe_ptr(addr, pe) != NULL;
}

/* ---- RMA between emulated PEs ------------------------------------------------ */
/* The local side may be a symmetric (device) address or host memory. */

static int is_sym(const void *p) {
  const char *c = (const char *)p;
  return osc_g.initialized && c >= osc_g.base &&
         c < osc_g.base + (size_t)osc_g.npes * osc_g.slice;
}

static void put_get(void *dst, const void *src, size_t nbytes, int blocking) {
  if (nbytes == 0) return;
  if (is_sym(dst) && is_sym(src)) {
    oscd_copy(dst, src, nbytes, osc_g.stream);
    if (blocking) oscd_stream_sync(osc_g.stream);
  } else if (is_sym(dst)) {
    oscd_stream_sync(osc_g.stream);
    oscd_memcpy_h2d(dst, src, nbytes);
  } else {
    oscd_stream_sync(osc_g.stream);
    oscd_memcpy_d2h(dst, src, nbytes);
  }
}

void shmem_putmem(void *dest, const void *source, size_t nbytes, int pe) {
  void *d = osc_pe_ptr(dest, pe);
  if (d != NULL) put_get(d, source, nbytes, 1);
}
void shmem_putmem_nbi(void *dest, const void *source, size_t nbytes, int pe) {
  void *d = osc_pe_ptr(dest, pe);
  if (d != NULL) put_get(d, source, nbytes, 0);
}
void shmem_getmem(void *dest, const void *source, size_t nbytes, int pe) {
  const void *s = osc_pe_ptr(source, pe);
  if (s != NULL) put_get(dest, s, nbytes, 1);
}
void shmem_getmem_nbi(void *dest, const void *source, size_t nbytes, int pe) {
  const void *s = osc_pe_ptr(source, pe);
  if (s != NULL) put_get(dest, s, nbytes, 0);
}

/* ---- collectives ---------------------------------------------------------------- */

static int finish(int rc) {
  if (rc != OSC_OK) return rc;
  rc = oscd_stream_sync(osc_g.stream);
  return rc == OSC_OK ? 0 : osc_hip_error(rc, "hipStreamSynchronize");
}

static int blocking_reduce(shmem_team_t team, int op, int dtype, void *dest,
                           const void *src, size_t n) {
  if (dtype < 0)
    return osc_set_error(OSC_ERR_UNSUPPORTED, "long double reduction (no 80-bit type on gfx950)");
  return finish(osc_reduce(team, op, dtype, dest, src, n, NULL));
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
    return finish(osc_fcollect(team, dest, source, nelems * sizeof(_type),     \
                               NULL));                                         \
  }                                                                            \
  int shmem_##_typename##_alltoall(shmem_team_t team, _type *dest,             \
                                   const _type *source, size_t nelems) {       \
    return finish(osc_alltoall(team, dest, source, nelems * sizeof(_type),     \
                               NULL));                                         \
  }                                                                            \
  int shmem_##_typename##_broadcast(shmem_team_t team, _type *dest,            \
                                    const _type *source, size_t nelems,        \
                                    int PE_root) {                             \
    return finish(osc_broadcast(team, dest, source, nelems * sizeof(_type),    \
                                PE_root, NULL));                               \
  }
OSC_RMA_TYPES(OSC_DEF_MOVE)

int shmem_fcollectmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes) {
  return finish(osc_fcollect(team, dest, source, nbytes, NULL));
}
int shmem_alltoallmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes) {
  return finish(osc_alltoall(team, dest, source, nbytes, NULL));
}
int shmem_broadcastmem(shmem_team_t team, void *dest, const void *source,
                       size_t nbytes, int PE_root) {
  return finish(osc_broadcast(team, dest, source, nbytes, PE_root, NULL));
}
