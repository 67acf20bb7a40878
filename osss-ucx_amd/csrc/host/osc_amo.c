/*
 * osc_amo.c - atomic memory operations on symmetric words (OpenSHMEM 1.5
 * section 9.8; reference declarations include/shmem/api.h:2030-2990,
 * implementation src/api/atomics/ over the UCX atomics of
 * src/shmemc/ucx/).
 *
 * Every operation is one device atomic on the target PE's copy, done by a
 * one-wave kernel on the library stream (oscd_amo), so operations from any
 * host thread -- SPMD PE threads included -- are atomic with respect to
 * each other and ordered after the puts enqueued before them.  Values
 * travel as bit patterns of their width (4 or 8 bytes): float and double
 * take part only in fetch, set and swap, where bits are all that matter.
 */
#include <sched.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include "../../../include/shmem.h"
#include "osc_internal.h"

static uint64_t bits_of(const void *v, size_t w) {
  uint64_t u = 0;
  memcpy(&u, v, w);
  return u;
}

/* the operation on dest's copy on PE pe of context ctx (its stream; a
 * team context's PE numbers are team indices); old != NULL waits for the
 * value */
static int amo_ctx(shmem_ctx_t ctx, const void *dest, size_t w, int pe, int op,
                   uint64_t v, uint64_t cond, uint64_t *old) {
  size_t off;
  char *target;
  void *st = NULL;
  int rc;
  if (old != NULL) *old = 0;
  pe = osc_ctx_target(ctx, pe, &st);
  if (pe < 0) return osc_last_error_code();
  if (pe >= osc_g.npes) return osc_set_error(OSC_ERR_ARG, "PE %d", pe);
  rc = osc_sym_range(dest, w, &off);
  if (rc != OSC_OK) return rc;
  target = osc_g.base + (size_t)pe * osc_g.slice + off;
  if (((uintptr_t)target % w) != 0)
    return osc_set_error(OSC_ERR_PTR, "atomic target %p is not %zu-byte aligned", dest, w);
  rc = oscd_amo(target, (int)w, op, v, cond, old, st, osc_g.wait_spin_us);
  if (rc == OSC_ERR_HIP) return osc_hip_error(rc, "atomic operation");
  return rc == OSC_OK ? OSC_OK : osc_set_error(rc, "atomic operation");
}

/* [p, p + w) inside one PE's usable heap (no error recorded) */
static int in_heap(const void *p, size_t w) {
  const char *c = (const char *)p;
  size_t d;
  if (c < osc_g.base || c >= osc_g.base + (size_t)osc_g.npes * osc_g.slice) return 0;
  d = (size_t)(c - osc_g.base) % osc_g.slice;
  return d < osc_g.heap_bytes && w <= osc_g.heap_bytes - d;
}

/* *fetch = the w bytes of value; fetch may be host memory or symmetric */
static void deliver(void *fetch, uint64_t value, size_t w) {
  if (fetch == NULL) {
    osc_set_error(OSC_ERR_ARG, "NULL fetch address");
  } else if (in_heap(fetch, w)) {
    const int rc = oscd_memcpy_h2d(fetch, &value, w);
    if (rc != OSC_OK) osc_hip_error(rc, "atomic fetch delivery");
  } else {
    memcpy(fetch, &value, w);
  }
}

#define OSC_AMO_FROM(_type, _typename)                                         \
  static _type from_##_typename(uint64_t u) {                                  \
    _type r;                                                                   \
    memcpy(&r, &u, sizeof(r));                                                 \
    return r;                                                                  \
  }

/* fetching operation returning _type */
#define OSC_AMO_GET(_type, _typename, _ctx, _dest, _pe, _op, _v, _cond)        \
  do {                                                                         \
    uint64_t old_;                                                             \
    amo_ctx(_ctx, _dest, sizeof(_type), _pe, _op, _v, _cond, &old_);           \
    return from_##_typename(old_);                                             \
  } while (0)

/* fetching operation delivered to *fetch */
#define OSC_AMO_NBI(_type, _ctx, _fetch, _dest, _pe, _op, _v, _cond)           \
  do {                                                                         \
    uint64_t old_;                                                             \
    if (amo_ctx(_ctx, _dest, sizeof(_type), _pe, _op, _v, _cond, &old_) ==     \
        OSC_OK)                                                                \
      deliver(_fetch, old_, sizeof(_type));                                    \
  } while (0)

#define B(_x) bits_of(&(_x), sizeof(_x))

/* Every routine is its context form on SHMEM_CTX_DEFAULT. */
#define OSC_DEF_AMO_EXT(_type, _typename)                                      \
  _type shmem_ctx_##_typename##_atomic_fetch(shmem_ctx_t ctx,                  \
      const _type *source, int pe) {                                           \
    OSC_AMO_GET(_type, _typename, ctx, source, pe, OSCD_AMO_FETCH, 0, 0);      \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch(const _type *source, int pe) {        \
    return shmem_ctx_##_typename##_atomic_fetch(SHMEM_CTX_DEFAULT, source,     \
        pe);                                                                   \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_set(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe) {                                                   \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_SET, B(value), 0, NULL);    \
  }                                                                            \
  void shmem_##_typename##_atomic_set(_type *dest, _type value, int pe) {      \
    shmem_ctx_##_typename##_atomic_set(SHMEM_CTX_DEFAULT, dest, value, pe);    \
  }                                                                            \
  _type shmem_ctx_##_typename##_atomic_swap(shmem_ctx_t ctx, _type *dest,      \
      _type value, int pe) {                                                   \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_SWAP, B(value), 0);  \
  }                                                                            \
  _type shmem_##_typename##_atomic_swap(_type *dest, _type value, int pe) {    \
    return shmem_ctx_##_typename##_atomic_swap(SHMEM_CTX_DEFAULT, dest, value, \
        pe);                                                                   \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_nbi(shmem_ctx_t ctx, _type *fetch, \
      const _type *source, int pe) {                                           \
    OSC_AMO_NBI(_type, ctx, fetch, source, pe, OSCD_AMO_FETCH, 0, 0);          \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_nbi(_type *fetch, const _type *source, \
      int pe) {                                                                \
    shmem_ctx_##_typename##_atomic_fetch_nbi(SHMEM_CTX_DEFAULT, fetch, source, \
        pe);                                                                   \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_swap_nbi(shmem_ctx_t ctx, _type *fetch,  \
      _type *dest, _type value, int pe) {                                      \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_SWAP, B(value), 0);      \
  }                                                                            \
  void shmem_##_typename##_atomic_swap_nbi(_type *fetch, _type *dest,          \
      _type value, int pe) {                                                   \
    shmem_ctx_##_typename##_atomic_swap_nbi(SHMEM_CTX_DEFAULT, fetch, dest,    \
        value, pe);                                                            \
  }

#define OSC_DEF_AMO_STD(_type, _typename)                                      \
  _type shmem_ctx_##_typename##_atomic_compare_swap(shmem_ctx_t ctx,           \
      _type *dest, _type cond, _type value, int pe) {                          \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_CSWAP, B(value),     \
        B(cond));                                                              \
  }                                                                            \
  _type shmem_##_typename##_atomic_compare_swap(_type *dest, _type cond,       \
      _type value, int pe) {                                                   \
    return shmem_ctx_##_typename##_atomic_compare_swap(SHMEM_CTX_DEFAULT,      \
        dest, cond, value, pe);                                                \
  }                                                                            \
  _type shmem_ctx_##_typename##_atomic_fetch_inc(shmem_ctx_t ctx, _type *dest, \
      int pe) {                                                                \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_ADD, 1, 0);          \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch_inc(_type *dest, int pe) {            \
    return shmem_ctx_##_typename##_atomic_fetch_inc(SHMEM_CTX_DEFAULT, dest,   \
        pe);                                                                   \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_inc(shmem_ctx_t ctx, _type *dest,        \
      int pe) {                                                                \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_ADD, 1, 0, NULL);           \
  }                                                                            \
  void shmem_##_typename##_atomic_inc(_type *dest, int pe) {                   \
    shmem_ctx_##_typename##_atomic_inc(SHMEM_CTX_DEFAULT, dest, pe);           \
  }                                                                            \
  _type shmem_ctx_##_typename##_atomic_fetch_add(shmem_ctx_t ctx, _type *dest, \
      _type value, int pe) {                                                   \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_ADD, B(value), 0);   \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch_add(_type *dest, _type value,         \
      int pe) {                                                                \
    return shmem_ctx_##_typename##_atomic_fetch_add(SHMEM_CTX_DEFAULT, dest,   \
        value, pe);                                                            \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_add(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe) {                                                   \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_ADD, B(value), 0, NULL);    \
  }                                                                            \
  void shmem_##_typename##_atomic_add(_type *dest, _type value, int pe) {      \
    shmem_ctx_##_typename##_atomic_add(SHMEM_CTX_DEFAULT, dest, value, pe);    \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_compare_swap_nbi(shmem_ctx_t ctx,        \
      _type *fetch, _type *dest, _type cond, _type value, int pe) {            \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_CSWAP, B(value),         \
        B(cond));                                                              \
  }                                                                            \
  void shmem_##_typename##_atomic_compare_swap_nbi(_type *fetch, _type *dest,  \
      _type cond, _type value, int pe) {                                       \
    shmem_ctx_##_typename##_atomic_compare_swap_nbi(SHMEM_CTX_DEFAULT, fetch,  \
        dest, cond, value, pe);                                                \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_inc_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, int pe) {                                     \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_ADD, 1, 0);              \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_inc_nbi(_type *fetch, _type *dest,     \
      int pe) {                                                                \
    shmem_ctx_##_typename##_atomic_fetch_inc_nbi(SHMEM_CTX_DEFAULT, fetch,     \
        dest, pe);                                                             \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_add_nbi(shmem_ctx_t ctx,           \
      _type *fetch, const _type *dest, _type value, int pe) {                  \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_ADD, B(value), 0);       \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_add_nbi(_type *fetch,                  \
      const _type *dest, _type value, int pe) {                                \
    shmem_ctx_##_typename##_atomic_fetch_add_nbi(SHMEM_CTX_DEFAULT, fetch,     \
        dest, value, pe);                                                      \
  }

#define OSC_DEF_AMO_BITWISE(_type, _typename)                                  \
  _type shmem_ctx_##_typename##_atomic_fetch_and(shmem_ctx_t ctx, _type *dest, \
      _type value, int pe) {                                                   \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_AND, B(value), 0);   \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch_and(_type *dest, _type value,         \
      int pe) {                                                                \
    return shmem_ctx_##_typename##_atomic_fetch_and(SHMEM_CTX_DEFAULT, dest,   \
        value, pe);                                                            \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_and(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe) {                                                   \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_AND, B(value), 0, NULL);    \
  }                                                                            \
  void shmem_##_typename##_atomic_and(_type *dest, _type value, int pe) {      \
    shmem_ctx_##_typename##_atomic_and(SHMEM_CTX_DEFAULT, dest, value, pe);    \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_and_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, _type value, int pe) {                        \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_AND, B(value), 0);       \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_and_nbi(_type *fetch, _type *dest,     \
      _type value, int pe) {                                                   \
    shmem_ctx_##_typename##_atomic_fetch_and_nbi(SHMEM_CTX_DEFAULT, fetch,     \
        dest, value, pe);                                                      \
  }                                                                            \
  _type shmem_ctx_##_typename##_atomic_fetch_or(shmem_ctx_t ctx, _type *dest,  \
      _type value, int pe) {                                                   \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_OR, B(value), 0);    \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch_or(_type *dest, _type value,          \
      int pe) {                                                                \
    return shmem_ctx_##_typename##_atomic_fetch_or(SHMEM_CTX_DEFAULT, dest,    \
        value, pe);                                                            \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_or(shmem_ctx_t ctx, _type *dest,         \
      _type value, int pe) {                                                   \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_OR, B(value), 0, NULL);     \
  }                                                                            \
  void shmem_##_typename##_atomic_or(_type *dest, _type value, int pe) {       \
    shmem_ctx_##_typename##_atomic_or(SHMEM_CTX_DEFAULT, dest, value, pe);     \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_or_nbi(shmem_ctx_t ctx,            \
      _type *fetch, _type *dest, _type value, int pe) {                        \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_OR, B(value), 0);        \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_or_nbi(_type *fetch, _type *dest,      \
      _type value, int pe) {                                                   \
    shmem_ctx_##_typename##_atomic_fetch_or_nbi(SHMEM_CTX_DEFAULT, fetch,      \
        dest, value, pe);                                                      \
  }                                                                            \
  _type shmem_ctx_##_typename##_atomic_fetch_xor(shmem_ctx_t ctx, _type *dest, \
      _type value, int pe) {                                                   \
    OSC_AMO_GET(_type, _typename, ctx, dest, pe, OSCD_AMO_XOR, B(value), 0);   \
  }                                                                            \
  _type shmem_##_typename##_atomic_fetch_xor(_type *dest, _type value,         \
      int pe) {                                                                \
    return shmem_ctx_##_typename##_atomic_fetch_xor(SHMEM_CTX_DEFAULT, dest,   \
        value, pe);                                                            \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_xor(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe) {                                                   \
    amo_ctx(ctx, dest, sizeof(_type), pe, OSCD_AMO_XOR, B(value), 0, NULL);    \
  }                                                                            \
  void shmem_##_typename##_atomic_xor(_type *dest, _type value, int pe) {      \
    shmem_ctx_##_typename##_atomic_xor(SHMEM_CTX_DEFAULT, dest, value, pe);    \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_fetch_xor_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, _type value, int pe) {                        \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_XOR, B(value), 0);       \
  }                                                                            \
  void shmem_##_typename##_atomic_fetch_xor_nbi(_type *fetch, _type *dest,     \
      _type value, int pe) {                                                   \
    shmem_ctx_##_typename##_atomic_fetch_xor_nbi(SHMEM_CTX_DEFAULT, fetch,     \
        dest, value, pe);                                                      \
  }                                                                            \
  void shmem_ctx_##_typename##_atomic_xor_nbi(shmem_ctx_t ctx, _type *fetch,   \
      _type *dest, _type value, int pe) {                                      \
    OSC_AMO_NBI(_type, ctx, fetch, dest, pe, OSCD_AMO_XOR, B(value), 0);       \
  }                                                                            \
  void shmem_##_typename##_atomic_xor_nbi(_type *fetch, _type *dest,           \
      _type value, int pe) {                                                   \
    shmem_ctx_##_typename##_atomic_xor_nbi(SHMEM_CTX_DEFAULT, fetch, dest,     \
        value, pe);                                                            \
  }

#define OSC_DEF_AMO_DEPR_EXT(_type, _typename)                                 \
  _type shmem_##_typename##_fetch(const _type *target, int pe) {               \
    return shmem_##_typename##_atomic_fetch(target, pe);                       \
  }                                                                            \
  void shmem_##_typename##_set(_type *target, _type value, int pe) {           \
    shmem_##_typename##_atomic_set(target, value, pe);                         \
  }                                                                            \
  _type shmem_##_typename##_swap(_type *target, _type value, int pe) {         \
    return shmem_##_typename##_atomic_swap(target, value, pe);                 \
  }

#define OSC_DEF_AMO_DEPR_INT(_type, _typename)                                 \
  _type shmem_##_typename##_cswap(_type *target, _type cond, _type value,      \
                                  int pe) {                                    \
    return shmem_##_typename##_atomic_compare_swap(target, cond, value, pe);   \
  }                                                                            \
  _type shmem_##_typename##_finc(_type *target, int pe) {                      \
    return shmem_##_typename##_atomic_fetch_inc(target, pe);                   \
  }                                                                            \
  void shmem_##_typename##_inc(_type *target, int pe) {                        \
    shmem_##_typename##_atomic_inc(target, pe);                                \
  }                                                                            \
  _type shmem_##_typename##_fadd(_type *target, _type value, int pe) {         \
    return shmem_##_typename##_atomic_fetch_add(target, value, pe);            \
  }                                                                            \
  void shmem_##_typename##_add(_type *target, _type value, int pe) {           \
    shmem_##_typename##_atomic_add(target, value, pe);                         \
  }

OSC_AMO_EXT_TYPES(OSC_AMO_FROM)
OSC_AMO_EXT_TYPES(OSC_DEF_AMO_EXT)
OSC_AMO_STD_TYPES(OSC_DEF_AMO_STD)
OSC_AMO_BITWISE_TYPES(OSC_DEF_AMO_BITWISE)
OSC_AMO_DEPR_INT_TYPES(OSC_DEF_AMO_DEPR_EXT)
OSC_AMO_DEPR_INT_TYPES(OSC_DEF_AMO_DEPR_INT)
OSC_DEF_AMO_DEPR_EXT(float, float)
OSC_DEF_AMO_DEPR_EXT(double, double)

/* ---- distributed locks ----------------------------------------------------
 * The lock word is PE 0's copy of the symmetric long: 0 when free, else
 * the holder's PE + 1 (compare_swap takes it, swap releases it). */
int shmem_test_lock(long *lock) {
  uint64_t old = 1;
  const long me = (long)osc_my_pe_get() + 1;
  if (amo_ctx(SHMEM_CTX_DEFAULT, lock, sizeof(long), 0, OSCD_AMO_CSWAP,
              (uint64_t)me, 0, &old) != OSC_OK)
    return 1;
  return old != 0;
}

void shmem_set_lock(long *lock) {
  const long me = (long)osc_my_pe_get() + 1;
  for (;;) {
    uint64_t old = 1;
    if (amo_ctx(SHMEM_CTX_DEFAULT, lock, sizeof(long), 0, OSCD_AMO_CSWAP,
              (uint64_t)me, 0, &old) != OSC_OK)
      return; /* error recorded: a bad lock address never succeeds */
    if (old == 0) return;
    sched_yield();
  }
}

void shmem_clear_lock(long *lock) {
  /* stream order puts the release after every store of the critical
   * section, blocking or not */
  amo_ctx(SHMEM_CTX_DEFAULT, lock, sizeof(long), 0, OSCD_AMO_SET, 0, 0, NULL);
}

/* ---- deprecated cache routines: no effect ------------------------------- */
void shmem_clear_cache_inv(void) {}
void shmem_set_cache_inv(void) {}
void shmem_clear_cache_line_inv(void *dest) { (void)dest; }
void shmem_set_cache_line_inv(void *dest) { (void)dest; }
void shmem_udcflush(void) {}
void shmem_udcflush_line(void *dest) { (void)dest; }

double shmemx_wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
