/*
 * osc_p2p.c - point-to-point synchronization on symmetric variables:
 * shmem_<T>_test / wait_until and their _all, _any, _some and _vector
 * forms, and the deprecated shmem_<T>_wait (OpenSHMEM 1.5 section 9.10;
 * reference declarations include/shmem/api.h:1472-2020, implementation
 * src/api/pt2pt_synch/ over src/shmemc/ucx/test.c and waituntil.c).
 *
 * The variables live in device memory (the symmetric heap), so the host
 * reads them with a copy.  A wait first lets the work this process has
 * already enqueued finish, then polls; it blocks until another thread's
 * put (an SPMD PE thread, or any host thread) satisfies it, as in the
 * reference.  Elements whose status entry is nonzero are excluded.  A bad
 * address, comparison or count returns at once with the error recorded
 * and the routine's "nothing satisfied" value.
 *
 * Two behaviours follow the specification where the reference differs:
 * test_all is 1 when every included element satisfies the comparison (the
 * reference compares the count with nelems, so any excluded element makes
 * it 0), and unsigned types compare as unsigned (the reference compares
 * through the signed type of the same width).
 */
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/shmem.h"
#include "osc_internal.h"

typedef int (*ok_fn)(const void *v, int cmp, const void *x);

#define OSC_P2P_OK(_type, _typename)                                           \
  static int ok_##_typename(const void *pv, int cmp, const void *px) {         \
    _type v, x;                                                                \
    memcpy(&v, pv, sizeof(v));                                                 \
    memcpy(&x, px, sizeof(x));                                                 \
    switch (cmp) {                                                             \
      case SHMEM_CMP_EQ: return v == x;                                        \
      case SHMEM_CMP_NE: return v != x;                                        \
      case SHMEM_CMP_GT: return v > x;                                         \
      case SHMEM_CMP_LE: return v <= x;                                        \
      case SHMEM_CMP_LT: return v < x;                                         \
      default: return v >= x;                                                  \
    }                                                                          \
  }

enum { P_ALL, P_ANY, P_SOME };

/* Host copy of nelems variables of es bytes at the symmetric address. */
static int fetch(const void *ivars, size_t es, size_t nelems, void *out) {
  size_t off;
  int rc;
  if (ivars == NULL || ((uintptr_t)ivars % es) != 0 ||
      osc_sym_range(ivars, es * nelems, &off) != OSC_OK)
    return osc_set_error(OSC_ERR_PTR, "synchronization variables at %p", ivars);
  rc = oscd_memcpy_d2h(out, ivars, es * nelems);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "synchronization read");
}

/* One test, or a wait, over nelems variables.  cv is the comparison value,
 * or per element when cv_step is the element size.  Returns, for P_ALL,
 * 1 if every included element satisfies the comparison (also when none
 * is included), else 0; for P_ANY, the index of the first satisfying
 * included element, or SIZE_MAX; for P_SOME, the number of satisfying
 * included elements, whose indices go to indices[] in ascending order.
 * A wait returns once the test would succeed; with no included element it
 * returns at once. */
static size_t p2p(const void *ivars, size_t es, size_t nelems,
                  const int *status, int cmp, const void *cv, size_t cv_step,
                  ok_fn ok, int kind, size_t *indices, int wait) {
  const size_t fail = kind == P_ANY ? SIZE_MAX : 0;
  char one[16];
  char *vals = one;
  size_t i, included = 0, hits = 0, first = SIZE_MAX;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised"), fail;
  if (cmp < SHMEM_CMP_EQ || cmp > SHMEM_CMP_GE)
    return osc_set_error(OSC_ERR_ARG, "comparison %d", cmp), fail;
  if (cv == NULL) return osc_set_error(OSC_ERR_ARG, "NULL comparison values"), fail;
  if (kind == P_SOME && indices == NULL && nelems > 0)
    return osc_set_error(OSC_ERR_ARG, "NULL indices"), fail;
  if (nelems > ((size_t)-1) / es) return osc_set_error(OSC_ERR_ARG, "nelems overflow"), fail;
  for (i = 0; i < nelems; ++i) included += status == NULL || status[i] == 0;
  if (included == 0) return kind == P_ALL ? 1 : fail;
  if (es * nelems > sizeof(one)) {
    vals = (char *)malloc(es * nelems);
    if (vals == NULL) return osc_set_error(OSC_ERR_NOMEM, "synchronization buffer"), fail;
  }
  if (wait) oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  for (;;) {
    int done;
    if (fetch(ivars, es, nelems, vals) != OSC_OK) {
      hits = 0;
      first = SIZE_MAX;
      break;
    }
    hits = 0;
    first = SIZE_MAX;
    for (i = 0; i < nelems; ++i) {
      if (status != NULL && status[i] != 0) continue;
      if (!ok(vals + i * es, cmp, (const char *)cv + i * cv_step)) continue;
      if (first == SIZE_MAX) first = i;
      if (kind == P_SOME) indices[hits] = i;
      ++hits;
    }
    done = kind == P_ALL ? hits == included : hits > 0;
    if (done || !wait) break;
    sched_yield();
  }
  if (vals != one) free(vals);
  if (kind == P_ALL) return hits == included;
  return kind == P_ANY ? first : hits;
}

#define OSC_DEF_TEST_WAIT(_type, _typename)                                    \
  int shmem_##_typename##_test(_type *ivar, int cmp, _type cmp_value) {        \
    return (int)p2p(ivar, sizeof(_type), 1, NULL, cmp, &cmp_value, 0,          \
                    ok_##_typename, P_ALL, NULL, 0);                           \
  }                                                                            \
  void shmem_##_typename##_wait_until(_type *ivar, int cmp, _type cmp_value) { \
    (void)p2p(ivar, sizeof(_type), 1, NULL, cmp, &cmp_value, 0,                \
              ok_##_typename, P_ALL, NULL, 1);                                 \
  }

#define OSC_DEF_P2P_SETS(_type, _typename)                                     \
  int shmem_##_typename##_test_all(_type *ivars, size_t nelems,                \
                                   const int *status, int cmp,                 \
                                   _type cmp_value) {                          \
    return (int)p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,  \
                    ok_##_typename, P_ALL, NULL, 0);                           \
  }                                                                            \
  size_t shmem_##_typename##_test_any(_type *ivars, size_t nelems,             \
                                      const int *status, int cmp,              \
                                      _type cmp_value) {                       \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,       \
               ok_##_typename, P_ANY, NULL, 0);                                \
  }                                                                            \
  size_t shmem_##_typename##_test_some(_type *ivars, size_t nelems,            \
                                       size_t *indices, const int *status,     \
                                       int cmp, _type cmp_value) {             \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,       \
               ok_##_typename, P_SOME, indices, 0);                            \
  }                                                                            \
  void shmem_##_typename##_wait_until_all(_type *ivars, size_t nelems,         \
                                          const int *status, int cmp,          \
                                          _type cmp_value) {                   \
    (void)p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,        \
              ok_##_typename, P_ALL, NULL, 1);                                 \
  }                                                                            \
  size_t shmem_##_typename##_wait_until_any(_type *ivars, size_t nelems,       \
                                            const int *status, int cmp,        \
                                            _type cmp_value) {                 \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,       \
               ok_##_typename, P_ANY, NULL, 1);                                \
  }                                                                            \
  size_t shmem_##_typename##_wait_until_some(_type *ivars, size_t nelems,      \
                                             size_t *indices,                  \
                                             const int *status, int cmp,       \
                                             _type cmp_value) {                \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, &cmp_value, 0,       \
               ok_##_typename, P_SOME, indices, 1);                            \
  }                                                                            \
  int shmem_##_typename##_test_all_vector(_type *ivars, size_t nelems,         \
                                          const int *status, int cmp,          \
                                          _type *cmp_values) {                 \
    return (int)p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,     \
                    sizeof(_type), ok_##_typename, P_ALL, NULL, 0);            \
  }                                                                            \
  size_t shmem_##_typename##_test_any_vector(_type *ivars, size_t nelems,      \
                                             const int *status, int cmp,       \
                                             _type *cmp_values) {              \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,          \
               sizeof(_type), ok_##_typename, P_ANY, NULL, 0);                 \
  }                                                                            \
  size_t shmem_##_typename##_test_some_vector(                                 \
      _type *ivars, size_t nelems, size_t *indices, const int *status,         \
      int cmp, _type *cmp_values) {                                            \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,          \
               sizeof(_type), ok_##_typename, P_SOME, indices, 0);             \
  }                                                                            \
  void shmem_##_typename##_wait_until_all_vector(_type *ivars, size_t nelems,  \
                                                 const int *status, int cmp,   \
                                                 _type *cmp_values) {          \
    (void)p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,           \
              sizeof(_type), ok_##_typename, P_ALL, NULL, 1);                  \
  }                                                                            \
  size_t shmem_##_typename##_wait_until_any_vector(                            \
      _type *ivars, size_t nelems, const int *status, int cmp,                 \
      _type *cmp_values) {                                                     \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,          \
               sizeof(_type), ok_##_typename, P_ANY, NULL, 1);                 \
  }                                                                            \
  size_t shmem_##_typename##_wait_until_some_vector(                           \
      _type *ivars, size_t nelems, size_t *indices, const int *status,         \
      int cmp, _type *cmp_values) {                                            \
    return p2p(ivars, sizeof(_type), nelems, status, cmp, cmp_values,          \
               sizeof(_type), ok_##_typename, P_SOME, indices, 1);             \
  }

/* deprecated: wait while *ivar == cmp_value */
#define OSC_DEF_WAIT(_type, _typename)                                         \
  void shmem_##_typename##_wait(_type *ivar, _type cmp_value) {                \
    shmem_##_typename##_wait_until(ivar, SHMEM_CMP_NE, cmp_value);             \
  }

OSC_P2P_SYNC_TYPES(OSC_P2P_OK)
OSC_P2P_SYNC_TYPES(OSC_DEF_TEST_WAIT)
OSC_P2P_SYNC_TYPES(OSC_DEF_P2P_SETS)
OSC_P2P_DEPR_TYPES(OSC_P2P_OK)
OSC_P2P_DEPR_TYPES(OSC_DEF_TEST_WAIT)
OSC_P2P_SYNC_TYPES(OSC_DEF_WAIT)
OSC_P2P_DEPR_TYPES(OSC_DEF_WAIT)
