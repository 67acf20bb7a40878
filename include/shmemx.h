/*
 * shmemx.h - extensions beside the OpenSHMEM API (the reference's
 * include/shmemx.h); only the wall-clock timer has a counterpart here.
 */
#ifndef OSC_SHMEMX_H
#define OSC_SHMEMX_H

#include "shmem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* seconds from an arbitrary fixed point, monotonic (reference
 * include/shmemx.h:43) */
double shmemx_wtime(void);

/* the device address of addr's copy on PE pe, as shmem_ptr */
void *shmemx_lookup_remote_addr(void *addr, int pe);

/* 1 when the (context's) operations issued so far are complete, without
 * waiting; fences always hold (operations of a context are ordered) */
int shmemx_ctx_fence_test(shmem_ctx_t ctx);
int shmemx_ctx_quiet_test(shmem_ctx_t ctx);
int shmemx_fence_test(void);
int shmemx_quiet_test(void);
/* communication-session hints: no effect */
void shmemx_ctx_session_start(shmem_ctx_t ctx);
void shmemx_ctx_session_stop(shmem_ctx_t ctx);

/* Named heaps.  As in the reference (src/shmemc/heaps.c: one heap), only
 * index 0 can allocate: it is the symmetric heap.  A name gets the next
 * index the first time it is looked up; the calls by name use that index. */
typedef int shmemx_heap_index_t;
shmemx_heap_index_t shmemx_name_to_index(const char *name);
const char *shmemx_index_to_name(shmemx_heap_index_t index);
void *shmemx_malloc_by_index(shmemx_heap_index_t index, size_t s);
void *shmemx_calloc_by_index(shmemx_heap_index_t index, size_t n, size_t s);
void shmemx_free_by_index(shmemx_heap_index_t index, void *p);
void *shmemx_realloc_by_index(shmemx_heap_index_t index, void *p, size_t s);
void *shmemx_align_by_index(shmemx_heap_index_t index, size_t a, size_t s);
void *shmemx_malloc_by_name(const char *name, size_t s);
void *shmemx_calloc_by_name(const char *name, size_t n, size_t s);
void shmemx_free_by_name(const char *name, void *p);
void *shmemx_realloc_by_name(const char *name, void *p, size_t s);
void *shmemx_align_by_name(const char *name, size_t a, size_t s);
#if defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L && !defined(__cplusplus)
#define OSC_GEN_HEAP(_arg1, _op)                                               \
  _Generic((_arg1), shmemx_heap_index_t: shmemx_##_op##_by_index,              \
           char *: shmemx_##_op##_by_name, const char *: shmemx_##_op##_by_name)
#define shmemx_malloc(_arg1, _s) OSC_GEN_HEAP(_arg1, malloc)(_arg1, _s)
#define shmemx_calloc(_arg1, _n, _s) OSC_GEN_HEAP(_arg1, calloc)(_arg1, _n, _s)
#define shmemx_free(_arg1, _p) OSC_GEN_HEAP(_arg1, free)(_arg1, _p)
#define shmemx_realloc(_arg1, _p, _s) OSC_GEN_HEAP(_arg1, realloc)(_arg1, _p, _s)
#define shmemx_align(_arg1, _a, _s) OSC_GEN_HEAP(_arg1, align)(_arg1, _a, _s)
#endif

/* interoperability with other models (reference include/shmemx.h):
 * none is provided, so every query returns 0 */
enum interoperability {
  UPC_THREADS_ARE_PES = 0,
  MPI_PROCESSES_ARE_PES,
  SHMEM_INITIALIZES_MPI,
  MPI_INITIALIZES_SHMEM
};
int shmemx_query_interoperability(int property);

#ifdef __cplusplus
}
#endif
#endif /* OSC_SHMEMX_H */
