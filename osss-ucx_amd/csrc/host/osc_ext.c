/*
 * osc_ext.c - the extensions of include/shmemx.h beside the OpenSHMEM API
 * (reference include/shmemx.h and src/api/extensions/): completion tests,
 * session hints, named heaps, the interoperability query and the
 * remote-address lookup.  shmemx_wtime lives with the other setup calls in
 * osc_amo.c.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/shmemx.h"
#include "osc_internal.h"

void *shmemx_lookup_remote_addr(void *addr, int pe) { return shmem_ptr(addr, pe); }

/* 1 if the stream has no pending work, 0 if it has, 0 with the error
 * recorded on failure */
static int idle(void *stream) {
  const int q = oscd_stream_query(stream);
  if (q < 0) return osc_hip_error(q, "stream query"), 0;
  return q;
}

int shmemx_ctx_quiet_test(shmem_ctx_t ctx) {
  void *st = NULL;
  if (osc_ctx_target(ctx, 0, &st) < 0) return 0;
  return idle(st);
}

int shmemx_ctx_fence_test(shmem_ctx_t ctx) {
  void *st = NULL;
  return osc_ctx_target(ctx, 0, &st) >= 0;
}

int shmemx_quiet_test(void) { return shmemx_ctx_quiet_test(SHMEM_CTX_DEFAULT); }
int shmemx_fence_test(void) { return shmemx_ctx_fence_test(SHMEM_CTX_DEFAULT); }
void shmemx_ctx_session_start(shmem_ctx_t ctx) { (void)ctx; }
void shmemx_ctx_session_stop(shmem_ctx_t ctx) { (void)ctx; }

int shmemx_query_interoperability(int property) {
  (void)property;
  return 0;
}

/* ---- named heaps ----------------------------------------------------------
 * Names get indices in first-lookup order, as the reference's allocator
 * does (src/api/allocator/xmemalloc.c:84-92); the table lives for the
 * process, as there. */
#define OSC_MAX_HEAP_NAMES 256
static pthread_mutex_t g_names_mu = PTHREAD_MUTEX_INITIALIZER;
static char *g_names[OSC_MAX_HEAP_NAMES];
static int g_nnames;

shmemx_heap_index_t shmemx_name_to_index(const char *name) {
  int i, r = -1;
  if (name == NULL) return osc_set_error(OSC_ERR_ARG, "NULL heap name"), -1;
  pthread_mutex_lock(&g_names_mu);
  for (i = 0; i < g_nnames; ++i)
    if (strcmp(g_names[i], name) == 0) r = i;
  if (r < 0 && g_nnames < OSC_MAX_HEAP_NAMES) {
    g_names[g_nnames] = strdup(name);
    if (g_names[g_nnames] != NULL) r = g_nnames++;
  }
  pthread_mutex_unlock(&g_names_mu);
  if (r < 0) osc_set_error(OSC_ERR_NOMEM, "heap name table");
  return r;
}

const char *shmemx_index_to_name(shmemx_heap_index_t index) {
  const char *n = NULL;
  pthread_mutex_lock(&g_names_mu);
  if (index >= 0 && index < g_nnames) n = g_names[index];
  pthread_mutex_unlock(&g_names_mu);
  return n;
}

/* one heap, index 0 */
static int heap_ok(shmemx_heap_index_t index) {
  if (index == 0) return 1;
  osc_set_error(OSC_ERR_ARG, "heap index %d: there is one heap (index 0)", index);
  return 0;
}

void *shmemx_malloc_by_index(shmemx_heap_index_t index, size_t s) {
  return heap_ok(index) ? shmem_malloc(s) : NULL;
}
void *shmemx_calloc_by_index(shmemx_heap_index_t index, size_t n, size_t s) {
  return heap_ok(index) ? shmem_calloc(n, s) : NULL;
}
void shmemx_free_by_index(shmemx_heap_index_t index, void *p) {
  if (heap_ok(index)) shmem_free(p);
}
void *shmemx_realloc_by_index(shmemx_heap_index_t index, void *p, size_t s) {
  return heap_ok(index) ? shmem_realloc(p, s) : NULL;
}
void *shmemx_align_by_index(shmemx_heap_index_t index, size_t a, size_t s) {
  return heap_ok(index) ? shmem_align(a, s) : NULL;
}
void *shmemx_malloc_by_name(const char *name, size_t s) {
  return shmemx_malloc_by_index(shmemx_name_to_index(name), s);
}
void *shmemx_calloc_by_name(const char *name, size_t n, size_t s) {
  return shmemx_calloc_by_index(shmemx_name_to_index(name), n, s);
}
void shmemx_free_by_name(const char *name, void *p) {
  shmemx_free_by_index(shmemx_name_to_index(name), p);
}
void *shmemx_realloc_by_name(const char *name, void *p, size_t s) {
  return shmemx_realloc_by_index(shmemx_name_to_index(name), p, s);
}
void *shmemx_align_by_name(const char *name, size_t a, size_t s) {
  return shmemx_align_by_index(shmemx_name_to_index(name), a, s);
}
