/*
 * osc_heap.c - the symmetric offset allocator.
 *
 * Replaces the reference's dlmalloc mspace on a UCX-registered region
 * (src/api/allocator/memalloc.c:29-31, src/ptmalloc/ptmalloc.c).  Because
 * every emulated PE has its slice at the same place in one device arena,
 * one allocator over the offset space [0, heap_bytes) serves all PEs at
 * once: an allocation at offset `off` is the symmetric object
 * base + pe*slice + off on every PE.  No barrier is needed (the reference
 * runs shmem_barrier_all around every malloc/free, src/api/shmalloc.c:56,
 * :131) because one host call acts for all PEs.
 *
 * First fit over an address-ordered block list with coalescing on free.
 * Minimum alignment 256 bytes so every object starts on a 16-byte (wide
 * load) and cache-line boundary.
 */
#include <stdlib.h>

#include "osc_internal.h"

#define OSC_MIN_ALIGN 256u

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int osc_heap_init(size_t heap_bytes) {
  osc_block *b = (osc_block *)calloc(1, sizeof(*b));
  if (b == NULL) return OSC_ERR_NOMEM;
  b->off = 0;
  b->size = heap_bytes;
  b->used = 0;
  b->next = NULL;
  osc_g.blocks = b;
  return OSC_OK;
}

void osc_heap_fini(void) {
  osc_block *b = osc_g.blocks;
  while (b != NULL) {
    osc_block *n = b->next;
    free(b);
    b = n;
  }
  osc_g.blocks = NULL;
}

/* Split block b so that a new block starts at offset `at` (inside b). */
static int split_at(osc_block *b, size_t at) {
  osc_block *n;
  if (at == b->off || at >= b->off + b->size) return OSC_OK;
  n = (osc_block *)calloc(1, sizeof(*n));
  if (n == NULL) return OSC_ERR_NOMEM;
  n->off = at;
  n->size = b->off + b->size - at;
  n->used = b->used;
  n->next = b->next;
  b->size = at - b->off;
  b->next = n;
  return OSC_OK;
}

int osc_heap_alloc(size_t size, size_t align, size_t *off) {
  osc_block *b;
  if (align < OSC_MIN_ALIGN) align = OSC_MIN_ALIGN;
  if ((align & (align - 1)) != 0) return OSC_ERR_ARG;
  if (size == 0) size = 1;
  /* Every offset and block end lies in [0, heap_bytes]: a request larger
   * than the heap can never fit, and rejecting it here keeps the rounding
   * and the fit test below from wrapping (the reference's dlmalloc returns
   * NULL for such sizes as well). */
  if (size > osc_g.heap_bytes || size > ((size_t)-1) - OSC_MIN_ALIGN ||
      align > osc_g.heap_bytes)
    return OSC_ERR_NOMEM;
  size = round_up(size, OSC_MIN_ALIGN);
  for (b = osc_g.blocks; b != NULL; b = b->next) {
    size_t start, end;
    if (b->used) continue;
    start = round_up(b->off, align);
    end = b->off + b->size;
    if (start < b->off || start > end || size > end - start) continue;
    if (split_at(b, start) != OSC_OK) return OSC_ERR_NOMEM;
    if (start != b->off) b = b->next; /* the aligned tail */
    if (split_at(b, start + size) != OSC_OK) return OSC_ERR_NOMEM;
    b->used = 1;
    *off = b->off;
    return OSC_OK;
  }
  return OSC_ERR_NOMEM;
}

int osc_heap_free(size_t off) {
  osc_block *b, *prev = NULL;
  for (b = osc_g.blocks; b != NULL; prev = b, b = b->next) {
    if (b->off != off) continue;
    if (!b->used) return OSC_ERR_PTR;
    b->used = 0;
    /* merge with the following free block */
    if (b->next != NULL && !b->next->used) {
      osc_block *n = b->next;
      b->size += n->size;
      b->next = n->next;
      free(n);
    }
    /* merge into the preceding free block */
    if (prev != NULL && !prev->used) {
      prev->size += b->size;
      prev->next = b->next;
      free(b);
    }
    return OSC_OK;
  }
  return OSC_ERR_PTR;
}

int osc_heap_block_size(size_t off, size_t *size) {
  osc_block *b;
  for (b = osc_g.blocks; b != NULL; b = b->next)
    if (b->off == off && b->used) {
      *size = b->size;
      return OSC_OK;
    }
  return OSC_ERR_PTR;
}

size_t osc_heap_used(void) {
  size_t u = 0;
  osc_block *b;
  for (b = osc_g.blocks; b != NULL; b = b->next)
    if (b->used) u += b->size;
  return u;
}
