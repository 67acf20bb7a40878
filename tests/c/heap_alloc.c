/*
 * Host-only check of the symmetric offset allocator (osc_heap.c): sizes
 * near SIZE_MAX must fail without changing the block list, and later
 * allocations must not alias live objects.  Built and run by
 * tests/test_heap_host.py; prints "ok" and exits 0 on success.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "osc_internal.h"

osc_state osc_g;

static int fails;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++fails;                                                \
    }                                                         \
  } while (0)

static int nblocks(void) {
  int n = 0;
  for (osc_block *b = osc_g.blocks; b != NULL; b = b->next) ++n;
  return n;
}

int main(void) {
  const size_t heap = (size_t)1 << 20;
  size_t off = 12345, a, b, c, used;
  osc_g.heap_bytes = heap;
  CHECK(osc_heap_init(heap) == OSC_OK);

  /* empty heap: huge sizes fail, nothing changes */
  CHECK(osc_heap_alloc(SIZE_MAX, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_alloc(SIZE_MAX - 10, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_alloc(SIZE_MAX - 255, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_alloc(heap + 1, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_alloc(16, (size_t)1 << 62, &off) == OSC_ERR_NOMEM);
  CHECK(off == 12345);
  CHECK(nblocks() == 1 && osc_heap_used() == 0);

  /* after one small object, a near-SIZE_MAX request still fails and the
   * next allocation does not alias the live object */
  CHECK(osc_heap_alloc(300, 0, &a) == OSC_OK && a == 0);
  used = osc_heap_used();
  CHECK(osc_heap_alloc(SIZE_MAX - 600, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_alloc(SIZE_MAX - 600, 4096, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_used() == used && nblocks() == 2);
  CHECK(osc_heap_alloc(4096, 0, &b) == OSC_OK && b == 512);

  /* the rest of the heap fits exactly once */
  CHECK(osc_heap_alloc(heap - 512 - 4096, 0, &c) == OSC_OK && c == 4608);
  CHECK(osc_heap_alloc(1, 0, &off) == OSC_ERR_NOMEM);
  CHECK(osc_heap_used() == heap);

  /* free and coalesce back to one block; the whole heap is one object */
  CHECK(osc_heap_free(b) == OSC_OK);
  CHECK(osc_heap_free(a) == OSC_OK);
  CHECK(osc_heap_free(c) == OSC_OK);
  CHECK(nblocks() == 1 && osc_heap_used() == 0);
  CHECK(osc_heap_alloc(heap, 0, &a) == OSC_OK && a == 0);
  CHECK(osc_heap_free(a) == OSC_OK);
  osc_heap_fini();
  if (fails == 0) printf("ok\n");
  return fails != 0;
}
