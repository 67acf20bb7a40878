/* latency.c - wall time of blocking team reductions, measured from C.
 *
 *   bin/latency            # one JSON line on stdout
 *
 * The calls are the plain OpenSHMEM entry points, as a C program makes
 * them: shmem_int_sum_reduce on 8 PEs x 65,536 int32 (values in
 * [-1000, 1000]), and shmem_long_sum_reduce / shmem_double_sum_reduce of 16
 * elements on teams of 64 and 128 PEs.  Each is called 200 times to warm
 * up, then 2000 times, each call timed with CLOCK_MONOTONIC; p50, p90 and
 * p99 are reported in microseconds, with the first result checked against
 * a sum on the host. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "osc.h"
#include "shmem.h"

#define WARM 200
#define REPS 2000

static double now_us(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec * 1e6 + (double)t.tv_nsec / 1e3;
}

static int cmp(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}

static double v[REPS];

static void emit(const char *name, int last) {
  qsort(v, REPS, sizeof(double), cmp);
  printf("\"%s\": {\"p50_us\": %.2f, \"p90_us\": %.2f, \"p99_us\": %.2f}%s",
         name, v[REPS / 2], v[REPS * 9 / 10], v[REPS * 99 / 100],
         last ? "" : ", ");
}

static int fail(const char *what) {
  fprintf(stderr, "latency: %s: %s\n", what, osc_last_error());
  return 1;
}

int main(void) {
  const size_t n = 65536, m = 16;
  int i, k;
  /* 8 PEs x 65536 int32 */
  if (osc_init(8, 4 << 20, 0) != 0) return fail("osc_init(8)");
  {
    int *s = shmem_malloc(n * sizeof(int)), *d = shmem_malloc(n * sizeof(int));
    int *h = malloc(8 * n * sizeof(int)), *got = malloc(n * sizeof(int));
    if (!s || !d || !h || !got) return fail("allocation");
    for (k = 0; k < 8; ++k)
      for (i = 0; i < (int)n; ++i)
        h[k * n + i] = (int)(((unsigned)k * 7919u + (unsigned)i * 104729u) % 2001u) - 1000;
    for (k = 0; k < 8; ++k) shmem_putmem(s, h + k * n, n * sizeof(int), k);
    for (i = 0; i < WARM; ++i) shmem_int_sum_reduce(SHMEM_TEAM_WORLD, d, s, n);
    for (i = 0; i < REPS; ++i) {
      const double t0 = now_us();
      if (shmem_int_sum_reduce(SHMEM_TEAM_WORLD, d, s, n) != 0)
        return fail("shmem_int_sum_reduce");
      v[i] = now_us() - t0;
    }
    shmem_getmem(got, d, n * sizeof(int), 5);
    for (i = 0; i < (int)n; ++i) {
      int want = 0;
      for (k = 0; k < 8; ++k) want += h[k * n + i];
      if (got[i] != want) {
        fprintf(stderr, "latency: wrong int sum at %d\n", i);
        return 1;
      }
    }
    printf("{");
    emit("int_sum_8pe_64K", 0);
    free(h);
    free(got);
  }
  osc_finalize();
  /* 16 elements on 128 PEs and on a team of the first 64 */
  if (osc_init(128, 1 << 20, 0) != 0) return fail("osc_init(128)");
  {
    long *sl = shmem_malloc(m * sizeof(long)), *dl = shmem_malloc(m * sizeof(long));
    double *sd = shmem_malloc(m * sizeof(double)), *dd = shmem_malloc(m * sizeof(double));
    shmem_team_t t64;
    long hl[16];
    double hd[16];
    int size;
    if (!sl || !dl || !sd || !dd) return fail("allocation");
    if (shmem_team_split_strided(SHMEM_TEAM_WORLD, 0, 1, 64, NULL, 0, &t64) != 0)
      return fail("team split");
    for (k = 0; k < 128; ++k) {
      for (i = 0; i < (int)m; ++i) {
        hl[i] = (long)((k * 31 + i * 17) % 2001) - 1000;
        hd[i] = (double)hl[i] * 0.5;
      }
      shmem_putmem(sl, hl, sizeof hl, k);
      shmem_putmem(sd, hd, sizeof hd, k);
    }
    for (size = 64; size <= 128; size *= 2) {
      shmem_team_t t = size == 64 ? t64 : SHMEM_TEAM_WORLD;
      char name[64];
      for (i = 0; i < WARM; ++i) shmem_long_sum_reduce(t, dl, sl, m);
      for (i = 0; i < REPS; ++i) {
        const double t0 = now_us();
        if (shmem_long_sum_reduce(t, dl, sl, m) != 0)
          return fail("shmem_long_sum_reduce");
        v[i] = now_us() - t0;
      }
      snprintf(name, sizeof name, "long_sum_%dpe_16", size);
      emit(name, 0);
      for (i = 0; i < WARM; ++i) shmem_double_sum_reduce(t, dd, sd, m);
      for (i = 0; i < REPS; ++i) {
        const double t0 = now_us();
        if (shmem_double_sum_reduce(t, dd, sd, m) != 0)
          return fail("shmem_double_sum_reduce");
        v[i] = now_us() - t0;
      }
      snprintf(name, sizeof name, "double_sum_%dpe_16", size);
      emit(name, size == 128);
    }
    {
      long got[16];
      shmem_getmem(got, dl, sizeof got, 3);
      for (i = 0; i < (int)m; ++i) {
        long want = 0;
        for (k = 0; k < 128; ++k) want += (long)((k * 31 + i * 17) % 2001) - 1000;
        if (got[i] != want) {
          fprintf(stderr, "latency: wrong long sum at %d\n", i);
          return 1;
        }
      }
    }
    printf("}\n");
  }
  osc_finalize();
  return 0;
}
