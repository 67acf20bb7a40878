/*
 * spmd_collectives.c - an SPMD OpenSHMEM program: every PE runs main() and
 * calls each collective with its own arguments, as with the reference's
 * oshrun.  Built with -Dmain=osc_pe_main against libosc_spmd_main.a, which
 * runs one host thread per PE.  Symmetric memory is device memory, so
 * host values move in and out with shmem_putmem/getmem on the PE's own
 * copy.  Each check prints on failure; the PE's exit status is the number
 * of failures.
 */
#include <shmem.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define N 4096

static int check(int ok, int me, const char *what) {
  if (!ok) printf("PE %d FAIL: %s\n", me, what);
  return ok ? 0 : 1;
}

int main(void) {
  int fails = 0, me, npes, i, k;
  long *src, *dst, *slot, *data;
  uint64_t *sig;
  long *host, *out;
  int *isrc, *idst, ihost[64], iout[64];
  shmem_team_t even = SHMEM_TEAM_INVALID;

  int level = -1;
  if (shmem_init_thread(SHMEM_THREAD_FUNNELED, &level) != 0) return 1;
  me = shmem_my_pe();
  npes = shmem_n_pes();
  fails += check(level == SHMEM_THREAD_FUNNELED && shmem_initialized() &&
                     !shmem_finalized() && shmem_pe_accessible(npes - 1) &&
                     !shmem_pe_accessible(npes),
                 me, "init_thread / initialized / pe_accessible");
  host = (long *)malloc(sizeof(long) * N * (size_t)npes);
  out = (long *)malloc(sizeof(long) * N * (size_t)npes);
  src = (long *)shmem_malloc(sizeof(long) * N * (size_t)npes);
  dst = (long *)shmem_malloc(sizeof(long) * N * (size_t)npes);

  /* 1. sum reduce over the world */
  for (i = 0; i < N; ++i) host[i] = (long)me * 1000 + i;
  shmem_putmem(src, host, sizeof(long) * N, me);
  shmem_long_sum_reduce(SHMEM_TEAM_WORLD, dst, src, N);
  shmem_getmem(out, dst, sizeof(long) * N, me);
  for (i = 0, k = 0; i < N; ++i)
    k |= out[i] != 1000L * npes * (npes - 1) / 2 + (long)npes * i;
  fails += check(!k, me, "long sum reduce");

  /* 2. max reduce, int */
  isrc = (int *)shmem_malloc(sizeof(int) * 64);
  idst = (int *)shmem_malloc(sizeof(int) * 64);
  for (i = 0; i < 64; ++i) ihost[i] = (i % npes == me) ? 1000 + i : -i;
  shmem_putmem(isrc, ihost, sizeof(ihost), me);
  shmem_max_reduce(SHMEM_TEAM_WORLD, idst, isrc, 64); /* C11 generic */
  shmem_getmem(iout, idst, sizeof(iout), me);
  for (i = 0, k = 0; i < 64; ++i) k |= iout[i] != 1000 + i;
  fails += check(!k, me, "int max reduce");

  /* 3. fcollect: block of PE p holds p*N + i */
  for (i = 0; i < N; ++i) host[i] = (long)me * N + i;
  shmem_putmem(src, host, sizeof(long) * N, me);
  shmem_long_fcollect(SHMEM_TEAM_WORLD, dst, src, N);
  shmem_getmem(out, dst, sizeof(long) * N * npes, me);
  for (i = 0, k = 0; i < N * npes; ++i) k |= out[i] != i;
  fails += check(!k, me, "fcollect");

  /* 4. collect: PE p brings p + 1 elements of value p */
  for (i = 0; i <= me; ++i) host[i] = me;
  shmem_putmem(src, host, sizeof(long) * (me + 1), me);
  shmem_long_collect(SHMEM_TEAM_WORLD, dst, src, (size_t)me + 1);
  shmem_getmem(out, dst, sizeof(long) * npes * (npes + 1) / 2, me);
  for (i = 0, k = 0; i < npes; ++i) {
    int j, base = i * (i + 1) / 2;
    for (j = 0; j <= i; ++j) k |= out[base + j] != i;
  }
  fails += check(!k, me, "collect (variable sizes)");

  /* 5. alltoall: block j of PE p's source is p*100 + j */
  for (i = 0; i < npes; ++i)
    for (k = 0; k < 8; ++k) host[i * 8 + k] = (long)me * 100 + i;
  shmem_putmem(src, host, sizeof(long) * 8 * npes, me);
  shmem_alltoall(SHMEM_TEAM_WORLD, dst, src, 8); /* C11 generic */
  shmem_getmem(out, dst, sizeof(long) * 8 * npes, me);
  for (i = 0, k = 0; i < npes * 8; ++i) k |= out[i] != (long)(i / 8) * 100 + me;
  fails += check(!k, me, "alltoall");

  /* 6. broadcast from the last PE */
  for (i = 0; i < N; ++i) host[i] = (long)me * 7 + i;
  shmem_putmem(src, host, sizeof(long) * N, me);
  shmem_long_broadcast(SHMEM_TEAM_WORLD, dst, src, N, npes - 1);
  shmem_getmem(out, dst, sizeof(long) * N, me);
  for (i = 0, k = 0; i < N; ++i) k |= out[i] != (long)(npes - 1) * 7 + i;
  fails += check(!k, me, "broadcast");

  /* 7. split the even PEs off and reduce among them */
  shmem_team_split_strided(SHMEM_TEAM_WORLD, 0, 2, (npes + 1) / 2, NULL, 0,
                           &even);
  fails += check((even != SHMEM_TEAM_INVALID) == (me % 2 == 0), me,
                 "split membership");
  if (even != SHMEM_TEAM_INVALID) {
    for (i = 0; i < 64; ++i) ihost[i] = me + i;
    shmem_putmem(isrc, ihost, sizeof(ihost), me);
    shmem_int_sum_reduce(even, idst, isrc, 64);
    shmem_getmem(iout, idst, sizeof(iout), me);
    for (i = 0, k = 0; i < 64; ++i) {
      int want = 0, p;
      for (p = 0; p < npes; p += 2) want += p + i;
      k |= iout[i] != want;
    }
    fails += check(!k, me, "reduce on the split team");
    fails += check(shmem_team_my_pe(even) == me / 2, me, "team rank");
    shmem_team_destroy(even);
  }

  /* 8. active-set form over all PEs */
  {
    static long pSync[SHMEM_REDUCE_SYNC_SIZE];
    static long pWrk[SHMEM_REDUCE_MIN_WRKDATA_SIZE];
    for (i = 0; i < N; ++i) host[i] = (i % npes == me) ? -5 : i;
    shmem_putmem(src, host, sizeof(long) * N, me);
    shmem_min_to_all(dst, src, N, 0, 0, npes, pWrk, pSync); /* C11 generic */
    shmem_getmem(out, dst, sizeof(long) * N, me);
    for (i = 0, k = 0; i < N; ++i) k |= out[i] != -5;
    fails += check(!k, me, "long min to_all");
  }

  /* 9. put ring: write my PE number into the next PE's slot */
  slot = (long *)shmem_calloc(1, sizeof(long));
  {
    long v = me;
    shmem_long_p(slot, v, (me + 1) % npes);
  }
  shmem_barrier_all();
  fails += check(shmem_long_g(slot, me) == (me + npes - 1) % npes, me,
                 "put ring");

  /* 10. put with signal: the next PE waits for the data's signal */
  data = (long *)shmem_malloc(sizeof(long) * 16);
  sig = (uint64_t *)shmem_calloc(1, sizeof(uint64_t));
  for (i = 0; i < 16; ++i) host[i] = (long)me * 16 + i;
  shmem_long_put_signal(data, host, 16, sig, 1, SHMEM_SIGNAL_ADD,
                        (me + 1) % npes);
  shmem_signal_wait_until(sig, SHMEM_CMP_GE, 1);
  shmem_getmem(out, data, sizeof(long) * 16, me);
  for (i = 0, k = 0; i < 16; ++i)
    k |= out[i] != (long)((me + npes - 1) % npes) * 16 + i;
  fails += check(!k, me, "put_signal ring");

  shmem_barrier_all();
  shmem_free(sig);
  shmem_free(data);
  shmem_free(slot);
  shmem_free(idst);
  shmem_free(isrc);
  shmem_free(dst);
  shmem_free(src);
  free(host);
  free(out);
  shmem_finalize();
  fails += check(shmem_finalized() && !shmem_initialized(), me, "finalized");
  printf("PE %d of %d: %s\n", me, npes, fails ? "FAIL" : "PASS");
  return fails;
}
