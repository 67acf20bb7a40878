/* The C11 type-generic routines of <shmem.h>.
 *
 * uses(): every generic with element types of its table; built with
 * -Werror, so a selection whose parameter types do not match the arguments
 * (an incompatible pointer) fails the build.  It is compiled and linked,
 * never run (it would need a GPU).
 * main(): the routine each type selects, compared by address, for the
 * canonical C types and the <stdint.h>/<stddef.h> typedefs. */
#include <shmem.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

typedef void (*fn_t)(void);
static int bad;

static void same(fn_t got, fn_t want, const char *what) {
  if (got != want) {
    printf("%s: selected the wrong routine\n", what);
    ++bad;
  }
}
#define SAME(sel, want) same((fn_t)(sel), (fn_t)(want), #sel " -> " #want)

void uses(shmem_team_t t, int pe) {
  long *ld = 0;
  const long *ls = 0;
  uint64_t *sig = 0;
  long pwrk[8];
  long psync[SHMEM_REDUCE_SYNC_SIZE];
  long v;
  shmem_put(ld, ls, 4, pe);
  shmem_get(ld, ls, 4, pe);
  shmem_put_nbi(ld, ls, 4, pe);
  shmem_get_nbi(ld, ls, 4, pe);
  shmem_p(ld, 7L, pe);
  v = shmem_g(ls, pe);
  (void)v;
  shmem_iput(ld, ls, 2, 1, 4, pe);
  shmem_iget(ld, ls, 1, 2, 4, pe);
  shmem_put_signal(ld, ls, 4, sig, 1, SHMEM_SIGNAL_SET, pe);
  shmem_put_signal_nbi(ld, ls, 4, sig, 1, SHMEM_SIGNAL_ADD, pe);
  shmem_sum_reduce(t, ld, ls, 4);
  shmem_prod_reduce(t, ld, ls, 4);
  shmem_min_reduce(t, ld, ls, 4);
  shmem_max_reduce(t, ld, ls, 4);
  shmem_and_reduce(t, ld, ls, 4);
  shmem_or_reduce(t, (unsigned *)0, (const unsigned *)0, 4);
  shmem_xor_reduce(t, (uint64_t *)0, (const uint64_t *)0, 4);
  shmem_and_reduce(t, (int8_t *)0, (const int8_t *)0, 4);
  shmem_sum_reduce(t, (float *)0, (const float *)0, 4);
  shmem_sum_reduce(t, (double _Complex *)0, (const double _Complex *)0, 4);
  shmem_prod_reduce(t, (float _Complex *)0, (const float _Complex *)0, 4);
  shmem_max_reduce(t, (size_t *)0, (const size_t *)0, 4);
  shmem_min_reduce(t, (ptrdiff_t *)0, (const ptrdiff_t *)0, 4);
  shmem_sum_reduce(t, (char *)0, (const char *)0, 4);
  shmem_fcollect(t, (int32_t *)0, (const int32_t *)0, 4);
  shmem_collect(t, (double *)0, (const double *)0, 4);
  shmem_alltoall(t, (int64_t *)0, (const int64_t *)0, 4);
  shmem_alltoalls(t, (short *)0, (const short *)0, 2, 3, 4);
  shmem_broadcast(t, (long double *)0, (const long double *)0, 4, 0);
  shmem_and_to_all((short *)0, (const short *)0, 4, 0, 0, 2, (short *)0, psync);
  shmem_or_to_all(ld, ls, 4, 0, 0, 2, pwrk, psync);
  shmem_xor_to_all((long long *)0, (const long long *)0, 4, 0, 0, 2,
                   (long long *)0, psync);
  shmem_min_to_all((float *)0, (const float *)0, 4, 0, 0, 2, (float *)0, psync);
  shmem_max_to_all((int *)0, (const int *)0, 4, 0, 0, 2, (int *)0, psync);
  shmem_sum_to_all((double _Complex *)0, (const double _Complex *)0, 4, 0, 0,
                   2, (double _Complex *)0, psync);
  shmem_prod_to_all(ld, ls, 4, 0, 0, 2, pwrk, psync);
  {
    int st[4] = {0, 0, 0, 0};
    size_t idx[4];
    long cv[4] = {1, 2, 3, 4};
    unsigned short *us = 0;
    size_t k = 0;
    k += (size_t)shmem_test(ld, SHMEM_CMP_EQ, 1L);
    shmem_wait_until((int *)0, SHMEM_CMP_GE, 1);
    shmem_wait(us, (unsigned short)0);
    k += (size_t)shmem_test_all(ld, 4, st, SHMEM_CMP_NE, 0L);
    k += shmem_test_any((uint64_t *)0, 4, st, SHMEM_CMP_GT, (uint64_t)0);
    k += shmem_test_some((size_t *)0, 4, idx, st, SHMEM_CMP_LT, (size_t)9);
    shmem_wait_until_all(ld, 4, NULL, SHMEM_CMP_EQ, 1L);
    k += shmem_wait_until_any(ld, 4, st, SHMEM_CMP_EQ, 1L);
    k += shmem_wait_until_some(ld, 4, idx, st, SHMEM_CMP_EQ, 1L);
    k += (size_t)shmem_test_all_vector(ld, 4, st, SHMEM_CMP_LE, cv);
    k += shmem_test_any_vector(ld, 4, st, SHMEM_CMP_LE, cv);
    k += shmem_test_some_vector(ld, 4, idx, st, SHMEM_CMP_LE, cv);
    shmem_wait_until_all_vector(ld, 4, st, SHMEM_CMP_GE, cv);
    k += shmem_wait_until_any_vector(ld, 4, st, SHMEM_CMP_GE, cv);
    k += shmem_wait_until_some_vector(ld, 4, idx, st, SHMEM_CMP_GE, cv);
    (void)k;
  }
  {
    long lf = 0, *lp = 0;
    float ff = 0, *fp = 0;
    unsigned *up = 0;
    long r = 0;
    r += shmem_atomic_fetch(ls, pe);
    shmem_atomic_set(fp, 1.0f, pe);
    ff = shmem_atomic_swap(fp, ff, pe);
    shmem_atomic_fetch_nbi(&ff, fp, pe);
    shmem_atomic_swap_nbi(&lf, lp, 2L, pe);
    r += shmem_atomic_compare_swap(lp, 1L, 2L, pe);
    r += shmem_atomic_fetch_inc(lp, pe);
    shmem_atomic_inc(lp, pe);
    r += shmem_atomic_fetch_add(lp, 3L, pe);
    shmem_atomic_add(lp, 3L, pe);
    shmem_atomic_compare_swap_nbi(&lf, lp, 1L, 2L, pe);
    shmem_atomic_fetch_inc_nbi(&lf, lp, pe);
    shmem_atomic_fetch_add_nbi(&lf, lp, 4L, pe);
    r += (long)shmem_atomic_fetch_and(up, 1u, pe);
    r += (long)shmem_atomic_fetch_or(up, 1u, pe);
    r += (long)shmem_atomic_fetch_xor(up, 1u, pe);
    shmem_atomic_and(lp, 1L, pe);
    shmem_atomic_or(up, 1u, pe);
    shmem_atomic_xor(up, 1u, pe);
    shmem_atomic_fetch_and_nbi((unsigned *)&lf, up, 1u, pe);
    shmem_atomic_fetch_or_nbi(&lf, lp, 1L, pe);
    shmem_atomic_fetch_xor_nbi(&lf, lp, 1L, pe);
    (void)r;
    (void)ff;
  }
  { /* the context forms: one more leading argument */
    shmem_ctx_t ctx = SHMEM_CTX_DEFAULT;
    long *lp = 0, lf = 0;
    uint64_t *sig = 0;
    shmem_put(ctx, lp, ls, 4, pe);
    shmem_get(ctx, lp, ls, 4, pe);
    shmem_put_nbi(ctx, lp, ls, 4, pe);
    shmem_get_nbi(ctx, lp, ls, 4, pe);
    shmem_p(ctx, lp, 3L, pe);
    lf = shmem_g(ctx, ls, pe);
    shmem_iput(ctx, lp, ls, 2, 1, 4, pe);
    shmem_iget(ctx, lp, ls, 1, 2, 4, pe);
    shmem_put_signal(ctx, lp, ls, 4, sig, 1, SHMEM_SIGNAL_SET, pe);
    shmem_put_signal_nbi(ctx, lp, ls, 4, sig, 1, SHMEM_SIGNAL_ADD, pe);
    lf += shmem_atomic_fetch(ctx, ls, pe);
    lf += shmem_atomic_fetch_inc(ctx, lp, pe);
    shmem_atomic_add(ctx, lp, 2L, pe);
    lf += shmem_atomic_compare_swap(ctx, lp, 1L, 2L, pe);
    shmem_atomic_fetch_add_nbi(ctx, &lf, lp, 2L, pe);
    shmem_atomic_fetch_xor_nbi(ctx, &lf, lp, 2L, pe);
    shmem_atomic_xor_nbi(&lf, lp, 2L, pe);
    shmem_atomic_xor_nbi(ctx, (uint64_t *)&lf, (uint64_t *)lp, (uint64_t)2, pe);
    (void)lf;
  }
}

int main(void) {
  int *pi = 0;
  long *pl = 0;
  const float *cf = 0;
  double _Complex *pz = 0;
  float _Complex *pc = 0;
  long double *pq = 0;
  int8_t *p8 = 0;
  int16_t *p16 = 0;
  int32_t *p32 = 0;
  int64_t *p64 = 0;
  uint8_t *u8 = 0;
  uint16_t *u16 = 0;
  uint32_t *u32 = 0;
  uint64_t *u64 = 0;
  size_t *psz = 0;
  ptrdiff_t *ppd = 0;
  char *pch = 0;
  signed char *psc = 0;
  unsigned long long *pull = 0;
  long long *pll = 0;
  /* canonical types */
  SAME(OSC_GEN_STD(pi, put), shmem_int_put);
  SAME(OSC_GEN_STD(pl, get), shmem_long_get);
  SAME(OSC_GEN_STD(cf, g), shmem_float_g); /* const source */
  SAME(OSC_GEN_STD(pq, broadcast), shmem_longdouble_broadcast);
  SAME(OSC_GEN_STD(pch, fcollect), shmem_char_fcollect);
  SAME(OSC_GEN_STD(psc, alltoall), shmem_schar_alltoall);
  SAME(OSC_GEN_STD(pll, min_reduce), shmem_longlong_min_reduce);
  SAME(OSC_GEN_STD(pull, max_reduce), shmem_ulonglong_max_reduce);
  SAME(OSC_GEN_ARITH(pi, sum_reduce), shmem_int_sum_reduce);
  SAME(OSC_GEN_ARITH(pz, prod_reduce), shmem_complexd_prod_reduce);
  SAME(OSC_GEN_ARITH(pc, sum_reduce), shmem_complexf_sum_reduce);
  SAME(OSC_GEN_BITWISE(pi, and_reduce), shmem_int32_and_reduce);
  SAME(OSC_GEN_BITWISE(pl, xor_reduce), shmem_int64_xor_reduce);
  SAME(OSC_GEN_BITWISE(pull, or_reduce), shmem_ulonglong_or_reduce);
  SAME(OSC_GEN_TO_ALL_ARITH(pz, sum_to_all), shmem_complexd_sum_to_all);
  SAME(OSC_GEN_TO_ALL_BITWISE(pll, and_to_all), shmem_longlong_and_to_all);
  /* typedefs select the routine of their C type (LP64 Linux) */
  SAME(OSC_GEN_BITWISE(p8, and_reduce), shmem_int8_and_reduce);
  SAME(OSC_GEN_BITWISE(p16, or_reduce), shmem_int16_or_reduce);
  SAME(OSC_GEN_BITWISE(p32, xor_reduce), shmem_int32_xor_reduce);
  SAME(OSC_GEN_BITWISE(p64, and_reduce), shmem_int64_and_reduce);
  SAME(OSC_GEN_BITWISE(u8, and_reduce), shmem_uchar_and_reduce);
  SAME(OSC_GEN_BITWISE(u16, and_reduce), shmem_ushort_and_reduce);
  SAME(OSC_GEN_BITWISE(u32, and_reduce), shmem_uint_and_reduce);
  SAME(OSC_GEN_BITWISE(u64, xor_reduce), shmem_ulong_xor_reduce);
  SAME(OSC_GEN_BITWISE(psz, or_reduce), shmem_ulong_or_reduce);
  SAME(OSC_GEN_ARITH(p64, sum_reduce), shmem_long_sum_reduce);
  SAME(OSC_GEN_ARITH(ppd, sum_reduce), shmem_long_sum_reduce);
  SAME(OSC_GEN_STD(u64, put), shmem_ulong_put);
  SAME(OSC_GEN_STD(p32, fcollect), shmem_int_fcollect);
  SAME(OSC_GEN_SYNC(u64, test_any), shmem_ulong_test_any);
  SAME(OSC_GEN_SYNC(psz, wait_until_all), shmem_ulong_wait_until_all);
  SAME(OSC_GEN_SYNC1(p16, test), shmem_short_test);
  SAME(OSC_GEN_SYNC1(u16, wait), shmem_ushort_wait);
  SAME(OSC_GEN_AMO_EXT(cf, atomic_fetch), shmem_float_atomic_fetch);
  SAME(OSC_GEN_AMO_STD(p64, atomic_fetch_inc), shmem_long_atomic_fetch_inc);
  SAME(OSC_GEN_AMO_STD(psz, atomic_add), shmem_ulong_atomic_add);
  SAME(OSC_GEN_AMO_BITWISE(p32, atomic_fetch_xor), shmem_int32_atomic_fetch_xor);
  SAME(OSC_GEN_AMO_BITWISE(u64, atomic_or), shmem_ulong_atomic_or);
  SAME(OSC_GEN_STD_P(shmem_ctx_, p64, put), shmem_ctx_long_put);
  SAME(OSC_GEN_AMO_STD_P(shmem_ctx_, psz, atomic_fetch_inc),
       shmem_ctx_ulong_atomic_fetch_inc);
  SAME(OSC_GEN_AMO_EXT_P(shmem_ctx_, cf, atomic_fetch),
       shmem_ctx_float_atomic_fetch);
  if (bad == 0) printf("generics ok\n");
  return bad != 0;
}
