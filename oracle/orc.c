/*
 * orc.c - CPU oracle for the collective hot path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and the benchmark's CPU baseline, never
 * by the product library.
 *
 * Each PE's memory is a separate host buffer; the algorithms below perform,
 * for all PEs, the same puts/gets and local combines, in the same order, as
 * the reference code cited at each function.  Built with -ffp-contract=off
 * on x86-64 (SSE2: no excess precision), as the reference is.
 */
#define _GNU_SOURCE
#include "orc.h"

#include <complex.h>
#include <math.h>
#include <sched.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/wait.h>
#include <time.h>
#include <unistd.h>

enum { AND, OR, XOR, MIN, MAX, SUM, PROD };
enum { U8, U16, U32, U64, I8, I16, I32, I64, F32, F64, C32, C64, NTYPES };

static size_t esize(int t) {
  static const size_t sz[NTYPES] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8, 8, 16};
  return (t < 0 || t >= NTYPES) ? 0 : sz[t];
}

/* ---- local_<T>_<op>_reduce: reduce.c:35-43 with the operators of
 *      reduce.c:730-736.  Integer + and * wrap (the reference's C signed
 *      arithmetic on x86-64 wraps in practice); WT is wide enough that the
 *      product of two promoted values cannot overflow a signed int. ---- */
#define DEF_LOCAL_INT(NAME, ST, UT, WT)                                        \
  static int local_##NAME(int op, void *d_, const void *a_, const void *b_,    \
                          size_t n) {                                          \
    ST *d = (ST *)d_;                                                          \
    const ST *a = (const ST *)a_, *b = (const ST *)b_;                         \
    size_t i;                                                                  \
    switch (op) {                                                              \
      case AND: for (i = 0; i < n; i++) d[i] = (ST)((UT)a[i] & (UT)b[i]); break; \
      case OR: for (i = 0; i < n; i++) d[i] = (ST)((UT)a[i] | (UT)b[i]); break;  \
      case XOR: for (i = 0; i < n; i++) d[i] = (ST)((UT)a[i] ^ (UT)b[i]); break; \
      case MIN: for (i = 0; i < n; i++) d[i] = (a[i] < b[i]) ? a[i] : b[i]; break; \
      case MAX: for (i = 0; i < n; i++) d[i] = (a[i] > b[i]) ? a[i] : b[i]; break; \
      case SUM:                                                                \
        for (i = 0; i < n; i++) d[i] = (ST)(UT)((WT)(UT)a[i] + (WT)(UT)b[i]);  \
        break;                                                                 \
      case PROD:                                                               \
        for (i = 0; i < n; i++) d[i] = (ST)(UT)((WT)(UT)a[i] * (WT)(UT)b[i]);  \
        break;                                                                 \
      default: return -1;                                                      \
    }                                                                          \
    return 0;                                                                  \
  }
DEF_LOCAL_INT(u8, uint8_t, uint8_t, unsigned)
DEF_LOCAL_INT(u16, uint16_t, uint16_t, unsigned)
DEF_LOCAL_INT(u32, uint32_t, uint32_t, uint32_t)
DEF_LOCAL_INT(u64, uint64_t, uint64_t, uint64_t)
DEF_LOCAL_INT(i8, int8_t, uint8_t, unsigned)
DEF_LOCAL_INT(i16, int16_t, uint16_t, unsigned)
DEF_LOCAL_INT(i32, int32_t, uint32_t, uint32_t)
DEF_LOCAL_INT(i64, int64_t, uint64_t, uint64_t)

#define DEF_LOCAL_REAL(NAME, T)                                                \
  static int local_##NAME(int op, void *d_, const void *a_, const void *b_,    \
                          size_t n) {                                          \
    T *d = (T *)d_;                                                            \
    const T *a = (const T *)a_, *b = (const T *)b_;                            \
    size_t i;                                                                  \
    switch (op) {                                                              \
      case MIN: for (i = 0; i < n; i++) d[i] = (a[i] < b[i]) ? a[i] : b[i]; break; \
      case MAX: for (i = 0; i < n; i++) d[i] = (a[i] > b[i]) ? a[i] : b[i]; break; \
      case SUM: for (i = 0; i < n; i++) d[i] = a[i] + b[i]; break;             \
      case PROD: for (i = 0; i < n; i++) d[i] = a[i] * b[i]; break;            \
      default: return -1;                                                      \
    }                                                                          \
    return 0;                                                                  \
  }
DEF_LOCAL_REAL(f32, float)
DEF_LOCAL_REAL(f64, double)

/* C99 complex arithmetic exactly as the reference compiles it */
#define DEF_LOCAL_CPLX(NAME, T)                                                \
  static int local_##NAME(int op, void *d_, const void *a_, const void *b_,    \
                          size_t n) {                                          \
    T *d = (T *)d_;                                                            \
    const T *a = (const T *)a_, *b = (const T *)b_;                            \
    size_t i;                                                                  \
    switch (op) {                                                              \
      case SUM: for (i = 0; i < n; i++) d[i] = a[i] + b[i]; break;             \
      case PROD: for (i = 0; i < n; i++) d[i] = a[i] * b[i]; break;            \
      default: return -1;                                                      \
    }                                                                          \
    return 0;                                                                  \
  }
DEF_LOCAL_CPLX(c32, float _Complex)
DEF_LOCAL_CPLX(c64, double _Complex)

typedef int (*local_fn)(int, void *, const void *, const void *, size_t);

static local_fn pick_local(int dtype) {
  static const local_fn f[NTYPES] = {local_u8,  local_u16, local_u32, local_u64,
                                     local_i8,  local_i16, local_i32, local_i64,
                                     local_f32, local_f64, local_c32, local_c64};
  return (dtype < 0 || dtype >= NTYPES) ? NULL : f[dtype];
}

/* ---- REDUCE_HELPER_REC_DBL, reduce.c:182-293, for all PEs of the active
 *      set at once.  Members of the power-of-two set are me_as with
 *      ceil(floor(me_as*p2s/P)*P/p2s) == me_as (:207-214). ---- */
static int p2s_of(int P) {
  int p2s;
  for (p2s = 1; p2s * 2 <= P; p2s *= 2) {
  }
  return p2s;
}

static int member_p2s(int me, int P, int p2s) {
  int m = me * p2s / P;
  return ((m * P + p2s - 1) / p2s == me) ? m : -1;
}

int orc_reduce_rec_dbl(int op, int dtype, int P, size_t n,
                       const void *const *src, void *const *dest) {
  const local_fn f = pick_local(dtype);
  const size_t nb = n * esize(dtype);
  int p2s, me, mask, rc = 0;
  int *mp;
  char **tmp;
  if (f == NULL || P < 1) return -1;
  p2s = p2s_of(P);
  mp = (int *)malloc(sizeof(int) * (size_t)P);
  tmp = (char **)calloc((size_t)P, sizeof(char *));
  if (mp == NULL || tmp == NULL) {
    free(mp);
    free(tmp);
    return -1;
  }
  for (me = 0; me < P; ++me) mp[me] = member_p2s(me, P, p2s);
  /* :218-244 tmp buffer; a member whose right neighbour is outside the
   * power-of-two set pulls that neighbour's source into its dest and
   * computes tmp = op(dest, source) = op(x[me+1], x[me]). */
  for (me = 0; me < P; ++me) {
    if (mp[me] < 0) continue;
    tmp[me] = (char *)malloc(nb ? nb : 1);
    if (tmp[me] == NULL) {
      rc = -1;
      goto out;
    }
    if ((me + 1) * p2s / P == mp[me]) {
      memcpy(dest[me], src[me + 1], nb);
      rc |= f(op, tmp[me], dest[me], src[me], n);
    } else {
      memcpy(tmp[me], src[me], nb);
    }
  }
  /* :250-272 butterfly: in round `mask` every member receives its
   * partner's tmp into dest (putmem) and then computes tmp = op(tmp, dest).
   * All puts of a round land before any combine of that round. */
  for (mask = 1; mask < p2s; mask <<= 1) {
    for (me = 0; me < P; ++me) {
      int peer_p2s, peer_as;
      if (mp[me] < 0) continue;
      peer_p2s = mp[me] ^ mask;
      peer_as = (peer_p2s * P + p2s - 1) / p2s;
      memcpy(dest[me], tmp[peer_as], nb);
    }
    for (me = 0; me < P; ++me)
      if (mp[me] >= 0) rc |= f(op, tmp[me], tmp[me], dest[me], n);
  }
  /* :274-288 final copy; the folded neighbour receives the member's dest */
  for (me = 0; me < P; ++me)
    if (mp[me] >= 0) memcpy(dest[me], tmp[me], nb);
  for (me = 0; me < P; ++me)
    if (mp[me] >= 0 && (me + 1) * p2s / P == mp[me])
      memcpy(dest[me + 1], dest[me], nb);
out:
  for (me = 0; me < P; ++me) free(tmp[me]);
  free(tmp);
  free(mp);
  return rc;
}

/* ---- REDUCE_HELPER_LINEAR, reduce.c:55-91: the root (active-set index
 *      0) copies its source to tmp, then for i = 1..P-1 gets PE i's source
 *      into its dest and computes tmp = op(tmp, dest); its dest = tmp is
 *      then broadcast (shcoll_broadcast8_linear) to every PE. ---- */
int orc_reduce_linear(int op, int dtype, int P, size_t n,
                      const void *const *src, void *const *dest) {
  const local_fn f = pick_local(dtype);
  const size_t nb = n * esize(dtype);
  char *tmp;
  int i, rc = 0;
  if (f == NULL || P < 1) return -1;
  tmp = (char *)malloc(nb ? nb : 1);
  if (tmp == NULL) return -1;
  memcpy(tmp, src[0], nb);
  for (i = 1; i < P; ++i) {
    memcpy(dest[0], src[i], nb);
    rc |= f(op, tmp, tmp, dest[0], n);
  }
  memcpy(dest[0], tmp, nb);
  for (i = 1; i < P; ++i) memcpy(dest[i], dest[0], nb);
  free(tmp);
  return rc;
}

/* ---- REDUCE_HELPER_BINOMIAL, reduce.c:102-171.  PE a copies its source
 *      to dest, then receives from each child a | mask (mask below a's
 *      lowest set bit, a | mask < P): tmp = dest; dest = the child's dest
 *      (getmem); dest = op(dest, tmp).  The reference takes the children
 *      in arrival order (the lowest bit of the arrived ones first); this
 *      restatement is the schedule in which every child has finished and
 *      reported before its parent's first poll, so the parent takes them in
 *      ascending order.  PEs are run from the highest index down, which
 *      finishes every child before its parent.  Root's dest is then
 *      broadcast (shcoll_broadcast8_binomial_tree). ---- */
int orc_reduce_binomial(int op, int dtype, int P, size_t n,
                        const void *const *src, void *const *dest) {
  const local_fn f = pick_local(dtype);
  const size_t nb = n * esize(dtype);
  char *tmp;
  int me, rc = 0;
  if (f == NULL || P < 1) return -1;
  tmp = (char *)malloc(nb ? nb : 1);
  if (tmp == NULL) return -1;
  for (me = P - 1; me >= 0; --me) {
    unsigned mask;
    long to_receive = 0;
    memcpy(dest[me], src[me], nb);
    for (mask = 1; !(me & mask) && ((me | mask) < (unsigned)P); mask <<= 1)
      to_receive |= (long)mask;
    while (to_receive != 0) {
      const long recv = to_receive & -to_receive; /* lowest pending child */
      memcpy(tmp, dest[me], nb);
      memcpy(dest[me], dest[me | (int)recv], nb);
      rc |= f(op, dest[me], dest[me], tmp, n);
      to_receive &= ~recv;
    }
  }
  for (me = 1; me < P; ++me) memcpy(dest[me], dest[0], nb);
  free(tmp);
  return rc;
}

/* bit reversal of the lowest nbits bits (bithacks.c:11 reverse_bits) */
static int rev_bits(int x, int nbits) {
  int r = 0, i;
  for (i = 0; i < nbits; ++i) r |= ((x >> i) & 1) << (nbits - 1 - i);
  return r;
}

/* ---- REDUCE_HELPER_RABENSEIFNER, reduce.c:306-511 (ring = 0), and
 *      REDUCE_HELPER_RABENSEIFNER2, reduce.c:521-724 (ring = 1), for all
 *      PEs at once:
 *  1. fold of non-power-of-two pairs (:362-404): the non-member u = m+1
 *     gets the member's upper half of source into its dest, computes
 *     op(x_m, x_u) there and puts it back to the member's dest; the
 *     member gets u's lower half of source into its dest and computes
 *     op(x_u, x_m); other members copy their source;
 *  2. reduce-scatter by recursive halving (:409-446): in the round of
 *     distance d a member keeps the lower half of its block range when
 *     bit d of its index is clear, the upper half otherwise, reads the
 *     partner's dest on that half and computes dest = op(dest, partner's);
 *  3. allgather: recursive doubling over the block ranges starting at
 *     block reverse_bits(k) (:452-480), or a ring in which member k sends
 *     block reverse_bits(k - i) to member k+1 in step i (:652-676);
 *  4. the member puts its whole dest to the folded neighbour (:482-494).
 *  Block b covers elements [b*n/p2s, (b+1)*n/p2s). ---- */
int orc_reduce_rabenseifner(int op, int dtype, int P, size_t n, int ring,
                            const void *const *src, void *const *dest) {
  const local_fn f = pick_local(dtype);
  const size_t es = esize(dtype), nb = n * es;
  int p2s, lg, me, d, i, rc = 0;
  int *mp, *bb, *be;
  char *tmp;
  if (f == NULL || P < 1) return -1;
  p2s = p2s_of(P);
  for (lg = 0; (1 << lg) < p2s; ++lg) {
  }
  mp = (int *)malloc(sizeof(int) * (size_t)P);
  bb = (int *)malloc(sizeof(int) * (size_t)P);
  be = (int *)malloc(sizeof(int) * (size_t)P);
  tmp = (char *)malloc((size_t)P * nb > 0 ? (size_t)P * nb : 1);
  if (mp == NULL || bb == NULL || be == NULL || tmp == NULL) {
    rc = -1;
    goto out;
  }
  for (me = 0; me < P; ++me) mp[me] = member_p2s(me, P, p2s);
  /* 1. fold */
  for (me = 0; me < P; ++me) {
    const size_t h = n / 2;
    if (mp[me] < 0) {
      const int m = me - 1;
      memcpy((char *)dest[me] + h * es, (const char *)src[m] + h * es,
             (n - h) * es);
      rc |= f(op, (char *)dest[me] + h * es, (char *)dest[me] + h * es,
              (const char *)src[me] + h * es, n - h);
      memcpy((char *)dest[m] + h * es, (char *)dest[me] + h * es,
             (n - h) * es);
    } else if ((me + 1) * p2s / P == mp[me]) {
      memcpy(dest[me], src[me + 1], h * es);
      rc |= f(op, dest[me], dest[me], src[me], h);
    } else {
      memcpy(dest[me], src[me], nb);
    }
  }
  /* 2. reduce-scatter; every read of a round precedes its combines */
  for (me = 0; me < P; ++me) {
    bb[me] = 0;
    be[me] = p2s;
  }
  for (d = 1; d < p2s; d <<= 1) {
    for (me = 0; me < P; ++me) {
      int peer;
      size_t o, o2;
      if (mp[me] < 0) continue;
      if ((mp[me] & d) == 0)
        be[me] = (bb[me] + be[me]) / 2;
      else
        bb[me] = (bb[me] + be[me]) / 2;
      peer = ((mp[me] ^ d) * P + p2s - 1) / p2s;
      o = ((size_t)bb[me] * n) / (size_t)p2s;
      o2 = ((size_t)be[me] * n) / (size_t)p2s;
      memcpy(tmp + (size_t)me * nb, (const char *)dest[peer] + o * es,
             (o2 - o) * es);
    }
    for (me = 0; me < P; ++me) {
      size_t o, o2;
      if (mp[me] < 0) continue;
      o = ((size_t)bb[me] * n) / (size_t)p2s;
      o2 = ((size_t)be[me] * n) / (size_t)p2s;
      rc |= f(op, (char *)dest[me] + o * es, (char *)dest[me] + o * es,
              tmp + (size_t)me * nb, o2 - o);
    }
  }
  /* 3. allgather of the owned blocks */
  if (!ring) {
    for (me = 0; me < P; ++me) {
      if (mp[me] < 0) continue;
      bb[me] = rev_bits(mp[me], lg);
      be[me] = bb[me] + 1;
    }
    for (d = p2s / 2; d > 0; d >>= 1) {
      for (me = 0; me < P; ++me) { /* puts of this round (disjoint) */
        int peer;
        size_t o, o2;
        if (mp[me] < 0) continue;
        peer = ((mp[me] ^ d) * P + p2s - 1) / p2s;
        o = ((size_t)bb[me] * n) / (size_t)p2s;
        o2 = ((size_t)be[me] * n) / (size_t)p2s;
        memcpy((char *)dest[peer] + o * es, (char *)dest[me] + o * es,
               (o2 - o) * es);
      }
      for (me = 0; me < P; ++me) {
        if (mp[me] < 0) continue;
        if ((mp[me] & d) == 0)
          be[me] += be[me] - bb[me];
        else
          bb[me] -= be[me] - bb[me];
      }
    }
  } else {
    for (i = 0; i < p2s; ++i)
      for (me = 0; me < P; ++me) {
        int peer, blk;
        size_t o, o2;
        if (mp[me] < 0) continue;
        peer = (((mp[me] + 1) % p2s) * P + p2s - 1) / p2s;
        blk = rev_bits((mp[me] - i + p2s) % p2s, lg);
        o = ((size_t)blk * n) / (size_t)p2s;
        o2 = ((size_t)(blk + 1) * n) / (size_t)p2s;
        memcpy((char *)dest[peer] + o * es, (char *)dest[me] + o * es,
               (o2 - o) * es);
      }
  }
  /* 4. the folded neighbour gets its member's result */
  for (me = 0; me < P; ++me)
    if (mp[me] >= 0 && (me + 1) * p2s / P == mp[me])
      memcpy(dest[me + 1], dest[me], nb);
out:
  free(mp);
  free(bb);
  free(be);
  free(tmp);
  return rc;
}

/* ---- high-precision references ------------------------------------------ */

/* Shewchuk's exact expansion sum with the correctly rounded final step, as
 * Python's math.fsum (J. R. Shewchuk, "Adaptive Precision Floating-Point
 * Arithmetic and Fast Robust Geometric Predicates", 1997). */
static double fsum_partials(const double *x, size_t m, double *p, size_t *np_) {
  size_t np = 0, i, j;
  for (i = 0; i < m; ++i) {
    double v = x[i];
    size_t k = 0;
    for (j = 0; j < np; ++j) {
      double y = p[j], hi, lo;
      if (fabs(v) < fabs(y)) {
        double t = v;
        v = y;
        y = t;
      }
      hi = v + y;
      lo = y - (hi - v);
      if (lo != 0.0) p[k++] = lo;
      v = hi;
    }
    p[k++] = v;
    np = k;
  }
  *np_ = np;
  {
    double hi = 0.0;
    if (np > 0) {
      size_t k = np;
      double lo = 0.0, yr;
      hi = p[--k];
      while (k > 0) {
        double v = hi, y = p[--k];
        hi = v + y;
        yr = hi - v;
        lo = y - yr;
        if (lo != 0.0) break;
      }
      /* round-half-even correction across the remaining partials */
      if (k > 0 && ((lo < 0 && p[k - 1] < 0) || (lo > 0 && p[k - 1] > 0))) {
        double y = lo * 2.0, v = hi + y;
        yr = v - hi;
        if (y == yr) hi = v;
      }
    }
    return hi;
  }
}

double orc_fsum(const double *x, size_t m) {
  double *p = (double *)malloc(sizeof(double) * (m + 1));
  size_t np;
  double r;
  if (p == NULL) return NAN;
  r = fsum_partials(x, m, p, &np);
  free(p);
  return r;
}

/* exact sum as hi (correctly rounded) + lo (rounded remainder) */
static void fsum_dd(double *x, size_t m, double *hi, double *lo) {
  double h = orc_fsum(x, m);
  x[m] = -h;
  *hi = h;
  *lo = orc_fsum(x, m + 1);
}

typedef struct {
  double hi, lo;
} dd_t;
static dd_t two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  dd_t r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static dd_t fast_two_sum(double a, double b) {
  double s = a + b;
  dd_t r = {s, b - (s - a)};
  return r;
}
static dd_t dd_mul_d(dd_t x, double y) {
  double p = x.hi * y;
  double e = fma(x.hi, y, -p);
  e = fma(x.lo, y, e);
  return fast_two_sum(p, e);
}
static dd_t dd_add(dd_t x, dd_t y) {
  dd_t s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

static double get_real(const void *p, int dtype, size_t i, int comp) {
  switch (dtype) {
    case F32: return (double)((const float *)p)[i];
    case F64: return ((const double *)p)[i];
    case C32: return (double)((const float *)p)[2 * i + comp];
    case C64: return ((const double *)p)[2 * i + comp];
  }
  return NAN;
}

int orc_reduce_accurate(int op, int dtype, int P, size_t n,
                        const void *const *src, double *hi, double *lo) {
  const int cplx = (dtype == C32 || dtype == C64);
  size_t i;
  int k;
  double *buf;
  if (dtype != F32 && dtype != F64 && !cplx) return -1;
  if (op != SUM && op != PROD) return -1;
  buf = (double *)malloc(sizeof(double) * ((size_t)P + 2));
  if (buf == NULL) return -1;
  for (i = 0; i < n; ++i) {
    if (op == SUM) {
      int comp;
      for (comp = 0; comp < (cplx ? 2 : 1); ++comp) {
        size_t o = cplx ? 2 * i + comp : i;
        for (k = 0; k < P; ++k) buf[k] = get_real(src[k], dtype, i, comp);
        fsum_dd(buf, (size_t)P, &hi[o], &lo[o]);
      }
    } else if (!cplx) {
      dd_t acc = {get_real(src[0], dtype, i, 0), 0.0};
      for (k = 1; k < P; ++k) acc = dd_mul_d(acc, get_real(src[k], dtype, i, 0));
      hi[i] = acc.hi;
      lo[i] = acc.lo;
    } else {
      dd_t re = {get_real(src[0], dtype, i, 0), 0.0};
      dd_t im = {get_real(src[0], dtype, i, 1), 0.0};
      for (k = 1; k < P; ++k) {
        const double xr = get_real(src[k], dtype, i, 0);
        const double xi = get_real(src[k], dtype, i, 1);
        dd_t a = dd_mul_d(re, xr), b = dd_mul_d(im, xi);
        dd_t c = dd_mul_d(re, xi), d = dd_mul_d(im, xr);
        b.hi = -b.hi;
        b.lo = -b.lo;
        re = dd_add(a, b);
        im = dd_add(c, d);
      }
      hi[2 * i] = re.hi;
      lo[2 * i] = re.lo;
      hi[2 * i + 1] = im.hi;
      lo[2 * i + 1] = im.lo;
    }
  }
  free(buf);
  return 0;
}

/* ---- data movement ------------------------------------------------------- */

/* util/rotate.c:22-54: right-rotate arr by dist bytes along gcd cycles */
void orc_rotate_inplace(char *arr, size_t size, size_t dist) {
  size_t i, groups, a, b;
  if (dist == 0) return;
  a = dist;
  b = size;
  for (;;) {
    size_t r = a % b;
    if (r == 0) break;
    a = b;
    b = r;
  }
  groups = b;
  for (i = 0; i < groups; i++) {
    char temp = arr[i];
    size_t j = i;
    for (;;) {
      size_t k = (j < dist) ? size - dist + j : j - dist;
      if (k == i) break;
      arr[j] = arr[k];
      j = k;
    }
    arr[j] = temp;
  }
}

/* fcollect.c:370-403 fcollect_helper_bruck_inplace for all PEs.  In round
 * `distance` PE me puts dest[0, to_send) into peer (me - distance) at
 * offset sent_bytes; the source and target ranges of one PE never overlap
 * (to_send <= sent_bytes), so a round's puts commute. */
int orc_fcollect_bruck(int P, size_t nbytes, const void *const *src,
                       void *const *dest) {
  const size_t total = (size_t)P * nbytes;
  size_t distance, sent = nbytes;
  int me;
  for (me = 0; me < P; ++me) memcpy(dest[me], src[me], nbytes);
  for (distance = 1; distance < (size_t)P; distance <<= 1) {
    const size_t to_send = (2 * sent <= total) ? sent : total - sent;
    for (me = 0; me < P; ++me) {
      const int peer = (int)(((size_t)me - distance + (size_t)P) % (size_t)P);
      memcpy((char *)dest[peer] + sent, dest[me], to_send);
    }
    sent += distance * nbytes;
  }
  for (me = 0; me < P; ++me)
    orc_rotate_inplace((char *)dest[me], total, (size_t)me * nbytes);
  return 0;
}

/* alltoall.c:75-110 with SHIFT_PEER (:206): own block by memcpy, then
 * P-1 puts src_me[block peer] -> dest_peer[block me], peer = me + i. */
int orc_alltoall_shift(int P, size_t nbytes, const void *const *src,
                       void *const *dest) {
  int me, i;
  for (me = 0; me < P; ++me) {
    memcpy((char *)dest[me] + (size_t)me * nbytes,
           (const char *)src[me] + (size_t)me * nbytes, nbytes);
    for (i = 1; i < P; ++i) {
      const int peer = (me + i) % P;
      memcpy((char *)dest[peer] + (size_t)me * nbytes,
             (const char *)src[me] + (size_t)peer * nbytes, nbytes);
    }
  }
  return 0;
}

/* broadcast.c:527-531 (root copies its source) + broadcast_helper_linear
 * (:58-71: every other PE gets from the root) */
int orc_broadcast_linear(int P, int root, size_t nbytes,
                         const void *const *src, void *const *dest) {
  int me;
  if (root < 0 || root >= P) return -1;
  memcpy(dest[root], src[root], nbytes);
  for (me = 0; me < P; ++me)
    if (me != root) memcpy(dest[me], src[root], nbytes);
  return 0;
}

/* util/rotate.c:56-64: right-rotate arr by dist bytes through a copy */
void orc_rotate(char *arr, size_t size, size_t dist) {
  char *tmp;
  if (dist == 0 || size == 0) return;
  tmp = (char *)malloc(dist);
  memcpy(tmp, arr + (size - dist), dist);
  memmove(arr + dist, arr, size - dist);
  memcpy(arr, tmp, dist);
  free(tmp);
}

/* collect.c:406-492 collect_helper_bruck for all PEs, with PE me
 * contributing nbytes[me].  The exclusive prefix sum (util/scan.c) gives
 * block_offset; the last PE's offset + size is the total every PE learns
 * (broadcast_size).  In the round with distance d, PE me learns the
 * received byte count of recv_from = me + d (block_sizes[round]), clips it
 * to the total, and gets that prefix of recv_from's dest into its own dest
 * at its received count.  A round's gets read [0, recv(peer)) of the peer
 * and write [recv(me), ...) of their own dest, so with the counts taken
 * before the round they commute.  Finally dest is rotated right by
 * block_offset (rotate(), not rotate_inplace). */
int orc_collect_bruck(int P, const size_t *nbytes, const void *const *src,
                      void *const *dest) {
  size_t *off = (size_t *)malloc(sizeof(size_t) * (size_t)P);
  size_t *recv = (size_t *)malloc(sizeof(size_t) * (size_t)P);
  size_t *take = (size_t *)malloc(sizeof(size_t) * (size_t)P);
  size_t total = 0, d;
  int me;
  if (off == NULL || recv == NULL || take == NULL) {
    free(off);
    free(recv);
    free(take);
    return -1;
  }
  for (me = 0; me < P; ++me) {
    off[me] = total;
    total += nbytes[me];
  }
  for (me = 0; me < P; ++me) {
    memcpy(dest[me], src[me], nbytes[me]);
    recv[me] = nbytes[me];
  }
  for (d = 1; d < (size_t)P; d <<= 1) {
    for (me = 0; me < P; ++me) {
      const int from = (int)(((size_t)me + d) % (size_t)P);
      size_t r = recv[from];
      take[me] = recv[me] + r < total ? r : total - recv[me];
    }
    for (me = 0; me < P; ++me) {
      const int from = (int)(((size_t)me + d) % (size_t)P);
      memcpy((char *)dest[me] + recv[me], dest[from], take[me]);
    }
    for (me = 0; me < P; ++me) recv[me] += take[me];
  }
  for (me = 0; me < P; ++me) orc_rotate((char *)dest[me], total, off[me]);
  free(off);
  free(recv);
  free(take);
  return 0;
}

/* alltoalls.c:54-90 alltoalls_helper_shift_exchange_barrier for all PEs:
 * the self block, then for i = 1..P-1 peer = me + i: element t of
 * src_me at (peer*n + t)*sst goes to dest_peer at (me*n + t)*dst. */
int orc_alltoalls_shift(int P, size_t es, size_t dst, size_t sst, size_t n,
                        const void *const *src, void *const *dest) {
  int me, i;
  size_t t;
  for (me = 0; me < P; ++me) {
    for (t = 0; t < n; ++t)
      memcpy((char *)dest[me] + (((size_t)me * n + t) * dst) * es,
             (const char *)src[me] + (((size_t)me * n + t) * sst) * es, es);
    for (i = 1; i < P; ++i) {
      const int peer = (me + i) % P;
      for (t = 0; t < n; ++t)
        memcpy((char *)dest[peer] + (((size_t)me * n + t) * dst) * es,
               (const char *)src[me] + (((size_t)peer * n + t) * sst) * es,
               es);
    }
  }
  return 0;
}

/* broadcast.c:451-470 sized (active-set) form: no root copy, and the
 * helpers only write non-roots (binomial :150-190, `me_as != PE_root`). */
int orc_broadcast_active_set(int P, int root, size_t nbytes,
                             const void *const *src, void *const *dest) {
  int me;
  if (root < 0 || root >= P) return -1;
  for (me = 0; me < P; ++me)
    if (me != root) memcpy(dest[me], src[root], nbytes);
  return 0;
}

/* putget.h:130-186: strided put/get as one element copy per index, in
 * index order (later elements win on a repeated target) */
void orc_strided_copy(char *dst, const char *src, size_t es, ptrdiff_t tst,
                      ptrdiff_t sst, size_t n) {
  size_t t;
  for (t = 0; t < n; ++t)
    memcpy(dst + (ptrdiff_t)t * tst * (ptrdiff_t)es,
           src + (ptrdiff_t)t * sst * (ptrdiff_t)es, es);
}

/* ---- generator (same function as the device fill kernel) ---------------- */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint64_t orc_hash3(uint64_t seed, uint64_t pe, uint64_t idx) {
  const uint64_t G = 0x9E3779B97F4A7C15ull;
  return mix64(mix64(mix64(seed) + G * (pe + 1)) + G * (idx + 1));
}

/* ---- timing helpers of the CPU baseline ---- */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_d(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* ---- CPU baseline with one process per PE ---------------------------------
 * The reference runs one process per PE over UCX shared memory.  Here P
 * forked processes share one MAP_SHARED mapping that holds every PE's
 * source and destination slices, its pSync words and a barrier; a put or
 * get is a memcpy into or out of the peer's slice (a lower bound on a UCX
 * shared-memory copy), shmem_long_p a release store, and wait_until a
 * spin with sched_yield (src/shmemc/ucx/waituntil.c:18-27).  PE i runs on
 * CPU i mod C of the allowed set.  Children only touch this mapping and
 * end with _exit.  Per iteration every PE times the call between the
 * entry barrier and its return; the iteration's time is the maximum over
 * PEs, and the result is the median over iterations. */
enum { PS_WORDS = 64 };

typedef struct {
  long count, sense; /* sense-reversing barrier */
  long fail;
} mp_ctrl;

typedef struct {
  int kind, op, dtype, P, p2s, iters;
  size_t n, nb, src_b, dest_b, slot;
  local_fn f;
  mp_ctrl *ctrl;
  long *psync;   /* P x PS_WORDS */
  double *times; /* (iters + 1) x P */
  char *data;    /* P x slot: src then dest */
} mp_t;

static char *mp_src(mp_t *m, int pe) { return m->data + (size_t)pe * m->slot; }
static char *mp_dest(mp_t *m, int pe) {
  return m->data + (size_t)pe * m->slot + (m->src_b + 63) / 64 * 64;
}
static long *mp_ps(mp_t *m, int pe, int i) {
  return m->psync + (size_t)pe * PS_WORDS + (size_t)i;
}
static void mp_put_flag(mp_t *m, int pe, int i, long v) {
  __atomic_store_n(mp_ps(m, pe, i), v, __ATOMIC_RELEASE);
}
static void mp_wait_gt(mp_t *m, int me, int i, long v) {
  while (__atomic_load_n(mp_ps(m, me, i), __ATOMIC_ACQUIRE) <= v) sched_yield();
}
static void mp_wait_ne(mp_t *m, int me, int i, long v) {
  while (__atomic_load_n(mp_ps(m, me, i), __ATOMIC_ACQUIRE) == v) sched_yield();
}

static void mp_barrier(mp_t *m, long *local_sense) {
  const long s = 1 - *local_sense;
  *local_sense = s;
  if (__atomic_add_fetch(&m->ctrl->count, 1, __ATOMIC_ACQ_REL) == m->P) {
    __atomic_store_n(&m->ctrl->count, 0, __ATOMIC_RELAXED);
    __atomic_store_n(&m->ctrl->sense, s, __ATOMIC_RELEASE);
  } else {
    while (__atomic_load_n(&m->ctrl->sense, __ATOMIC_ACQUIRE) != s) sched_yield();
  }
}

/* REDUCE_HELPER_REC_DBL (reduce.c:182-293) for PE me, with its handshakes */
static int mp_rec_dbl(mp_t *m, int me) {
  const int P = m->P, p2s = m->p2s;
  const int me_p2s = member_p2s(me, P, p2s);
  char *dest = mp_dest(m, me);
  const char *src = mp_src(m, me);
  char *tmp = NULL;
  int rc = 0;
  if (me_p2s != -1) {
    tmp = (char *)malloc(m->nb ? m->nb : 1);
    if (tmp == NULL) return -1;
  }
  if (me_p2s == -1) {
    mp_put_flag(m, me - 1, 0, 1);
  } else if ((me + 1) * p2s / P == me_p2s) {
    mp_wait_ne(m, me, 0, 0);
    mp_put_flag(m, me, 0, 0);
    memcpy(dest, mp_src(m, me + 1), m->nb);
    rc |= m->f(m->op, tmp, dest, src, m->n);
  } else {
    memcpy(tmp, src, m->nb);
  }
  if (me_p2s != -1) {
    int mask, i;
    for (mask = 1, i = 1; mask < p2s; mask <<= 1, ++i) {
      const int peer = ((me_p2s ^ mask) * P + p2s - 1) / p2s;
      mp_put_flag(m, peer, i, 1);
      mp_wait_gt(m, me, i, 0);
      memcpy(mp_dest(m, peer), tmp, m->nb);
      mp_put_flag(m, peer, i, 2);
      mp_wait_gt(m, me, i, 1);
      rc |= m->f(m->op, tmp, tmp, dest, m->n);
      mp_put_flag(m, me, i, 0);
    }
    memcpy(dest, tmp, m->nb);
  }
  if (me_p2s == -1) {
    mp_wait_ne(m, me, 0, 0);
    mp_put_flag(m, me, 0, 0);
  } else if ((me + 1) * p2s / P == me_p2s) {
    memcpy(mp_dest(m, me + 1), dest, m->nb);
    mp_put_flag(m, me + 1, 0, 1);
  }
  free(tmp);
  return rc;
}

/* fcollect_helper_bruck_inplace (fcollect.c:370-403) for PE me */
static void mp_bruck(mp_t *m, int me) {
  const int P = m->P;
  const size_t total = (size_t)P * m->nb;
  size_t distance, sent = m->nb;
  int round;
  char *dest = mp_dest(m, me);
  memcpy(dest, mp_src(m, me), m->nb);
  for (distance = 1, round = 0; distance < (size_t)P; distance <<= 1, ++round) {
    const int peer = (int)(((size_t)me - distance + (size_t)P) % (size_t)P);
    const size_t to_send = (2 * sent <= total) ? sent : total - sent;
    memcpy(mp_dest(m, peer) + sent, dest, to_send);
    mp_put_flag(m, peer, round, 1);
    sent += distance * m->nb;
    mp_wait_ne(m, me, round, 0);
    mp_put_flag(m, me, round, 0);
  }
  orc_rotate_inplace(dest, total, (size_t)me * m->nb);
}

/* alltoall_helper_shift_exchange_barrier (alltoall.c:75-110) for PE me,
 * up to its final barrier */
static void mp_shift(mp_t *m, int me) {
  const int P = m->P;
  int i;
  memcpy(mp_dest(m, me) + (size_t)me * m->nb, mp_src(m, me) + (size_t)me * m->nb,
         m->nb);
  for (i = 1; i < P; ++i) {
    const int peer = (me + i) % P;
    memcpy(mp_dest(m, peer) + (size_t)me * m->nb,
           mp_src(m, me) + (size_t)peer * m->nb, m->nb);
  }
}

/* PE i's seeded input (SplitMix64 bits, or N(0,1) by Box-Muller for
 * floats), written by PE i itself */
static void mp_fill(mp_t *m, int i) {
  size_t k;
  if (m->kind == 0 && (m->dtype == F32 || m->dtype == F64)) {
    for (k = 0; k < m->n; ++k) {
      const uint64_t h = orc_hash3(0, (uint64_t)i, k);
      const double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;
      const double u2 = (double)(mix64(h ^ 0xD1B54A32D192ED03ull) >> 11) * 0x1.0p-53;
      const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      if (m->dtype == F32)
        ((float *)mp_src(m, i))[k] = (float)z;
      else
        ((double *)mp_src(m, i))[k] = z;
    }
  } else {
    for (k = 0; k + 8 <= m->src_b; k += 8) {
      const uint64_t h = orc_hash3(0, (uint64_t)i, k / 8);
      memcpy(mp_src(m, i) + k, &h, 8);
    }
  }
  memset(mp_dest(m, i), 0, m->dest_b);
}

static void mp_child(mp_t *m, int me, int cpu, int fill) {
  long sense = 0;
  int it, rc = 0;
  if (cpu >= 0) {
    cpu_set_t cs;
    CPU_ZERO(&cs);
    CPU_SET(cpu, &cs);
    (void)sched_setaffinity(0, sizeof(cs), &cs);
  }
  if (fill) mp_fill(m, me);
  for (it = 0; it < m->iters + 1; ++it) {
    double t0;
    mp_barrier(m, &sense);
    t0 = now_s();
    if (m->kind == 0) {
      /* the team wrapper's pWrk malloc and free: one WORLD barrier each
       * (reduce.c:1090-1100, shmalloc.c:56,131) */
      mp_barrier(m, &sense);
      rc |= mp_rec_dbl(m, me);
      mp_barrier(m, &sense);
    } else if (m->kind == 1) {
      mp_barrier(m, &sense); /* the wrapper's team sync (fcollect.c:574) */
      mp_bruck(m, me);
    } else {
      mp_shift(m, me);
      mp_barrier(m, &sense); /* quiet + the final barrier (alltoall.c:109) */
    }
    m->times[(size_t)it * (size_t)m->P + (size_t)me] = now_s() - t0;
  }
  if (rc) __atomic_store_n(&m->ctrl->fail, 1, __ATOMIC_RELAXED);
  _exit(0);
}

/* kind 0: rec_dbl reduce of n elements of dtype with op; 1: Bruck fcollect
 * of n bytes per PE; 2: shift-exchange alltoall of n-byte blocks.  src
 * (optional): P rows of the per-PE input (else the seeded generator);
 * dest_out (optional): receives every PE's destination after the last
 * iteration.  Returns 0, -1 on bad arguments, -2 if the mapping or a fork
 * failed. */
int orc_cpu_bench_procs(int kind, int op, int dtype, int npes, size_t n,
                        int iters, const void *const *src,
                        void *const *dest_out, double *sec_median) {
  mp_t m;
  size_t ctrl_b, ps_b, tm_b, total;
  char *map;
  pid_t *pids;
  int i, ncpu = 0, rc = 0, it;
  int cpus[1024];
  cpu_set_t cs;
  memset(&m, 0, sizeof(m));
  if (npes < 1 || iters < 1 || kind < 0 || kind > 2 || sec_median == NULL)
    return -1;
  m.kind = kind;
  m.op = op;
  m.dtype = kind == 0 ? dtype : U8;
  m.P = npes;
  m.p2s = p2s_of(npes);
  m.iters = iters;
  m.n = n;
  if (kind == 0) {
    m.f = pick_local(dtype);
    if (m.f == NULL) return -1;
    m.nb = n * esize(dtype);
    m.src_b = m.dest_b = m.nb;
  } else {
    m.nb = n;
    m.src_b = kind == 1 ? n : n * (size_t)npes;
    m.dest_b = n * (size_t)npes;
  }
  if (kind != 0 && npes > PS_WORDS * 64) return -1;
  m.slot = (m.src_b + 63) / 64 * 64 + (m.dest_b + 63) / 64 * 64;
  ctrl_b = 4096;
  ps_b = (size_t)npes * PS_WORDS * sizeof(long);
  tm_b = (size_t)(iters + 1) * (size_t)npes * sizeof(double);
  total = ctrl_b + (ps_b + 4095) / 4096 * 4096 + (tm_b + 4095) / 4096 * 4096 +
          (size_t)npes * m.slot;
  map = (char *)mmap(NULL, total, PROT_READ | PROT_WRITE,
                     MAP_SHARED | MAP_ANONYMOUS, -1, 0);
  if (map == MAP_FAILED) return -2;
  m.ctrl = (mp_ctrl *)map;
  m.psync = (long *)(map + ctrl_b);
  m.times = (double *)(map + ctrl_b + (ps_b + 4095) / 4096 * 4096);
  m.data = map + ctrl_b + (ps_b + 4095) / 4096 * 4096 + (tm_b + 4095) / 4096 * 4096;
  if (src != NULL)
    for (i = 0; i < npes; ++i) {
      memcpy(mp_src(&m, i), src[i], m.src_b);
      memset(mp_dest(&m, i), 0, m.dest_b);
    }
  if (sched_getaffinity(0, sizeof(cs), &cs) == 0)
    for (i = 0; i < CPU_SETSIZE && ncpu < 1024; ++i)
      if (CPU_ISSET(i, &cs)) cpus[ncpu++] = i;
  pids = (pid_t *)calloc((size_t)npes, sizeof(pid_t));
  if (pids == NULL) {
    munmap(map, total);
    return -2;
  }
  fflush(NULL);
  for (i = 0; i < npes; ++i) {
    pids[i] = fork();
    if (pids[i] == 0)
      mp_child(&m, i, ncpu > 0 ? cpus[i % ncpu] : -1, src == NULL);
    if (pids[i] < 0) {
      rc = -2;
      break;
    }
  }
  if (rc != 0) { /* cannot run: stop the PEs already started */
    int k;
    for (k = 0; k < i; ++k) kill(pids[k], SIGKILL);
  }
  for (i = 0; i < npes; ++i) {
    int st;
    if (pids[i] > 0) {
      waitpid(pids[i], &st, 0);
      if (!WIFEXITED(st) || WEXITSTATUS(st) != 0) rc = rc ? rc : -2;
    }
  }
  if (rc == 0 && m.ctrl->fail) rc = -1;
  if (rc == 0) {
    double *per = (double *)malloc(sizeof(double) * (size_t)iters);
    if (per == NULL) {
      rc = -2;
    } else {
      for (it = 1; it <= iters; ++it) {
        double mx = 0.0;
        for (i = 0; i < npes; ++i) {
          const double t = m.times[(size_t)it * (size_t)npes + (size_t)i];
          if (t > mx) mx = t;
        }
        per[it - 1] = mx;
      }
      qsort(per, (size_t)iters, sizeof(double), cmp_d);
      *sec_median = per[iters / 2];
      free(per);
    }
    if (dest_out != NULL)
      for (i = 0; i < npes; ++i) memcpy(dest_out[i], mp_dest(&m, i), m.dest_b);
  }
  free(pids);
  munmap(map, total);
  return rc;
}
