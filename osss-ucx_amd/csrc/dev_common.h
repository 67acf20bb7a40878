// dev_common.h - device helpers shared by the collective kernels.
//
// Compiled only by hipcc for gfx950.  All code is built with
// -ffp-contract=off: every a*b+c below is two roundings unless it is an
// explicit fma(), which the reference-order float path depends on.
#ifndef OSC_DEV_COMMON_H
#define OSC_DEV_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "osc_dev.h"

namespace osc {

// Complex element types: the layout of C99 float/double _Complex.
struct cf32 {
  float re, im;
};
struct cf64 {
  double re, im;
};

// Double-double value hi + lo with |lo| <= ulp(hi)/2.
struct dd {
  double hi, lo;
};

// Knuth's TwoSum (TAOCP vol. 2, 4.2.2): s + e == a + b exactly.
__device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

// Dekker's FastTwoSum, valid when |a| >= |b| or a == 0.
__device__ __forceinline__ dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

// TwoProduct with a fused multiply-add: p + e == a * b exactly.
__device__ __forceinline__ dd two_prod(double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  return {p, e};
}

// Double-double plus double (Hida, Li & Bailey, QD library, 2001).
__device__ __forceinline__ dd dd_add_d(dd x, double y) {
  dd s = two_sum(x.hi, y);
  s.lo += x.lo;
  return fast_two_sum(s.hi, s.lo);
}

// Double-double plus double-double, the "IEEE" variant of the QD library.
__device__ __forceinline__ dd dd_add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi);
  dd t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

// Double-double times double.
__device__ __forceinline__ dd dd_mul_d(dd x, double y) {
  dd p = two_prod(x.hi, y);
  p.lo = fma(x.lo, y, p.lo);
  return fast_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ dd dd_neg(dd x) { return {-x.hi, -x.lo}; }

// 64-bit address of team member i's copy of symmetric offset `off`.
__device__ __forceinline__ char *member_ptr(const oscd_geom &g, int i,
                                            uint64_t off) {
  return g.base + ((uint64_t)(int64_t)(g.start + i * g.stride)) * g.slice + off;
}

// 16-byte vector used for the wide HBM loads and stores.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <int BYTES>
struct Raw;
template <>
struct Raw<16> {
  typedef u32x4 type;
};
template <>
struct Raw<8> {
  typedef u32x2 type;
};
template <>
struct Raw<4> {
  typedef unsigned int type;
};
template <>
struct Raw<2> {
  typedef unsigned short type;
};
template <>
struct Raw<1> {
  typedef unsigned char type;
};

template <int BYTES>
__device__ __forceinline__ typename Raw<BYTES>::type ld(const void *p) {
  return *(const typename Raw<BYTES>::type *)p;
}

template <int BYTES>
__device__ __forceinline__ void st(void *p, typename Raw<BYTES>::type v) {
  *(typename Raw<BYTES>::type *)p = v;
}

// Streaming store: the result is not re-read by this kernel.
template <int BYTES>
__device__ __forceinline__ void st_nt(void *p, typename Raw<BYTES>::type v) {
  __builtin_nontemporal_store(v, (typename Raw<BYTES>::type *)p);
}

// Streaming load: every source byte is read exactly once.
template <int BYTES>
__device__ __forceinline__ typename Raw<BYTES>::type ld_nt(const void *p) {
  return __builtin_nontemporal_load((const typename Raw<BYTES>::type *)p);
}

}  // namespace osc

#endif  // OSC_DEV_COMMON_H
