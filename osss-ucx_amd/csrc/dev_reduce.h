// dev_reduce.h - the team reduction kernel (all ops x all element types).
//
// Semantics follow the reference's default team reduction, recursive
// doubling (src/shcoll/src/reduce.c:182-293, REDUCE_HELPER_REC_DBL), with
// the element operators of reduce.c:730-736.  Its result has a closed form
// that one pass over the P source slices computes directly:
//
//   p2s = largest power of two <= P, lg = log2(p2s)
//   member k (k < p2s) is team index m_k = ceil(k*P/p2s)
//   leaf k = op(x[m_k + 1], x[m_k]) if m_k + 1 is not a member, else x[m_k]
//     (reduce.c:228-244: the member pulls its right neighbour's source and
//      combines it as op(neighbour, own))
//   member k's result = balanced pairwise tree over leaves k^0, k^1, ...,
//     k^(p2s-1), each node op(left subtree, right subtree)
//     (reduce.c:250-272: round r combines op(own, partner k^2^r))
//   the non-member m_k + 1 receives member k's result (reduce.c:277-288).
//
// For + and * the two partners of a round compute op(a,b) and op(b,a),
// which are identical in IEEE arithmetic, so every member ends with the
// bits of member 0's tree; integer and bitwise ops are exact in any order.
// Only float MIN/MAX differ between members (ternary a>b?a:b picks the
// second operand on ties and NaN), and only where a NaN is present or the
// extremum is a signed zero.  Such elements take a block-cooperative exact
// path that evaluates every member's own tree.
//
// Layout: a 16-byte "unit" holds VEC elements.  A workgroup of 256 threads
// owns 256*UPT consecutive units; thread t handles units t, t+256, ...,
// t+256*(UPT-1), so each wave instruction touches 1 KiB contiguous per slice
// and a workgroup streams 256*UPT*16 bytes contiguous per slice.  Sources
// are read INFL slices at a time (UPT*INFL 16-byte loads in flight per
// thread), combined in registers, and the result is written to all P
// destination slices.  Loads and stores are non-temporal: every source
// byte is read once and every destination byte written once, so HBM
// traffic = 2*P*n*s.  (UPT, INFL) is a per-call "shape" chosen from a
// tuning table.
//
// Small integer calls (16+ members, at most 2048 elements) take a second
// kernel instead, k_reduce_lanes: a group of lanes shares one element,
// reads its members in parallel and combines them with a cross-lane
// butterfly (DPP, ds_swizzle, ds_bpermute), so the call costs one HBM round
// trip where the slice walk takes one per 8 members.  Small float and
// complex reductions (16..1024 members, at most 16 KiB per member) take
// k_reduce_gather: the lanes read the members in parallel into LDS and
// every lane folds them in team order, the walk's own order, so the result
// has the walk's bits (float MIN/MAX elements with a NaN or a signed-zero
// extremum get per-member results there too, and reference mode folds the
// reference's tree).
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

#ifndef OSC_DEV_REDUCE_H
#define OSC_DEV_REDUCE_H

namespace osc {

// ---- element operators (reduce.c:730-736) --------------------------------
__device__ __forceinline__ cf32 cadd(cf32 a, cf32 b) {
  return {a.re + b.re, a.im + b.im};
}
__device__ __forceinline__ cf64 cadd(cf64 a, cf64 b) {
  return {a.re + b.re, a.im + b.im};
}
// C99 complex product for finite operands: (ac - bd) + (ad + bc)i, each
// product and sum rounded separately (no contraction).
__device__ __forceinline__ cf32 cmul(cf32 a, cf32 b) {
  float ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
  return {ac - bd, ad + bc};
}
__device__ __forceinline__ cf64 cmul(cf64 a, cf64 b) {
  double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im,
         bc = a.im * b.re;
  return {ac - bd, ad + bc};
}

template <int OP>
struct Bin;
template <>
struct Bin<OSC_OP_AND> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (T)(a & b);
  }
};
template <>
struct Bin<OSC_OP_OR> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (T)(a | b);
  }
};
template <>
struct Bin<OSC_OP_XOR> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (T)(a ^ b);
  }
};
template <>
struct Bin<OSC_OP_MIN> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (a < b) ? a : b;
  }
};
template <>
struct Bin<OSC_OP_MAX> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (a > b) ? a : b;
  }
};
template <>
struct Bin<OSC_OP_SUM> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (T)(a + b);
  }
  __device__ static cf32 f(cf32 a, cf32 b) { return cadd(a, b); }
  __device__ static cf64 f(cf64 a, cf64 b) { return cadd(a, b); }
};
template <>
struct Bin<OSC_OP_PROD> {
  template <class T>
  __device__ static T f(T a, T b) {
    return (T)(a * b);
  }
  __device__ static cf32 f(cf32 a, cf32 b) { return cmul(a, b); }
  __device__ static cf64 f(cf64 a, cf64 b) { return cmul(a, b); }
};

// ---- accumulation policies -----------------------------------------------
// TREE    : combine in the recursive-doubling tree order (reference mode).
// SPECIAL : float MIN/MAX, whose per-member results can differ.
// lift/seq/comb/out: value -> accumulator, acc (+) value, acc (+) acc,
//                    accumulator -> value.

// Exact in any order: integers (wrap-around), bitwise, float MIN/MAX.
template <class T, int OP, bool SPEC>
struct PolSeq {
  typedef T Acc;
  static constexpr bool TREE = false;
  static constexpr int SHAPE = 0;  // default kernel shape
  static constexpr int WAVES81 = 0;
  static constexpr bool SPECIAL = SPEC;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(T x) { return x; }
  __device__ static Acc seq(Acc a, T x) { return mm(a, x); }
  __device__ static Acc comb(Acc a, Acc b) { return mm(a, b); }
  __device__ static T out(Acc a) { return a; }
  // f64 MIN/MAX under the special-element rule: an element whose inputs
  // hold a NaN or whose result is +-0 is recomputed exactly in the
  // reference's order (special_block), and for every other element the
  // hardware v_min_f64 / v_max_f64 returns the reference ternary's bits
  // (one instruction instead of a compare and two selects).  At 64 members
  // x 4M on MI355X the f64 min/max walks took 325.0 / 325.7 us against a
  // median of 322.8 us for the 8-byte integer walks of the same run; with
  // the ternary they ran 2.8-3% behind them.
  __device__ static Acc mm(Acc a, Acc b) {
    if constexpr (SPEC && std::is_same<T, double>::value &&
                  OP == OSC_OP_MIN) {
      return __builtin_fmin(a, b);
    } else if constexpr (SPEC && std::is_same<T, double>::value &&
                         OP == OSC_OP_MAX) {
      return __builtin_fmax(a, b);
    } else {
      return Bin<OP>::f(a, b);
    }
  }
};

// Float MIN/MAX in the reference Rabenseifner reductions (reduce.c:
// 306-511, 521-724): an element's result is its owner's recursive-doubling
// tree (the owner of block b of the reduce-scatter is member
// reverse_bits(b)), copied to every member; in the fold of a
// non-power-of-two pair the upper half of the elements combine as
// op(x[m], x[m+1]).  Only NaN and signed-zero elements can tell.
template <class T, int OP>
struct PolSeqRab : PolSeq<T, OP, true> {
  static constexpr bool RAB = true;
};

// The type's own arithmetic in the reference's tree order.
template <class T, int OP>
struct PolTree {
  typedef T Acc;
  static constexpr bool TREE = true;
  // default kernel shape: 1x8 for double (measured 5.26 vs 4.94 TB/s for
  // 2x4 at 64 PEs x 4M), 2x4 otherwise (float within 1% across shapes)
  static constexpr int SHAPE = std::is_same<T, double>::value ? 3 : 2;
  static constexpr int WAVES81 = 0;
  static constexpr bool SPECIAL = false;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(T x) { return x; }
  __device__ static Acc seq(Acc a, T x) { return Bin<OP>::f(a, x); }
  __device__ static Acc comb(Acc a, Acc b) { return Bin<OP>::f(a, b); }
  __device__ static T out(Acc a) { return a; }
};

// The type's own arithmetic in the reference binomial reduction's order
// (reduce.c:102-171): member a folds each child a + 2^j (2^j below a's
// lowest set bit, a + 2^j < P) as op(child's result, own partial), the
// children in ascending order, which is the order of the reference when
// every child has reported before its parent folds (the reference folds
// children in arrival order).  That is a pairwise tree over all P members
// in team order, node = op(right subtree, left subtree), truncated at P.
template <class T, int OP>
struct PolBinom : PolTree<T, OP> {
  static constexpr bool BINOM = true;
};

// FP32 SUM/PROD accumulated in FP64, rounded once.
template <int OP>
struct PolF32Acc {
  typedef double Acc;
  static constexpr bool TREE = false;
  static constexpr int SHAPE = 1;  // default kernel shape
  static constexpr int WAVES81 = 0;
  static constexpr bool SPECIAL = false;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(float x) { return (double)x; }
  __device__ static Acc seq(Acc a, float x) {
    return OP == OSC_OP_SUM ? a + (double)x : a * (double)x;
  }
  __device__ static Acc comb(Acc a, Acc b) {
    return OP == OSC_OP_SUM ? a + b : a * b;
  }
  __device__ static float out(Acc a) { return (float)a; }
};

// FP64 SUM/PROD accumulated in twice the working precision, rounded once.
// SUM is Sum2 of Ogita, Rump and Oishi ("Accurate sum and dot product",
// SIAM J. Sci. Comput. 26(6), 2005): TwoSum keeps the running sum's exact
// rounding error, which collects in a second accumulator added once at the
// end; the result is as accurate as a sum in double-double rounded to
// double.  PROD multiplies in double-double.
template <int OP>
struct PolF64Acc {
  typedef dd Acc;
  static constexpr bool TREE = false;
  // The accurate sum keeps 2 doubles per element: at 8x2 it takes 180
  // VGPRs (2 waves per SIMD).  8x1 held to 4 waves per SIMD measured 3%
  // faster at 64 PEs x 4M when the walk stored every member; the walk that
  // stores member 0 only (split calls) is 1.5% faster at 8x2 (663 vs 674
  // us per call).
  static constexpr int SHAPE = OP == OSC_OP_SUM ? 1 : 0;  // default shape
  static constexpr int SHAPE_ONE = 0;  // default shape of a split call
  static constexpr int WAVES81 = OP == OSC_OP_SUM ? 4 : 0;
  static constexpr bool SPECIAL = false;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(double x) { return {x, 0.0}; }
  __device__ static Acc seq(Acc a, double x) {
    if constexpr (OP == OSC_OP_SUM) {
      const dd s = two_sum(a.hi, x);
      return {s.hi, a.lo + s.lo};
    } else {
      return dd_mul_d(a, x);
    }
  }
  __device__ static Acc comb(Acc a, Acc b) { return dd_add(a, b); }
  __device__ static double out(Acc a) {
    return OP == OSC_OP_SUM ? a.hi + a.lo : a.hi;
  }
};

// complex float SUM/PROD in FP64 components.
struct c64acc {
  double re, im;
};
template <int OP>
struct PolC32Acc {
  typedef c64acc Acc;
  static constexpr bool TREE = false;
  static constexpr int SHAPE = 0;  // default kernel shape
  static constexpr int WAVES81 = 0;
  static constexpr bool SPECIAL = false;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(cf32 x) { return {(double)x.re, (double)x.im}; }
  __device__ static Acc seq(Acc a, cf32 x) {
    const double xr = x.re, xi = x.im;
    if (OP == OSC_OP_SUM) return {a.re + xr, a.im + xi};
    double ac = a.re * xr, bd = a.im * xi, ad = a.re * xi, bc = a.im * xr;
    return {ac - bd, ad + bc};
  }
  __device__ static Acc comb(Acc a, Acc b) { return {a.re + b.re, a.im + b.im}; }
  __device__ static cf32 out(Acc a) { return {(float)a.re, (float)a.im}; }
};

// complex double SUM/PROD in double-double components.
struct cddacc {
  dd re, im;
};
template <int OP>
struct PolC64Acc {
  typedef cddacc Acc;
  static constexpr bool TREE = false;
  static constexpr int SHAPE = 0;  // default kernel shape
  static constexpr int WAVES81 = 0;
  static constexpr bool SPECIAL = false;
  static constexpr bool BINOM = false;
  static constexpr bool RAB = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(cf64 x) { return {{x.re, 0.0}, {x.im, 0.0}}; }
  __device__ static Acc seq(Acc a, cf64 x) {
    if (OP == OSC_OP_SUM) return {dd_add_d(a.re, x.re), dd_add_d(a.im, x.im)};
    dd re = dd_add(dd_mul_d(a.re, x.re), dd_neg(dd_mul_d(a.im, x.im)));
    dd im = dd_add(dd_mul_d(a.re, x.im), dd_mul_d(a.im, x.re));
    return {re, im};
  }
  __device__ static Acc comb(Acc a, Acc b) {
    return {dd_add(a.re, b.re), dd_add(a.im, b.im)};
  }
  __device__ static cf64 out(Acc a) { return {a.re.hi, a.im.hi}; }
};

// ---- names of the kernels a call launched --------------------------------
// Each launch records the address of a static string (two stores on the
// calling thread), so measurement tools can name the kernels a reduce call
// ran; oscd_last_reduce_kernels formats them as rocprof prints them
// ("k_reduce<double, osc::PolF64Acc<5>, 2, 8, 1, 1, true>").  A name is the
// __PRETTY_FUNCTION__ of a helper whose template parameters are the
// kernel's, or a plain name.
constexpr int kMaxLaunchNames = 8;
struct LaunchNote {
  const char *what;  // "k_reduce" ... or a plain name
  const char *pretty;  // __PRETTY_FUNCTION__ of the helper, or nullptr
};
extern thread_local LaunchNote t_launch[kMaxLaunchNames];
extern thread_local int t_nlaunch;
// Which part of a reduction this thread's next launches issue
// (oscd_set_split_phase): -1 all of it; 0 the walk of a split call, or the
// whole call when it does not split; 1 the replicate of a split call and
// nothing else.  Under 0 and 1 float MIN/MAX split only with span flags of
// their own (t_split_flags: one byte per 256 units of the call, the same
// for both phases), since a stream's flags serve one call at a time.
extern thread_local int t_split_phase;
extern thread_local uint8_t *t_split_flags;
inline void note_launch(const char *what, const char *pretty) {
  if (t_nlaunch < kMaxLaunchNames) t_launch[t_nlaunch++] = {what, pretty};
}
template <class T, class Pol, int VEC, int UPT, int INFL, int MAXLV, bool ONE>
const char *k_reduce_pf() {
  return __PRETTY_FUNCTION__;
}
template <class T, class Pol, int G>
const char *k_reduce_lanes_pf() {
  return __PRETTY_FUNCTION__;
}
template <class T, class Pol>
const char *k_reduce_gather_pf() {
  return __PRETTY_FUNCTION__;
}

// ---- unit load / store ---------------------------------------------------
template <class T, int VEC>
__device__ __forceinline__ void load_unit(const char *p, T (&x)[VEC],
                                          bool full, int cnt) {
  constexpr int B = VEC * (int)sizeof(T);
  if (B <= 16 && (B & (B - 1)) == 0 && full) {
    typename Raw<(B <= 16 ? B : 16)>::type r = ld_nt<(B <= 16 ? B : 16)>(p);
    __builtin_memcpy(&x[0], &r, B);
  } else {
#pragma unroll
    for (int v = 0; v < VEC; ++v)
      if (v < cnt) x[v] = ((const T *)p)[v];
  }
}

template <class T, int VEC>
__device__ __forceinline__ void store_unit(char *p, const T (&x)[VEC],
                                           bool full, int cnt,
                                           uint32_t skip) {
  constexpr int B = VEC * (int)sizeof(T);
  if (B <= 16 && (B & (B - 1)) == 0 && full && skip == 0) {
    typename Raw<(B <= 16 ? B : 16)>::type r;
    __builtin_memcpy(&r, &x[0], B);
    st_nt<(B <= 16 ? B : 16)>(p, r);
  } else {
#pragma unroll
    for (int v = 0; v < VEC; ++v)
      if (v < cnt && !((skip >> v) & 1u)) ((T *)p)[v] = x[v];
  }
}

template <class T>
__device__ __forceinline__ bool is_nan_or_zero(T) {
  return false;
}
__device__ __forceinline__ bool is_nan_or_zero(float x) {
  return x != x || x == 0.0f;
}
__device__ __forceinline__ bool is_nan_or_zero(double x) {
  return x != x || x == 0.0;
}
template <class T>
__device__ __forceinline__ bool is_nan(T) {
  return false;
}
__device__ __forceinline__ bool is_nan(float x) { return x != x; }
__device__ __forceinline__ bool is_nan(double x) { return x != x; }

// c ? a : b, component-wise for the complex accumulators (a select of the
// whole struct would go through scratch memory).
template <class A>
__device__ __forceinline__ A pick(bool c, const A &a, const A &b) {
  return c ? a : b;
}
__device__ __forceinline__ cf32 pick(bool c, const cf32 &a, const cf32 &b) {
  cf32 r;
  r.re = c ? a.re : b.re;
  r.im = c ? a.im : b.im;
  return r;
}
__device__ __forceinline__ cf64 pick(bool c, const cf64 &a, const cf64 &b) {
  cf64 r;
  r.re = c ? a.re : b.re;
  r.im = c ? a.im : b.im;
  return r;
}
// the accumulators' own component-wise selects: a select of the whole
// struct kept the FP64 walk's accumulators in scratch memory
__device__ __forceinline__ dd pick(bool c, const dd &a, const dd &b) {
  dd r;
  r.hi = c ? a.hi : b.hi;
  r.lo = c ? a.lo : b.lo;
  return r;
}
__device__ __forceinline__ c64acc pick(bool c, const c64acc &a,
                                       const c64acc &b) {
  c64acc r;
  r.re = c ? a.re : b.re;
  r.im = c ? a.im : b.im;
  return r;
}
__device__ __forceinline__ cddacc pick(bool c, const cddacc &a,
                                       const cddacc &b) {
  cddacc r;
  r.re = pick(c, a.re, b.re);
  r.im = pick(c, a.im, b.im);
  return r;
}

// Team index of leaf k's member: ceil(k*P / 2^lg).
__device__ __forceinline__ int member_of(int k, int P, int lg) {
  return (int)(((int64_t)k * P + ((1 << lg) - 1)) >> lg);
}

// The lowest lg bits of k in reverse order (the reference's reverse_bits,
// src/shcoll/src/util/bithacks.c:11).
__device__ __forceinline__ int rev_bits(int k, int lg) {
  return lg == 0 ? 0 : (int)(__builtin_bitreverse32((uint32_t)k) >> (32 - lg));
}

// Rabenseifner's owner of element e of n: the reduce-scatter leaves block
// b = [floor(b*n/p2s), floor((b+1)*n/p2s)) on member reverse_bits(b)
// (reduce.c:411-446, collect phase :452-455), so e belongs to the largest
// b with floor(b*n/p2s) <= e.
__device__ __forceinline__ int rab_owner(uint64_t e, uint64_t n, int lg) {
  const uint64_t p2s = 1ull << lg;
  const uint64_t b = ((e + 1) * p2s + n - 1) / n - 1;
  return rev_bits((int)b, lg);
}

// Member k's recursive-doubling result over snap[0..P) (leaves in the
// XOR-permuted order of its butterfly, reduce.c:250-272).  A folded pair
// is op(x[m+1], x[m]) (reduce.c:232-241), or op(x[m], x[m+1]) when swap
// (the upper half of a Rabenseifner fold, reduce.c:370-382).
template <class T, int OP>
__device__ T member_tree(const T *snap, int k, int P, int lg, bool swap) {
  const int p2s = 1 << lg;
  T stk[16];
  for (int j = 0; j < p2s; ++j) {
    const int kk = k ^ j;
    const int m = member_of(kk, P, lg);
    const int m1 = member_of(kk + 1, P, lg);
    T leaf = snap[m];
    if (m1 - m == 2)
      leaf = swap ? Bin<OP>::f(snap[m], snap[m + 1])
                  : Bin<OP>::f(snap[m + 1], snap[m]);
    int L = 0;
    while ((j >> L) & 1) {
      leaf = Bin<OP>::f(stk[L], leaf);
      ++L;
    }
    stk[L] = leaf;
  }
  return stk[lg];
}

// ---- exact per-member MIN/MAX for special elements ----------------------
// Cooperative over the workgroup: each special element is handled in turn.
// Its P inputs are copied into LDS before any result is written (so
// dest == source works), then thread t evaluates members t, t+256, ...
// over the LDS copy and writes their results (and the folded neighbour's).
// `mine` has bit k*VEC+v set for element v of the thread's k-th unit
// (unit index ubase + k*256).
// RAB: the Rabenseifner semantics instead, one owner's result to every
// member.
template <class T, int OP, int VEC, bool RAB>
__device__ void special_block(const oscd_geom &g, uint64_t src_off,
                              uint64_t dest_off, uint64_t ubase,
                              uint32_t mine, int lg, uint64_t n, T *snap) {
  __shared__ int s_owner;
  __shared__ unsigned long long s_e;
  __shared__ T s_r;
  const int P = g.npes;
  const int p2s = 1 << lg;
  for (;;) {
    if (threadIdx.x == 0) s_owner = 0x7fffffff;
    __syncthreads();
    if (mine) atomicMin(&s_owner, (int)threadIdx.x);
    __syncthreads();
    const int owner = s_owner;
    if (owner == 0x7fffffff) break;
    if ((int)threadIdx.x == owner) {
      const int b = __ffs(mine) - 1;
      mine &= mine - 1;
      s_e = (ubase + (uint64_t)(b / VEC) * blockDim.x) * VEC + (uint64_t)(b % VEC);
    }
    __syncthreads();
    const uint64_t e = s_e;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
      snap[i] = *(const T *)(member_ptr(g, i, src_off) + e * sizeof(T));
    __syncthreads();
    if constexpr (RAB) {
      if (threadIdx.x == 0)
        s_r = member_tree<T, OP>(snap, rab_owner(e, n, lg), P, lg, e >= n / 2);
      __syncthreads();
      const T r = s_r;
      for (int i = threadIdx.x; i < P; i += blockDim.x)
        *(T *)(member_ptr(g, i, dest_off) + e * sizeof(T)) = r;
    } else {
      for (int k = threadIdx.x; k < p2s; k += blockDim.x) {
        const T r = member_tree<T, OP>(snap, k, P, lg, false);
        const int mk = member_of(k, P, lg);
        const int mk1 = member_of(k + 1, P, lg);
        *(T *)(member_ptr(g, mk, dest_off) + e * sizeof(T)) = r;
        if (mk1 - mk == 2)
          *(T *)(member_ptr(g, mk + 1, dest_off) + e * sizeof(T)) = r;
      }
    }
    __syncthreads();
  }
}


// ---- the kernel ------------------------------------------------------------
constexpr int kBlock = 256;

// Shapes (units per thread, source slices in flight) a launch can use:
// 0: 8x2, 1: 8x1, 2: 2x4, 3: 1x8.
constexpr int kNumShapes = 4;
// Below this many default-shape spans (one per CU of the 256) a call takes
// the narrow shape 3 instead.
constexpr uint64_t kSmallSpans = 256;
__host__ __device__ constexpr int shape_upt(int s) {
  return s == 0 ? 8 : s == 1 ? 8 : s == 2 ? 2 : 1;
}
__host__ __device__ constexpr int shape_infl(int s) {
  return s == 0 ? 2 : s == 1 ? 1 : s == 2 ? 4 : 8;
}

// One block-iteration: the workgroup reduces units [ub, ub + 256*UPT).
// Addresses are a wave-uniform base (slice + block start, kept in SGPRs)
// plus a 32-bit lane offset (< 32 KiB), so the loads and stores use the
// scalar-base addressing form.  FAST: every unit of the block is complete
// and in range (all blocks but the last), so no per-unit guards.
// ONE: store the result to member 0 only (the replicate kernel copies it to
// every member afterwards).  With a SPECIAL policy (float MIN/MAX, whose
// special elements have per-member results) a block whose span holds a
// special element stores every member itself, as without ONE, and sets
// the flag of each of its 256-unit pieces in `flags` (one byte per piece of
// the call, written for every piece: 1 = skip it, 0 = replicate it).
template <bool FAST, class T, class Pol, int VEC, int UPT, int INFL, int MAXLV,
          bool ONE = false>
__device__ __forceinline__ void reduce_block(const oscd_geom &g,
                                             const uint64_t dest_off,
                                             const uint64_t src_off,
                                             const uint64_t n, const int lg,
                                             const uint64_t ub, T *lds,
                                             uint8_t *flags = nullptr) {
  typedef typename Pol::Acc Acc;
  constexpr uint32_t UB = VEC * (uint32_t)sizeof(T);
  const int P = g.npes;
  const int64_t step = (int64_t)g.stride * (int64_t)g.slice;
  const char *s0 = member_ptr(g, 0, src_off) + ub * UB;
  char *d0 = member_ptr(g, 0, dest_off) + ub * UB;
  const uint64_t nunits = (n + VEC - 1) / VEC;
  const uint64_t ubase = ub + threadIdx.x;
  uint32_t loff[UPT];
  bool full[UPT];
  int cnt[UPT];
  bool any = FAST;
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    loff[k] = (threadIdx.x + (uint32_t)k * kBlock) * UB;
    if (FAST) {
      full[k] = true;
      cnt[k] = VEC;
    } else {
      const uint64_t u = ubase + (uint64_t)k * kBlock;
      const uint64_t e0 = u * VEC;
      const bool act = u < nunits;
      full[k] = act && (e0 + VEC <= n);
      cnt[k] = !act ? 0 : (full[k] ? VEC : (int)(n - e0));
      any |= act;
    }
  }
  T r[UPT][VEC];
  uint32_t spec = 0;

  if (any) {
    Acc acc[UPT][VEC];
    if constexpr (!Pol::TREE) {
      uint32_t nanm = 0;
      // The small-call shapes (2x4, 1x8) load the seed slice and the leftover
      // slices in full batches, so a call makes ceil(P / INFL) dependent
      // round trips.  The wide shapes (8x2, 8x1) keep the seed slice and the
      // leftover slices apart: at 64 members x 4M the batched form ran the
      // matrix 0.7-2.7% slower (one box, alternating runs).
      if constexpr (INFL >= 4) {
        // Slices in batches of INFL whose loads are unconditional (so the
        // compiler's wait counts stay exact and the next batch's loads are
        // not held back by a drained queue).  The first batch holds slice 0,
        // which seeds the accumulator; the last holds the P % INFL leftover
        // slices.  In those two an index past P-1 re-reads slice P-1 and its
        // value is selected away, so every load still issues unconditionally
        // and a call makes ceil(P / INFL) dependent round trips: seeding from
        // slice 0 alone and taking the leftover slices one at a time made a
        // 1x8 call on 8 members take 8 round trips instead of one.
        // slices i .. i+INFL-1, those below P combined in order; FIRST: slice
        // i (= 0) seeds the accumulator
        auto batch = [&](const int i, const bool first, const bool tail)
                         __attribute__((always_inline)) {
          T buf[INFL][UPT][VEC];
  #pragma unroll
          for (int j = 0; j < INFL; ++j) {
            const int m = (!tail || i + j < P) ? i + j : P - 1;
            const char *sp = s0 + m * step;
  #pragma unroll
            for (int k = 0; k < UPT; ++k)
              if (FAST || cnt[k] > 0)
                load_unit<T, VEC>(sp + loff[k], buf[j][k], FAST || full[k],
                                  cnt[k]);
          }
  #pragma unroll
          for (int j = 0; j < INFL; ++j) {
            const bool in = !tail || i + j < P;
  #pragma unroll
            for (int k = 0; k < UPT; ++k)
  #pragma unroll
              for (int v = 0; v < VEC; ++v) {
                if (Pol::SPECIAL && in && is_nan(buf[j][k][v]))
                  nanm |= 1u << (k * VEC + v);
                if (first && j == 0) {
                  acc[k][v] = Pol::lift(buf[j][k][v]);
                } else {
                  acc[k][v] = pick(in, Pol::seq(acc[k][v], buf[j][k][v]),
                                   acc[k][v]);
                }
              }
          }
        };
        batch(0, true, INFL > P);
        int i = INFL;
        for (; i + INFL <= P; i += INFL) batch(i, false, false);
        if (i < P) batch(i, false, true);
      } else {
        // Slice 0 seeds the accumulator; then groups of INFL slices whose
        // loads are unconditional (so the compiler's wait counts stay exact
        // and the next group's loads are not held back by a drained queue),
        // then the < INFL leftover slices one at a time.
        {
          T buf[UPT][VEC];
  #pragma unroll
          for (int k = 0; k < UPT; ++k)
            if (FAST || cnt[k] > 0)
              load_unit<T, VEC>(s0 + loff[k], buf[k], FAST || full[k], cnt[k]);
  #pragma unroll
          for (int k = 0; k < UPT; ++k)
  #pragma unroll
            for (int v = 0; v < VEC; ++v) {
              if (Pol::SPECIAL && is_nan(buf[k][v])) nanm |= 1u << (k * VEC + v);
              acc[k][v] = Pol::lift(buf[k][v]);
            }
        }
        int i = 1;
        for (; i + INFL <= P; i += INFL) {
          T buf[INFL][UPT][VEC];
  #pragma unroll
          for (int j = 0; j < INFL; ++j) {
            const char *sp = s0 + (i + j) * step;
  #pragma unroll
            for (int k = 0; k < UPT; ++k)
              if (FAST || cnt[k] > 0)
                load_unit<T, VEC>(sp + loff[k], buf[j][k], FAST || full[k],
                                  cnt[k]);
          }
  #pragma unroll
          for (int j = 0; j < INFL; ++j)
  #pragma unroll
            for (int k = 0; k < UPT; ++k)
  #pragma unroll
              for (int v = 0; v < VEC; ++v) {
                if (Pol::SPECIAL && is_nan(buf[j][k][v]))
                  nanm |= 1u << (k * VEC + v);
                acc[k][v] = Pol::seq(acc[k][v], buf[j][k][v]);
              }
        }
        for (; i < P; ++i) {
          T buf[UPT][VEC];
          const char *sp = s0 + i * step;
  #pragma unroll
          for (int k = 0; k < UPT; ++k)
            if (FAST || cnt[k] > 0)
              load_unit<T, VEC>(sp + loff[k], buf[k], FAST || full[k], cnt[k]);
  #pragma unroll
          for (int k = 0; k < UPT; ++k)
  #pragma unroll
            for (int v = 0; v < VEC; ++v) {
              if (Pol::SPECIAL && is_nan(buf[k][v])) nanm |= 1u << (k * VEC + v);
              acc[k][v] = Pol::seq(acc[k][v], buf[k][v]);
            }
        }
      }
#pragma unroll
      for (int k = 0; k < UPT; ++k)
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          r[k][v] = Pol::out(acc[k][v]);
          const int b = k * VEC + v;
          if (Pol::SPECIAL && v < cnt[k] &&
              (((nanm >> b) & 1u) || is_nan_or_zero(r[k][v])))
            spec |= 1u << b;
        }
    } else {
      // Reference tree order: leaves in team-index order, folded pairs
      // op(x[m+1], x[m]), merged by a binary counter: stk[L] holds the
      // pending left subtree of 2^L leaves.
      Acc stk[MAXLV][UPT][VEC];
      T pend[UPT][VEC];
      int kl = 0;
      int mk = 0;
      int mn = member_of(1, P, lg);
      // one slice, in team-index order: fold, then carry the new subtree
      // up while level L is occupied.  Every stk index is a constant after
      // unrolling, and each level's store is unconditional (a select): a
      // store under a branch let the optimizer merge the levels' stores into
      // one store through a selected pointer, which kept stk in scratch
      // memory.  The merge itself stays under a branch that is uniform over
      // the workgroup (kl is the same for every thread).
      auto eat = [&](const T (&b)[UPT][VEC], const int i) {
        Acc leaf[UPT][VEC];
        bool ready = true;
        if constexpr (Pol::BINOM) {
          // binomial order: every member is a leaf, no fold
#pragma unroll
          for (int k = 0; k < UPT; ++k)
#pragma unroll
            for (int v = 0; v < VEC; ++v) leaf[k][v] = Pol::lift(b[k][v]);
        } else if (i == mk) {
          if (mn - mk == 2) {
#pragma unroll
            for (int k = 0; k < UPT; ++k)
#pragma unroll
              for (int v = 0; v < VEC; ++v) pend[k][v] = b[k][v];
            ready = false;
          } else {
#pragma unroll
            for (int k = 0; k < UPT; ++k)
#pragma unroll
              for (int v = 0; v < VEC; ++v) leaf[k][v] = Pol::lift(b[k][v]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < UPT; ++k)
#pragma unroll
            for (int v = 0; v < VEC; ++v)
              leaf[k][v] = Pol::comb(Pol::lift(b[k][v]), Pol::lift(pend[k][v]));
        }
        if (ready) {
          bool carry = true;
#pragma unroll
          for (int L = 0; L < MAXLV; ++L) {
            if (L <= lg) {  // the root is level lg; higher levels stay unused
              const bool occupied = ((kl >> L) & 1) != 0;
              const bool place = carry && !occupied;
              if (carry && occupied) {
                // recursive doubling: op(left, right); binomial: op(right,
                // left), the child's result first
#pragma unroll
                for (int k = 0; k < UPT; ++k)
#pragma unroll
                  for (int v = 0; v < VEC; ++v)
                    leaf[k][v] = Pol::BINOM ? Pol::comb(leaf[k][v], stk[L][k][v])
                                            : Pol::comb(stk[L][k][v], leaf[k][v]);
              }
#pragma unroll
              for (int k = 0; k < UPT; ++k)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                  stk[L][k][v] = pick(place, leaf[k][v], stk[L][k][v]);
              carry = carry && occupied;
            }
          }
          ++kl;
          if constexpr (!Pol::BINOM) {
            mk = mn;
            mn = member_of(kl + 1, P, lg);
          }
        }
      };
      // groups of INFL slices with unconditional loads, then the leftover
      // slices one at a time (as in the order-free branch above)
      int i = 0;
      for (; i + INFL <= P; i += INFL) {
        T buf[INFL][UPT][VEC];
#pragma unroll
        for (int j = 0; j < INFL; ++j) {
          const char *sp = s0 + (i + j) * step;
#pragma unroll
          for (int k = 0; k < UPT; ++k)
            if (FAST || cnt[k] > 0)
              load_unit<T, VEC>(sp + loff[k], buf[j][k], FAST || full[k],
                                cnt[k]);
        }
#pragma unroll
        for (int j = 0; j < INFL; ++j) eat(buf[j], i + j);
      }
      for (; i < P; ++i) {
        T buf[UPT][VEC];
        const char *sp = s0 + i * step;
#pragma unroll
        for (int k = 0; k < UPT; ++k)
          if (FAST || cnt[k] > 0)
            load_unit<T, VEC>(sp + loff[k], buf[k], FAST || full[k], cnt[k]);
        eat(buf, i);
      }
      if constexpr (Pol::BINOM) {
        // P leaves leave one subtree per set bit of P (the highest is the
        // left-most): merge them from the lowest up, each as op(what is
        // to its right, it)
        bool have = false;
#pragma unroll
        for (int L = 0; L < MAXLV; ++L)
          if (L <= lg && ((P >> L) & 1)) {
            if (have) {  // uniform: depends on P and L only
#pragma unroll
              for (int k = 0; k < UPT; ++k)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                  acc[k][v] = Pol::comb(acc[k][v], stk[L][k][v]);
            } else {
#pragma unroll
              for (int k = 0; k < UPT; ++k)
#pragma unroll
                for (int v = 0; v < VEC; ++v) acc[k][v] = stk[L][k][v];
            }
            have = true;
          }
      } else {
#pragma unroll
        for (int L = 0; L < MAXLV; ++L)
          if (L == lg) {
#pragma unroll
            for (int k = 0; k < UPT; ++k)
#pragma unroll
              for (int v = 0; v < VEC; ++v) acc[k][v] = stk[L][k][v];
          }
      }
#pragma unroll
      for (int k = 0; k < UPT; ++k)
#pragma unroll
        for (int v = 0; v < VEC; ++v) r[k][v] = Pol::out(acc[k][v]);
    }
  }

  auto store_members = [&](const int j0, const int j1) {
    for (int j = j0; j < j1; ++j) {
      char *dp = d0 + j * step;
#pragma unroll
      for (int k = 0; k < UPT; ++k)
        if (FAST || cnt[k] > 0)
          store_unit<T, VEC>(dp + loff[k], r[k], FAST || full[k], cnt[k],
                             Pol::SPECIAL ? (spec >> (k * VEC)) &
                                                ((1u << VEC) - 1u)
                                          : 0u);
    }
  };
  if (any) store_members(0, ONE ? 1 : P);
  if constexpr (ONE && !FAST) {
    // The call's last unit, when it holds fewer than VEC elements, goes to
    // every member from here: the replicate copies whole 16-byte units
    // only.  One thread of the grid has it.
#pragma unroll
    for (int k = 0; k < UPT; ++k)
      if (cnt[k] > 0 && !full[k])
        for (int j = 1; j < P; ++j)
          store_unit<T, VEC>(d0 + j * step + loff[k], r[k], false, cnt[k],
                             Pol::SPECIAL ? (spec >> (k * VEC)) &
                                                ((1u << VEC) - 1u)
                                          : 0u);
  }

  // Special elements last: the stores above skip them, so their inputs
  // are still intact (also when dest == source), and the block-wide vote
  // sits after every wave's stores instead of between the read and the
  // write phase, where it held each wave's stores back until the slowest
  // wave of the workgroup had loaded all P slices.
  if constexpr (Pol::SPECIAL) {
    const bool sp = __syncthreads_or(spec != 0);
    if constexpr (ONE) {
      // this span's pieces of 256 units (ub is a multiple of 256)
      if (threadIdx.x < (uint32_t)UPT) flags[ub / kBlock + threadIdx.x] = sp;
      if (sp && any) store_members(1, P);
    }
    if (sp)
      special_block<T, Pol::OPC, VEC, Pol::RAB>(g, src_off, dest_off, ubase,
                                                spec, lg, n, lds);
  }
}

// The ragged last block-iteration of a call, [ub, min(ub + 256*UPT,
// units)): 256 units at a time through the guarded one-unit-per-thread
// body with 8 slices in flight.  Keeping the guarded code out of the
// UPT-wide body keeps the kernel's register footprint at that of the
// guard-free body, which sets how many workgroups stay resident per CU.
// (k_reduce gives each 256-unit piece its own workgroup instead.)
template <class T, class Pol, int VEC, int UPT, int MAXLV>
__device__ __forceinline__ void reduce_tail(const oscd_geom &g,
                                            const uint64_t dest_off,
                                            const uint64_t src_off,
                                            const uint64_t n, const int lg,
                                            const uint64_t ub, T *lds) {
  const uint64_t nunits = (n + VEC - 1) / VEC;
  const uint64_t span_end = ub + (uint64_t)kBlock * UPT;
  const uint64_t end = span_end < nunits ? span_end : nunits;
  for (uint64_t sub = ub; sub < end; sub += kBlock)
    reduce_block<false, T, Pol, VEC, 1, 8, MAXLV>(g, dest_off, src_off, n, lg,
                                                  sub, lds);
}

// Waves per SIMD the register allocator must leave room for: a policy's
// WAVES81 at the 8x1 shape, else no constraint.
template <class Pol, int UPT, int INFL>
constexpr int kernel_waves() {
  return (UPT == 8 && INFL == 1 && Pol::WAVES81 > 0) ? Pol::WAVES81 : 1;
}

template <class T, class Pol, int VEC, int UPT, int INFL, int MAXLV,
          bool ONE = false>
__global__ __launch_bounds__(kBlock) __attribute__((
    amdgpu_waves_per_eu(kernel_waves<Pol, UPT, INFL>()))) void k_reduce(const oscd_geom g,
                                                   const uint64_t dest_off,
                                                   const uint64_t src_off,
                                                   const uint64_t n,
                                                   const int lg,
                                                   uint8_t *const flags) {
  static_assert(!Pol::SPECIAL || UPT * VEC <= 32,
                "special-element mask is 32 bits");
  extern __shared__ unsigned char k_reduce_lds[];
  constexpr uint64_t kSpan = (uint64_t)kBlock * UPT;
  const uint64_t nunits = (n + VEC - 1) / VEC;
  const uint64_t nfull = n / VEC;
  // one full span per workgroup, then one workgroup per 256-unit piece of
  // the ragged rest (the grid covers both): straight-line code measures ~4%
  // faster than a grid-stride loop around the same body, and the pieces of
  // a small call's only span run side by side instead of one after another
  const uint64_t nspan = nfull / kSpan;
  const uint64_t b = blockIdx.x;
  if (b < nspan) {
    reduce_block<true, T, Pol, VEC, UPT, INFL, MAXLV, ONE>(
        g, dest_off, src_off, n, lg, b * kSpan, (T *)k_reduce_lds, flags);
  } else {
    const uint64_t sub = nspan * kSpan + (b - nspan) * kBlock;
    if (sub < nunits)
      reduce_block<false, T, Pol, VEC, 1, 8, MAXLV, ONE>(
          g, dest_off, src_off, n, lg, sub, (T *)k_reduce_lds, flags);
  }
}

// ---- small calls: lanes over (element, member group) -----------------------
// With few elements and many members the slice walk above is bound by
// latency: one lane per element reads the P slices 8 at a time, one HBM
// round trip per group of 8.  Here a group of G lanes (G a power of two,
// G <= min(P, 64)) shares one element: lane i of the group reads members
// i, i+G, ... (independent loads, in flight together) and folds them, then
// the group combines its G partial values with a butterfly of cross-lane
// moves, so every lane ends with the element's result and writes it to its
// own members.  Integer and bitwise operators only: their results are exact
// in any order, so they equal the recursive-doubling tree's bit for bit.

// Partner value for butterfly step STEP (1, 2, 4, ... 32).  Inside a row of
// 16 lanes the moves are DPP: quad_perm [1,0,3,2] (lane^1), quad_perm
// [2,3,0,1] (lane^2), row_half_mirror (lane^7) and row_mirror (lane^15).
// The mirrors stand for lane^4 and lane^8 because the steps before them
// left every quad (then every 8 lanes) holding one value.  Across rows:
// ds_swizzle in bit-mask mode, xor 16 within 32 lanes, and ds_bpermute from
// lane^32.
template <int STEP>
__device__ __forceinline__ uint32_t xlane32(uint32_t v) {
  const int x = (int)v;
  if constexpr (STEP == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
  } else if constexpr (STEP == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
  } else if constexpr (STEP == 4) {
    return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false);
  } else if constexpr (STEP == 8) {
    return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false);
  } else if constexpr (STEP == 16) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle(x, 0x401F);
  } else {
    static_assert(STEP == 32, "butterfly steps are 1, 2, 4, 8, 16, 32");
    return (uint32_t)__builtin_amdgcn_ds_bpermute(
        (int)(((threadIdx.x & 63u) ^ 32u) * 4u), x);
  }
}

template <int STEP, class T>
__device__ __forceinline__ T xlane(T v) {
  static_assert(std::is_integral<T>::value && sizeof(T) <= 8, "integers");
  if constexpr (sizeof(T) == 8) {
    uint64_t u = (uint64_t)v;
    const uint32_t lo = xlane32<STEP>((uint32_t)u);
    const uint32_t hi = xlane32<STEP>((uint32_t)(u >> 32));
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)xlane32<STEP>((uint32_t)v);
  }
}

template <class T, class Pol>
constexpr bool lanes_ok() {
  return std::is_integral<T>::value && !Pol::TREE && !Pol::SPECIAL;
}

// Loads in flight per lane (members i, i+G, ... up to 16 at a time).
constexpr int kLanesChunk = 16;

template <class T, class Pol, int G>
__global__ __launch_bounds__(kBlock) void k_reduce_lanes(
    const oscd_geom g, const uint64_t dest_off, const uint64_t src_off,
    const uint64_t n) {
  static_assert(G >= 16 && G <= 64 && (G & (G - 1)) == 0, "G in 16..64");
  constexpr int E = 64 / G;  // elements per wave
  const int lane = (int)(threadIdx.x & 63u);
  const int gi = lane & (G - 1);
  const uint64_t e =
      ((uint64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64) * E +
      (uint64_t)(lane / G);
  const bool act = e < n;  // the same for all G lanes of a group
  const int P = g.npes;    // >= G: every lane has member gi
  const uint64_t eb = e * sizeof(T);
  T acc = 0;
  if (act) {
    acc = *(const T *)(member_ptr(g, gi, src_off) + eb);
    for (int j0 = gi + G; j0 < P; j0 += kLanesChunk * G) {
      T x[kLanesChunk];
#pragma unroll
      for (int k = 0; k < kLanesChunk; ++k)
        if (j0 + k * G < P)
          x[k] = *(const T *)(member_ptr(g, j0 + k * G, src_off) + eb);
#pragma unroll
      for (int k = 0; k < kLanesChunk; ++k)
        if (j0 + k * G < P) acc = Pol::seq(acc, x[k]);
    }
  }
  // every lane of the wave takes part in the moves, active or not
  acc = Pol::comb(acc, xlane<1>(acc));
  acc = Pol::comb(acc, xlane<2>(acc));
  acc = Pol::comb(acc, xlane<4>(acc));
  acc = Pol::comb(acc, xlane<8>(acc));
  if constexpr (G >= 32) acc = Pol::comb(acc, xlane<16>(acc));
  if constexpr (G >= 64) acc = Pol::comb(acc, xlane<32>(acc));
  if (act) {
    const T r = Pol::out(acc);
    for (int j = gi; j < P; j += G) *(T *)(member_ptr(g, j, dest_off) + eb) = r;
  }
}

// Group size for a team of P >= 16 members.
__host__ __device__ constexpr int lanes_group(int P) {
  return P >= 64 ? 64 : P >= 32 ? 32 : 16;
}

template <class T, class Pol>
static int launch_lanes(const oscd_geom *g, uint64_t doff, uint64_t soff,
                        uint64_t n, hipStream_t s) {
  static_assert(lanes_ok<T, Pol>(), "integer and bitwise operators only");
  const int P = g->npes;
  const uint64_t al = sizeof(T);
  // n <= 2^24 keeps the grid (n/4 workgroups at most) far inside HSA's
  // 32-bit work-item count
  if (P < 16 || n == 0 || n > (1ull << 24) || doff % al != 0 ||
      soff % al != 0 || g->slice % al != 0 || (uintptr_t)g->base % al != 0)
    return OSC_ERR_ARG;
  const int G = lanes_group(P);
  const uint64_t per_block = (uint64_t)(kBlock / 64) * (64 / G);
  const unsigned grid = (unsigned)((n + per_block - 1) / per_block);
  if (G == 64) {
    hipLaunchKernelGGL((k_reduce_lanes<T, Pol, 64>), dim3(grid), dim3(kBlock),
                       0, s, *g, doff, soff, n);
    note_launch("k_reduce_lanes", k_reduce_lanes_pf<T, Pol, 64>());
  } else if (G == 32) {
    hipLaunchKernelGGL((k_reduce_lanes<T, Pol, 32>), dim3(grid), dim3(kBlock),
                       0, s, *g, doff, soff, n);
    note_launch("k_reduce_lanes", k_reduce_lanes_pf<T, Pol, 32>());
  } else {
    hipLaunchKernelGGL((k_reduce_lanes<T, Pol, 16>), dim3(grid), dim3(kBlock),
                       0, s, *g, doff, soff, n);
    note_launch("k_reduce_lanes", k_reduce_lanes_pf<T, Pol, 16>());
  }
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

// ---- small float calls: lanes gather, one fold in team order ---------------
// Floating-point sums and products keep the slice walk's order (team index
// 0, 1, ..., P-1, or the reference's tree in reference mode), so a call
// gives the same bits on every path (alone, batched, served).  Here the 64
// lanes of a wave read one element's P members in parallel into LDS
// (independent loads, in flight together), and every lane then folds the P
// values in the walk's order with the policy's own operations, without the
// walk's dependent passes over the slices.
template <class T, class Pol>
constexpr bool gather_ok() {
  return !std::is_integral<T>::value;
}
// LDS holds 4 waves x P values: at most 64 KiB for 16-byte elements
constexpr int kGatherMaxPes = 1024;
constexpr int kGatherChunk = 8;  // loads in flight per lane

template <class T, class Pol>
__global__ __launch_bounds__(kBlock) void k_reduce_gather(
    const oscd_geom g, const uint64_t dest_off, const uint64_t src_off,
    const uint64_t n, const int lg) {
  // values travel as raw words of the element's size (a complex struct in
  // a local array would live in scratch memory)
  typedef typename Raw<(int)sizeof(T)>::type R;
  extern __shared__ unsigned char k_reduce_gather_lds[];
  const int lane = (int)(threadIdx.x & 63u);
  const int w = (int)(threadIdx.x >> 6);
  const uint64_t e = (uint64_t)blockIdx.x * (kBlock / 64) + (uint64_t)w;
  const int P = g.npes;
  R *x = (R *)k_reduce_gather_lds + (size_t)w * (size_t)P;
  const bool act = e < n;  // the same for the whole wave
  const uint64_t eb = e * sizeof(T);
  if (act) {
    for (int j0 = lane; j0 < P; j0 += 64 * kGatherChunk) {
      R v[kGatherChunk];
#pragma unroll
      for (int k = 0; k < kGatherChunk; ++k)
        if (j0 + 64 * k < P)
          v[k] = ld<(int)sizeof(T)>(member_ptr(g, j0 + 64 * k, src_off) + eb);
#pragma unroll
      for (int k = 0; k < kGatherChunk; ++k)
        if (j0 + 64 * k < P) x[j0 + 64 * k] = v[k];
    }
  }
  // every wave reaches the barrier; all loads of an element precede it and
  // all its stores follow it, so dest == source is safe
  __syncthreads();
  if (act) {
    auto value = [&](int j) {
      const R raw = x[j];
      T t;
      __builtin_memcpy(&t, &raw, sizeof(T));
      return t;
    };
    typename Pol::Acc acc;
    bool nan = false;
    if constexpr (Pol::TREE) {
      // the reference's order: member 0's recursive-doubling tree (leaves
      // in team order, a folded pair as op(x[m+1], x[m]), merged by a
      // binary counter), which every member's tree equals for + and *
      // (constant stack indices and unconditional selects, as in the
      // walk's tree, keep the stack in registers; lg <= 10 here)
      constexpr int kLv = 11;
      const int p2s = 1 << lg;
      typename Pol::Acc stk[kLv];
      if constexpr (Pol::BINOM) {
        // the binomial order: all P members are leaves, a node is
        // op(right subtree, left subtree), the subtrees left by a
        // non-power-of-two P merged from the lowest up (as in the walk)
        for (int j = 0; j < P; ++j) {
          typename Pol::Acc leaf = Pol::lift(value(j));
          bool carry = true;
#pragma unroll
          for (int L = 0; L < kLv; ++L) {
            const bool occupied = ((j >> L) & 1) != 0;
            if (carry && occupied) leaf = Pol::comb(leaf, stk[L]);
            stk[L] = pick(carry && !occupied, leaf, stk[L]);
            carry = carry && occupied;
          }
        }
        bool have = false;
#pragma unroll
        for (int L = 0; L < kLv; ++L)
          if ((P >> L) & 1) {
            if (have)
              acc = Pol::comb(acc, stk[L]);
            else
              acc = stk[L];
            have = true;
          }
      } else {
      for (int j = 0; j < p2s; ++j) {
        const int m = member_of(j, P, lg);
        const int m1 = member_of(j + 1, P, lg);
        typename Pol::Acc leaf =
            (m1 - m == 2) ? Pol::comb(Pol::lift(value(m + 1)), Pol::lift(value(m)))
                          : Pol::lift(value(m));
        bool carry = true;
#pragma unroll
        for (int L = 0; L < kLv; ++L) {
          const bool occupied = ((j >> L) & 1) != 0;
          if (carry && occupied) leaf = Pol::comb(stk[L], leaf);
          stk[L] = pick(carry && !occupied, leaf, stk[L]);
          carry = carry && occupied;
        }
      }
#pragma unroll
      for (int L = 0; L < kLv; ++L)
        if (L == lg) acc = stk[L];
      }
    } else {
      acc = Pol::lift(value(0));
      nan = is_nan(value(0));
      for (int j = 1; j < P; ++j) {
        const T v = value(j);
        if (Pol::SPECIAL) nan |= is_nan(v);
        acc = Pol::seq(acc, v);
      }
    }
    const T r = Pol::out(acc);
    bool special = false;
    if constexpr (Pol::SPECIAL) special = nan || is_nan_or_zero(r);
    if constexpr (Pol::SPECIAL) {
      if (special) {
      // float MIN/MAX with a NaN or a signed-zero extremum: each member
      // gets its own recursive-doubling result, evaluated over the LDS copy
      // as the slice walk's special path does (the element is the same
      // for the whole wave, so this branch is uniform)
      const int p2s = 1 << lg;
      const T *snap = (const T *)x;  // float or double: the raw words
      if constexpr (Pol::RAB) {
        // Rabenseifner: the owner's tree, to every member
        const T rk = member_tree<T, Pol::OPC>(snap, rab_owner(e, n, lg), P,
                                              lg, e >= n / 2);
        for (int j = lane; j < P; j += 64)
          *(T *)(member_ptr(g, j, dest_off) + eb) = rk;
      } else {
        for (int k = lane; k < p2s; k += 64) {
          const T rk = member_tree<T, Pol::OPC>(snap, k, P, lg, false);
          const int mk = member_of(k, P, lg);
          const int mk1 = member_of(k + 1, P, lg);
          *(T *)(member_ptr(g, mk, dest_off) + eb) = rk;
          if (mk1 - mk == 2)
            *(T *)(member_ptr(g, mk + 1, dest_off) + eb) = rk;
        }
      }
      }
    }
    if (!special) {
      R rr;
      __builtin_memcpy(&rr, &r, sizeof(T));
      for (int j = lane; j < P; j += 64)
        st<(int)sizeof(T)>(member_ptr(g, j, dest_off) + eb, rr);
    }
  }
}

template <class T, class Pol>
static int launch_gather(const oscd_geom *g, int lg, uint64_t doff,
                         uint64_t soff, uint64_t n, hipStream_t s) {
  static_assert(gather_ok<T, Pol>(), "float and complex types only");
  const int P = g->npes;
  const uint64_t al = sizeof(T);
  if (P < 16 || P > kGatherMaxPes || n == 0 || n > (1ull << 24) ||
      doff % al != 0 || soff % al != 0 || g->slice % al != 0 ||
      (uintptr_t)g->base % al != 0)
    return OSC_ERR_ARG;
  const unsigned grid = (unsigned)((n + (kBlock / 64) - 1) / (kBlock / 64));
  const size_t lds = (size_t)(kBlock / 64) * (size_t)P * sizeof(T);
  hipLaunchKernelGGL((k_reduce_gather<T, Pol>), dim3(grid), dim3(kBlock), lds,
                     s, *g, doff, soff, n, lg);
  note_launch("k_reduce_gather", k_reduce_gather_pf<T, Pol>());
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

// When the default launch takes the lanes kernel: teams of 16 members or
// more and at most 2048 elements.  Measured on MI355X (one launch, 64-bit
// and 32-bit sums, P = 16..256): the lanes kernel stays at the launch floor
// up to n = 2048 for every P, where the slice walk grows with P (P = 64,
// n = 1: 11.7 vs 6.2 us; P = 256: 34 vs 6.3 us); from n = 8192 on the
// walk's 16-byte coalesced loads usually win.
constexpr int kLanesMinPes = 16;
constexpr uint64_t kLanesMaxElems = 2048;
template <class T>
static bool lanes_pays(const oscd_geom *g, uint64_t doff, uint64_t soff,
                       uint64_t n) {
  const uint64_t al = sizeof(T);
  return g->npes >= kLanesMinPes && n <= kLanesMaxElems && doff % al == 0 &&
         soff % al == 0 && g->slice % al == 0 &&
         (uintptr_t)g->base % al == 0;
}

// The gather kernel's cut is in bytes per member: measured (one launch,
// f32/f64 sums and products, complex sums and products, P = 16..1024) it
// wins up to 2048 doubles at 64+ members (P = 256, n = 1: 41 vs 9 us;
// P = 1024: 153 vs 25 us), while 2048 complex doubles (32 KiB) and 8192
// doubles already run faster through the walk's coalesced loads.
constexpr uint64_t kGatherMaxBytes = 16384;
template <class T>
static bool gather_pays(const oscd_geom *g, uint64_t doff, uint64_t soff,
                        uint64_t n) {
  const uint64_t al = sizeof(T);
  return g->npes >= kLanesMinPes && g->npes <= kGatherMaxPes &&
         n <= kGatherMaxBytes / sizeof(T) && doff % al == 0 && soff % al == 0 &&
         g->slice % al == 0 && (uintptr_t)g->base % al == 0;
}

// ---- host-side launch ------------------------------------------------------
template <class T>
constexpr int vec_of() {
  return (16 / (int)sizeof(T)) > 8 ? 8 : (16 / (int)sizeof(T));
}

template <class T, class Pol, int VEC, int UPT, int INFL, bool ONE = false>
static void launch_shape(const oscd_geom *g, int lg, uint64_t doff,
                         uint64_t soff, uint64_t n, size_t lds,
                         hipStream_t s, uint8_t *flags = nullptr) {
  const uint64_t units = (n + VEC - 1) / VEC;
  const uint64_t span = (uint64_t)kBlock * UPT;
  // every full span gets its own workgroup, and every 256 units of the
  // ragged rest; launch() bounds n so that the grid stays within 2^23
  // workgroups
  const uint64_t nspan = (n / VEC) / span;
  const uint64_t grid = nspan + (units - nspan * span + kBlock - 1) / kBlock;
  if constexpr (Pol::TREE) {
    if (lg >= 8) {
      hipLaunchKernelGGL((k_reduce<T, Pol, VEC, UPT, INFL, 13, ONE>),
                         dim3((unsigned)grid), dim3(kBlock), lds, s, *g, doff,
                         soff, n, lg, flags);
      note_launch("k_reduce", k_reduce_pf<T, Pol, VEC, UPT, INFL, 13, ONE>());
      return;
    }
  }
  hipLaunchKernelGGL((k_reduce<T, Pol, VEC, UPT, INFL, Pol::TREE ? 8 : 1, ONE>),
                     dim3((unsigned)grid), dim3(kBlock), lds, s, *g, doff,
                     soff, n, lg, flags);
  note_launch("k_reduce",
              k_reduce_pf<T, Pol, VEC, UPT, INFL, Pol::TREE ? 8 : 1, ONE>());
}

// Large calls run in two kernels: the slice walk stores the result to
// member 0 only, then oscd_replicate copies it to the other members.  The
// walk then only reads (plus 1/P of writes) and the replicate only writes,
// and HBM serves each phase faster than it serves the two mixed: measured
// on MI355X at 64 members x 4M elements, the walk to member 0 reads at
// 6.8-7.2 TB/s for every element type (u64 xor 313 us, f64 sum 329 us at
// 8x2, f32 sum 159 us), while a workgroup that fans its span out to all 64
// members writes at 5.1-5.3 TB/s and a fused u64 cell takes 735 us; the
// replicate writes in an interleaved order that reaches 6.3-6.8 TB/s.
// Every type whose 16-byte unit holds whole elements (2, 4, 8 and 16
// bytes) takes it, from 1 MiB per member and 32 members (3 MiB per member
// and 8 members).  The replicate
// copies whole units; a last unit with fewer elements is stored to every
// member by the walk.  The split costs one extra launch and (P-1)/G cached
// re-reads of member 0's result.
// Float MIN/MAX (SPECIAL) split too: the walk flags every 256-unit piece of
// a span that holds a special element (whose members get per-member
// results), stores those spans to every member itself, and the replicate
// skips the flagged pieces.  The flags live in a per-stream device buffer
// (span_flags); without one (a stream under graph capture) the call runs
// as one kernel.
// Teams of fewer than kSplitWidePes members split from kSplitNarrowBytes
// per member: below it their walk and replicate are too short to pay for
// the second launch.  Measured on MI355X per call, split vs one kernel
// alternated in one process (the batched workload's world, 128 members
// about 1 GiB apart): 8-31 members at 1, 1.5 and 2-3 MiB per member
// 2.7-3.1 vs 3.3-3.5, 3.4-4.0 vs 4.3-4.8 and 3.9-4.0 vs 4.1-4.3 TB/s, at
// 3-4 MiB even; 32 members and more at 1 MiB 4.0-4.9 vs 4.0-4.6 and at
// 2-3 MiB 5.0-5.6 vs 4.9-5.2 TB/s.
constexpr int kSplitMinPes = 8, kSplitWidePes = 32;
constexpr uint64_t kSplitMinBytes = 1ull << 20;  // per member
constexpr uint64_t kSplitNarrowBytes = 3ull << 20;
// 0: no split; -1: replicate group picked by size; else the group size
extern int g_split_group;
// false: calls whose bytes per member are not a multiple of 16 run as one
// kernel (oscd_set_split_partial)
extern bool g_split_partial;
template <class T, class Pol>
static bool split_pays(const oscd_geom *g, uint64_t n) {
  return g_split_group != 0 && vec_of<T>() * sizeof(T) == 16 &&
         g->npes >= kSplitMinPes && n * sizeof(T) >= kSplitMinBytes &&
         (g->npes >= kSplitWidePes || n * sizeof(T) >= kSplitNarrowBytes) &&
         (g_split_partial || (n * sizeof(T)) % 16 == 0);
}
// One flag byte per 256-unit piece for a split float MIN/MAX call on stream
// s, or nullptr when the stream has none to give (under graph capture, more
// pieces than the heap slice holds, or no memory).
uint8_t *span_flags(hipStream_t s, uint64_t npieces);
// The replicate of a split call, with the group picked by measurement on
// the first call of its (team size, size class) when g_split_group is -1.
int replicate_tuned(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                    const uint8_t *skip, hipStream_t s);

// Default shape of the walk of a split call: Pol::SHAPE_ONE where the
// policy names one, else Pol::SHAPE.
template <class Pol, class = void>
struct shape_one {
  static constexpr int v = Pol::SHAPE;
};
template <class Pol>
struct shape_one<Pol, std::void_t<decltype(Pol::SHAPE_ONE)>> {
  static constexpr int v = Pol::SHAPE_ONE;
};

// Default shape per policy (Pol::SHAPE), from the 64-PE x 4M sweeps on
// MI355X: 8x2 for the order-free kernels, 8x1 for FP32 accumulated in FP64
// (its FP64 accumulators make 8x2 too register-heavy), 2x4 for the
// reference tree (its subtree stack takes registers).
template <class Pol>
constexpr int default_shape() {
  return Pol::SHAPE;
}
// The reference tree only runs the narrow shapes (its subtree stack takes
// MAXLV*UPT*VEC registers).
template <class Pol>
constexpr int clamp_shape(int shape) {
  return (shape < 0 || shape >= kNumShapes)
             ? default_shape<Pol>()
             : (Pol::TREE ? (shape == 3 ? 3 : 2) : shape);
}

// Walk shape of a split call when the library picks it, by bytes per
// member m (g_split_walk_rule, on by default; off: the policy's split shape
// from 8 MiB, 1x8 below).  The walk's grid is one workgroup per span, so
// the 8-unit shapes (32 KiB of each slice per workgroup) launch m / 32 KiB
// workgroups: 256 at 8 MiB, one per CU.  Measured on MI355X, every variant
// alternated in one process (a rebuilt library moves the heap's physical
// pages, which alone moves these rates by a few percent): on heaps whose
// members lie 6-36 MiB apart, 32-128 members, whole calls with 2x4 against
// the 1x8 pick at m = 4 MiB: i64 sum 5.24-6.14 vs 4.90-5.77 TB/s, f64 sum
// 5.38-6.38 vs 5.10-5.70, f32 sum 5.23-6.05 vs 4.94-5.77, f64 min
// 5.14-6.06 vs 4.97-5.77; at 2 MiB +5-8% except FP64 Sum2 (1x8 stays:
// 5.09-5.33 vs 4.51-5.00); at 8 MiB i64 sum +3-6% over 8x2, while the
// FP32/FP64 accumulated sums keep 8x2/8x1 (2x4: -3-6%) and f64 min/max
// takes 8x1 (+3%).  On the
// batched workload's world (128 members about 1 GiB apart) the same calls
// gain 0.7% overall (f32 at 2-4 MiB 3-9%), the rest within 1.5%.
extern bool g_split_walk_rule;
template <class T, class Pol>
static int split_walk_shape(uint64_t n, int picked) {
  if constexpr (Pol::TREE || Pol::RAB || !std::is_arithmetic<T>::value) {
    return picked;
  } else {
    if (!g_split_walk_rule) return picked;
    // FP64 SUM/PROD in twice the working precision
    constexpr bool dd = std::is_same<T, double>::value &&
                        !std::is_same<typename Pol::Acc, double>::value;
    const uint64_t m = n * sizeof(T);
    if (m < (2ull << 20)) return 3;
    if (m < (4ull << 20)) return dd ? 3 : 2;
    if (m < (8ull << 20)) return 2;
    if (m < (16ull << 20) && std::is_integral<T>::value) return 2;
    if (m < (16ull << 20) && Pol::SPECIAL && std::is_same<T, double>::value)
      return 1;
    return clamp_shape<Pol>(shape_one<Pol>::v);
  }
}

template <class T, class Pol>
static int launch(const oscd_geom *g, int lg, uint64_t doff, uint64_t soff,
                  uint64_t n, int shape, hipStream_t s) {
  constexpr int VB = vec_of<T>();
  constexpr uint64_t AL = VB * sizeof(T);
  const bool vec_ok = (doff % AL == 0) && (soff % AL == 0) &&
                      (g->slice % AL == 0) && ((uintptr_t)g->base % AL == 0);
  const size_t lds = Pol::SPECIAL ? (size_t)g->npes * sizeof(T) : 0;
  if (n == 0) return OSC_OK;
  // one workgroup per span of >= 256 units: at most 2^31 units keeps the
  // grid within 2^23 workgroups (the reference itself takes nreduce < 2^31)
  if ((n + (vec_ok ? VB : 1) - 1) / (vec_ok ? VB : 1) > (1ull << 31))
    return OSC_ERR_ARG;
  const bool chosen = shape < 0;  // the shape is the library's to pick
  const int phase = t_split_phase;
  if (shape < 0) {
    if constexpr (lanes_ok<T, Pol>()) {
      if (lanes_pays<T>(g, doff, soff, n))
        return phase == 1 ? OSC_OK : launch_lanes<T, Pol>(g, doff, soff, n, s);
    } else if constexpr (gather_ok<T, Pol>()) {
      if (gather_pays<T>(g, doff, soff, n))
        return phase == 1 ? OSC_OK
                          : launch_gather<T, Pol>(g, lg, doff, soff, n, s);
    }
    // Small calls: at the default shape they would fill fewer workgroups
    // than the chip has CUs; one unit per thread with 8 slices in flight
    // spreads the same work over UPT times more workgroups.
    const int d = default_shape<Pol>();
    const uint64_t spans = ((n + VB - 1) / VB) / ((uint64_t)kBlock * shape_upt(d));
    shape = spans < kSmallSpans ? 3 : d;
  }
  shape = clamp_shape<Pol>(shape);
  if (vec_ok && split_pays<T, Pol>(g, n)) {
    if (chosen && shape == default_shape<Pol>())
      shape = clamp_shape<Pol>(shape_one<Pol>::v);
    if (chosen) shape = split_walk_shape<T, Pol>(n, shape);
    uint8_t *fl = nullptr;
    if constexpr (Pol::SPECIAL)
      fl = phase < 0
               ? span_flags(s, ((n * sizeof(T) + 15) / 16 + kBlock - 1) / kBlock)
               : t_split_flags;
    if (!Pol::SPECIAL || fl != nullptr) {
      if (phase == 1)
        return replicate_tuned(g, doff, n * sizeof(T) / 16 * 16, fl, s);
      if constexpr (Pol::TREE) {
        if (shape == 3)
          launch_shape<T, Pol, VB, shape_upt(3), shape_infl(3), true>(g, lg, doff, soff, n, lds, s, fl);
        else
          launch_shape<T, Pol, VB, shape_upt(2), shape_infl(2), true>(g, lg, doff, soff, n, lds, s, fl);
      } else {
        switch (shape) {
          case 0: launch_shape<T, Pol, VB, shape_upt(0), shape_infl(0), true>(g, lg, doff, soff, n, lds, s, fl); break;
          case 1: launch_shape<T, Pol, VB, shape_upt(1), shape_infl(1), true>(g, lg, doff, soff, n, lds, s, fl); break;
          case 2: launch_shape<T, Pol, VB, shape_upt(2), shape_infl(2), true>(g, lg, doff, soff, n, lds, s, fl); break;
          default: launch_shape<T, Pol, VB, shape_upt(3), shape_infl(3), true>(g, lg, doff, soff, n, lds, s, fl); break;
        }
      }
      if (hipGetLastError() != hipSuccess) return OSC_ERR_HIP;
      if (phase == 0) return OSC_OK;
      return replicate_tuned(g, doff, n * sizeof(T) / 16 * 16, fl, s);
    }
  }
  if (phase == 1) return OSC_OK;  // one kernel: issued whole under phase 0
  if (!vec_ok) {
    launch_shape<T, Pol, 1, 1, 8>(g, lg, doff, soff, n, lds, s);
  } else if constexpr (Pol::TREE) {
    if (shape == 3)
      launch_shape<T, Pol, VB, shape_upt(3), shape_infl(3)>(g, lg, doff, soff, n, lds, s);
    else
      launch_shape<T, Pol, VB, shape_upt(2), shape_infl(2)>(g, lg, doff, soff, n, lds, s);
  } else {
    switch (shape) {
      case 0: launch_shape<T, Pol, VB, shape_upt(0), shape_infl(0)>(g, lg, doff, soff, n, lds, s); break;
      case 1: launch_shape<T, Pol, VB, shape_upt(1), shape_infl(1)>(g, lg, doff, soff, n, lds, s); break;
      case 2: launch_shape<T, Pol, VB, shape_upt(2), shape_infl(2)>(g, lg, doff, soff, n, lds, s); break;
      default: launch_shape<T, Pol, VB, shape_upt(3), shape_infl(3)>(g, lg, doff, soff, n, lds, s); break;
    }
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

// ---- batched tile loop -------------------------------------------------------
// Workgroup b processes tiles b, b+grid, ...; each tile is whole block
// iterations of one call (the planner aligns tile bounds to 256*UPT units),
// so the per-call kernel body runs unchanged.
template <class T, class Pol, int VEC, int UPT, int INFL, int MAXLV>
__global__ __launch_bounds__(kBlock) void k_batch_reduce(
    const oscd_bcall *__restrict__ calls, const oscd_tile *__restrict__ tiles,
    const uint32_t ntiles, char *const base, const uint64_t slice) {
  extern __shared__ unsigned char k_batch_lds[];
  constexpr uint64_t kSpan = (uint64_t)kBlock * UPT;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const oscd_tile tl = tiles[t];
    const oscd_bcall c = calls[tl.call];
    oscd_geom g;
    g.base = base;
    g.slice = slice;
    g.start = c.start;
    g.stride = c.stride;
    g.npes = c.npes;
    g.pad_ = 0;
    const uint64_t n = c.count;
    const uint64_t nfull = n / VEC;
    // tiles start on span boundaries; only a call's last span can be
    // ragged, and it ends the tile that holds it
    uint64_t ub = tl.begin;
    for (; ub < tl.end && ub + kSpan <= nfull; ub += kSpan)
      reduce_block<true, T, Pol, VEC, UPT, INFL, MAXLV>(
          g, c.dest_off, c.src_off, n, c.lg, ub, (T *)k_batch_lds);
    if (ub < tl.end)
      reduce_tail<T, Pol, VEC, UPT, MAXLV>(g, c.dest_off, c.src_off, n, c.lg,
                                           ub, (T *)k_batch_lds);
  }
}

struct BatchArgs {
  const oscd_bcall *calls;
  const oscd_tile *tiles;
  uint32_t ntiles, grid;
  size_t lds;
  char *base;
  uint64_t slice;
  int maxlg;
  hipStream_t s;
};

template <class T, class Pol, int UPT, int INFL>
static void batch_shape(const BatchArgs &a) {
  constexpr int VB = vec_of<T>();
  if constexpr (Pol::TREE) {
    if (a.maxlg >= 8) {
      hipLaunchKernelGGL((k_batch_reduce<T, Pol, VB, UPT, INFL, 13>),
                         dim3(a.grid), dim3(kBlock), a.lds, a.s, a.calls,
                         a.tiles, a.ntiles, a.base, a.slice);
      return;
    }
  }
  hipLaunchKernelGGL((k_batch_reduce<T, Pol, VB, UPT, INFL, Pol::TREE ? 8 : 1>),
                     dim3(a.grid), dim3(kBlock), a.lds, a.s, a.calls, a.tiles,
                     a.ntiles, a.base, a.slice);
}

// Batched families run the widest and the narrowest shape only (large
// calls vs. calls of a few thousand elements).
template <class T, class Pol>
static int batch_launch(const BatchArgs &a, int shape) {
  if (a.ntiles == 0) return OSC_OK;
  shape = clamp_shape<Pol>(shape);
  if constexpr (Pol::TREE) {
    if (shape == 3)
      batch_shape<T, Pol, shape_upt(3), shape_infl(3)>(a);
    else
      batch_shape<T, Pol, shape_upt(2), shape_infl(2)>(a);
  } else {
    if (shape == 3)
      batch_shape<T, Pol, shape_upt(3), shape_infl(3)>(a);
    else
      batch_shape<T, Pol, shape_upt(0), shape_infl(0)>(a);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

// Units a workgroup covers per block iteration at a shape, as the
// launchers above will run it.
template <class Pol>
static int span_units(int shape, bool batched) {
  shape = clamp_shape<Pol>(shape);
  if (batched && !Pol::TREE && shape != 3) shape = 0;
  return kBlock * shape_upt(shape);
}

// ---- policy selection ----------------------------------------------------------
// v.template run<T, Pol>() for the kernel policy of (dtype, op, mode).
// Integer bitwise/sum/prod use the unsigned type of the same width
// (two's-complement wrap, as the reference's C arithmetic on x86-64);
// MIN/MAX compare with the element type's own signedness.
template <class S, class U, class V>
static int visit_int_t(int op, V &v) {
  switch (op) {
    case OSC_OP_AND: return v.template run<U, PolSeq<U, OSC_OP_AND, false>>();
    case OSC_OP_OR: return v.template run<U, PolSeq<U, OSC_OP_OR, false>>();
    case OSC_OP_XOR: return v.template run<U, PolSeq<U, OSC_OP_XOR, false>>();
    case OSC_OP_SUM: return v.template run<U, PolSeq<U, OSC_OP_SUM, false>>();
    case OSC_OP_PROD: return v.template run<U, PolSeq<U, OSC_OP_PROD, false>>();
    case OSC_OP_MIN: return v.template run<S, PolSeq<S, OSC_OP_MIN, false>>();
    case OSC_OP_MAX: return v.template run<S, PolSeq<S, OSC_OP_MAX, false>>();
  }
  return OSC_ERR_ARG;
}

// SUM/PROD of a float or complex type by algorithm (float mode):
//   accurate (hip): wider accumulation, rounded once;
//   rec_dbl and rabenseifner: the recursive-doubling tree in the type's
//     own arithmetic (Rabenseifner's reduce-scatter gives each element its
//     owner's butterfly, whose + and * results equal every member's);
//   linear: team-order fold op(acc, x_i) (reduce.c:55-91);
//   binomial: the binomial tree (PolBinom).
template <class T, template <int> class AccPol, int OP, class V>
static int visit_arith(int mode, V &v) {
  switch (mode) {
    case OSC_FP_ACCURATE: return v.template run<T, AccPol<OP>>();
    case OSC_FP_LINEAR: return v.template run<T, PolSeq<T, OP, false>>();
    case OSC_FP_BINOMIAL: return v.template run<T, PolBinom<T, OP>>();
    case OSC_FP_REFERENCE:
    case OSC_FP_RABENSEIFNER: return v.template run<T, PolTree<T, OP>>();
  }
  return OSC_ERR_ARG;
}

// MIN/MAX of a float type: exact selection everywhere except elements with
// a NaN or a signed-zero extremum, where the operand order of the ternary
// decides: per-member recursive-doubling trees (accurate, rec_dbl), the
// owner's tree (rabenseifner), the team-order fold (linear) or the
// binomial tree, each as its reference algorithm leaves it.
template <class T, int OP, class V>
static int visit_sel(int mode, V &v) {
  switch (mode) {
    case OSC_FP_ACCURATE:
    case OSC_FP_REFERENCE: return v.template run<T, PolSeq<T, OP, true>>();
    case OSC_FP_LINEAR: return v.template run<T, PolSeq<T, OP, false>>();
    case OSC_FP_BINOMIAL: return v.template run<T, PolBinom<T, OP>>();
    case OSC_FP_RABENSEIFNER: return v.template run<T, PolSeqRab<T, OP>>();
  }
  return OSC_ERR_ARG;
}

template <class T, template <int> class AccPol, class V>
static int visit_real_t(int op, int mode, V &v) {
  switch (op) {
    case OSC_OP_MIN: return visit_sel<T, OSC_OP_MIN>(mode, v);
    case OSC_OP_MAX: return visit_sel<T, OSC_OP_MAX>(mode, v);
    case OSC_OP_SUM: return visit_arith<T, AccPol, OSC_OP_SUM>(mode, v);
    case OSC_OP_PROD: return visit_arith<T, AccPol, OSC_OP_PROD>(mode, v);
  }
  return OSC_ERR_UNSUPPORTED;
}

template <class T, template <int> class AccPol, class V>
static int visit_cplx_t(int op, int mode, V &v) {
  switch (op) {
    case OSC_OP_SUM: return visit_arith<T, AccPol, OSC_OP_SUM>(mode, v);
    case OSC_OP_PROD: return visit_arith<T, AccPol, OSC_OP_PROD>(mode, v);
  }
  return OSC_ERR_UNSUPPORTED;
}

// Visitors: per-call launch, batched launch, tile-span query.
struct PerCallV {
  const oscd_geom *g;
  int lg;
  uint64_t doff, soff, n;
  int shape;
  hipStream_t s;
  template <class T, class Pol>
  int run() {
    return launch<T, Pol>(g, lg, doff, soff, n, shape, s);
  }
};
struct BatchV {
  BatchArgs a;
  int shape;
  template <class T, class Pol>
  int run() {
    return batch_launch<T, Pol>(a, shape);
  }
};
struct SpanV {
  int shape;
  template <class T, class Pol>
  int run() {
    return span_units<Pol>(shape, true);
  }
};

// Family entry points, one translation unit each (compiled in parallel).
int visit_reduce_int_wide(int dtype, int op, int mode, PerCallV *pc, BatchV *bv, SpanV *sv);
int visit_reduce_int_narrow(int dtype, int op, int mode, PerCallV *pc, BatchV *bv, SpanV *sv);
int visit_reduce_f32(int dtype, int op, int mode, PerCallV *pc, BatchV *bv, SpanV *sv);
int visit_reduce_f64(int dtype, int op, int mode, PerCallV *pc, BatchV *bv, SpanV *sv);
int visit_reduce_cplx(int dtype, int op, int mode, PerCallV *pc, BatchV *bv, SpanV *sv);

// Exactly one of pc / bv / sv is non-null.
#define OSC_VISIT_FAMILY(NAME, ...)                                           \
  int NAME(int dtype, int op, int mode, PerCallV *pc, BatchV *bv,             \
           SpanV *sv) {                                                       \
    (void)mode;                                                               \
    if (pc) { PerCallV &v = *pc; __VA_ARGS__ }                                \
    if (bv) { BatchV &v = *bv; __VA_ARGS__ }                                  \
    SpanV &v = *sv;                                                           \
    __VA_ARGS__                                                               \
  }

int visit_reduce(int dtype, int op, int mode, PerCallV *pc, BatchV *bv,
                 SpanV *sv);

}  // namespace osc

#endif  // OSC_DEV_REDUCE_H
