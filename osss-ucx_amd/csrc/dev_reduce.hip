// dev_reduce.hip - the team reduction kernel (all ops x all element types).
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
// Layout: one thread owns VEC consecutive elements (16 bytes) of the
// reduced range, reads them from all P source slices with 16-byte loads
// (8 loads in flight), combines, and writes the 16-byte result to all P
// destination slices with streaming stores.  Every source byte is read
// once and every destination byte written once: HBM traffic = 2*P*n*s.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

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
  static constexpr bool SPECIAL = SPEC;
  static constexpr int OPC = OP;
  __device__ static Acc lift(T x) { return x; }
  __device__ static Acc seq(Acc a, T x) { return Bin<OP>::f(a, x); }
  __device__ static Acc comb(Acc a, Acc b) { return Bin<OP>::f(a, b); }
  __device__ static T out(Acc a) { return a; }
};

// The type's own arithmetic in the reference's tree order.
template <class T, int OP>
struct PolTree {
  typedef T Acc;
  static constexpr bool TREE = true;
  static constexpr bool SPECIAL = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(T x) { return x; }
  __device__ static Acc seq(Acc a, T x) { return Bin<OP>::f(a, x); }
  __device__ static Acc comb(Acc a, Acc b) { return Bin<OP>::f(a, b); }
  __device__ static T out(Acc a) { return a; }
};

// FP32 SUM/PROD accumulated in FP64, rounded once.
template <int OP>
struct PolF32Acc {
  typedef double Acc;
  static constexpr bool TREE = false;
  static constexpr bool SPECIAL = false;
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

// FP64 SUM/PROD accumulated in double-double, rounded once.
template <int OP>
struct PolF64Acc {
  typedef dd Acc;
  static constexpr bool TREE = false;
  static constexpr bool SPECIAL = false;
  static constexpr int OPC = OP;
  __device__ static Acc lift(double x) { return {x, 0.0}; }
  __device__ static Acc seq(Acc a, double x) {
    return OP == OSC_OP_SUM ? dd_add_d(a, x) : dd_mul_d(a, x);
  }
  __device__ static Acc comb(Acc a, Acc b) { return dd_add(a, b); }
  __device__ static double out(Acc a) { return a.hi; }
};

// complex float SUM/PROD in FP64 components.
struct c64acc {
  double re, im;
};
template <int OP>
struct PolC32Acc {
  typedef c64acc Acc;
  static constexpr bool TREE = false;
  static constexpr bool SPECIAL = false;
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
  static constexpr bool SPECIAL = false;
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

// ---- unit load / store ---------------------------------------------------
template <class T, int VEC>
__device__ __forceinline__ void load_unit(const char *p, T (&x)[VEC],
                                          bool full, int cnt) {
  constexpr int B = VEC * (int)sizeof(T);
  if (B <= 16 && (B & (B - 1)) == 0 && full) {
    typename Raw<(B <= 16 ? B : 16)>::type r = ld<(B <= 16 ? B : 16)>(p);
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

// Team index of leaf k's member: ceil(k*P / 2^lg).
__device__ __forceinline__ int member_of(int k, int P, int lg) {
  return (int)(((int64_t)k * P + ((1 << lg) - 1)) >> lg);
}

// ---- exact per-member MIN/MAX for special elements ----------------------
// Cooperative over the workgroup: each special element is handled in turn.
// Its P inputs are copied into LDS before any result is written (so
// dest == source works), then thread t evaluates members t, t+256, ...
// over the LDS copy and writes their results (and the folded neighbour's).
template <class T, int OP>
__device__ void special_block(const oscd_geom &g, uint64_t src_off,
                              uint64_t dest_off, uint64_t my_e0,
                              uint32_t mine, int lg, T *snap) {
  __shared__ int s_owner;
  __shared__ unsigned long long s_e;
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
      const int v = __ffs(mine) - 1;
      mine &= mine - 1;
      s_e = my_e0 + (uint64_t)v;
    }
    __syncthreads();
    const uint64_t e = s_e;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
      snap[i] = *(const T *)(member_ptr(g, i, src_off) + e * sizeof(T));
    __syncthreads();
    for (int k = threadIdx.x; k < p2s; k += blockDim.x) {
      T stk[16];
      for (int j = 0; j < p2s; ++j) {
        const int kk = k ^ j;
        const int m = member_of(kk, P, lg);
        const int m1 = member_of(kk + 1, P, lg);
        T leaf = (m1 - m == 2) ? Bin<OP>::f(snap[m + 1], snap[m]) : snap[m];
        int L = 0;
        while ((j >> L) & 1) {
          leaf = Bin<OP>::f(stk[L], leaf);
          ++L;
        }
        stk[L] = leaf;
      }
      const T r = stk[lg];
      const int mk = member_of(k, P, lg);
      const int mk1 = member_of(k + 1, P, lg);
      *(T *)(member_ptr(g, mk, dest_off) + e * sizeof(T)) = r;
      if (mk1 - mk == 2)
        *(T *)(member_ptr(g, mk + 1, dest_off) + e * sizeof(T)) = r;
    }
    __syncthreads();
  }
}

// ---- the kernel ------------------------------------------------------------
constexpr int kBlock = 256;
constexpr int kInflight = 8;  // source slices loaded before combining

template <class T, class Pol, int VEC, int MAXLV>
__global__ __launch_bounds__(kBlock) void k_reduce(const oscd_geom g,
                                                   const uint64_t dest_off,
                                                   const uint64_t src_off,
                                                   const uint64_t n,
                                                   const int lg) {
  typedef typename Pol::Acc Acc;
  extern __shared__ unsigned char k_reduce_lds[];
  const int P = g.npes;
  const int p2s = 1 << lg;
  const int64_t step = (int64_t)g.stride * (int64_t)g.slice;
  const char *s0 = member_ptr(g, 0, src_off);
  char *d0 = member_ptr(g, 0, dest_off);
  const uint64_t nunits = (n + VEC - 1) / VEC;

  for (uint64_t ub = (uint64_t)blockIdx.x * kBlock; ub < nunits;
       ub += (uint64_t)gridDim.x * kBlock) {
    const uint64_t u = ub + threadIdx.x;
    const bool active = u < nunits;
    const uint64_t e0 = u * VEC;
    const uint64_t boff = e0 * sizeof(T);
    const bool full = active && (e0 + VEC <= n);
    const int cnt = !active ? 0 : (full ? VEC : (int)(n - e0));
    T r[VEC];
    uint32_t spec = 0;

    if (active) {
      Acc acc[VEC];
      if constexpr (!Pol::TREE) {
        uint32_t nanm = 0;
        for (int i0 = 0; i0 < P; i0 += kInflight) {
          T buf[kInflight][VEC];
          const int m = (P - i0) < kInflight ? (P - i0) : kInflight;
#pragma unroll
          for (int j = 0; j < kInflight; ++j)
            if (j < m) load_unit<T, VEC>(s0 + (i0 + j) * step + boff, buf[j],
                                         full, cnt);
#pragma unroll
          for (int j = 0; j < kInflight; ++j) {
            if (j < m) {
#pragma unroll
              for (int v = 0; v < VEC; ++v) {
                if (Pol::SPECIAL && is_nan(buf[j][v])) nanm |= 1u << v;
                acc[v] = (i0 + j == 0) ? Pol::lift(buf[j][v])
                                       : Pol::seq(acc[v], buf[j][v]);
              }
            }
          }
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          r[v] = Pol::out(acc[v]);
          if (Pol::SPECIAL && v < cnt &&
              (((nanm >> v) & 1u) || is_nan_or_zero(r[v])))
            spec |= 1u << v;
        }
      } else {
        // Reference tree order: leaves in team-index order, folded pairs
        // op(x[m+1], x[m]), merged by a binary counter: stk[L] holds the
        // pending left subtree of 2^L leaves.
        Acc stk[MAXLV][VEC];
        T pend[VEC];
        int k = 0;
        int mk = 0;
        int mn = member_of(1, P, lg);
        for (int i0 = 0; i0 < P; i0 += kInflight) {
          T buf[kInflight][VEC];
          const int m = (P - i0) < kInflight ? (P - i0) : kInflight;
#pragma unroll
          for (int j = 0; j < kInflight; ++j)
            if (j < m) load_unit<T, VEC>(s0 + (i0 + j) * step + boff, buf[j],
                                         full, cnt);
#pragma unroll
          for (int j = 0; j < kInflight; ++j) {
            if (j < m) {
              const int i = i0 + j;
              Acc leaf[VEC];
              bool ready = true;
              if (i == mk) {
                if (mn - mk == 2) {
#pragma unroll
                  for (int v = 0; v < VEC; ++v) pend[v] = buf[j][v];
                  ready = false;
                } else {
#pragma unroll
                  for (int v = 0; v < VEC; ++v) leaf[v] = Pol::lift(buf[j][v]);
                }
              } else {
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                  leaf[v] = Pol::comb(Pol::lift(buf[j][v]), Pol::lift(pend[v]));
              }
              if (ready) {
#pragma unroll
                for (int L = 0; L < MAXLV; ++L) {
                  if ((k >> L) & 1) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                      leaf[v] = Pol::comb(stk[L][v], leaf[v]);
                  } else {
#pragma unroll
                    for (int v = 0; v < VEC; ++v) stk[L][v] = leaf[v];
                    break;
                  }
                }
                ++k;
                mk = mn;
                mn = member_of(k + 1, P, lg);
              }
            }
          }
        }
#pragma unroll
        for (int L = 0; L < MAXLV; ++L)
          if (L == lg) {
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] = stk[L][v];
          }
#pragma unroll
        for (int v = 0; v < VEC; ++v) r[v] = Pol::out(acc[v]);
        (void)p2s;
      }
    }

    if constexpr (Pol::SPECIAL) {
      if (__syncthreads_or(spec != 0))
        special_block<T, Pol::OPC>(g, src_off, dest_off, e0, spec, lg,
                                   (T *)k_reduce_lds);
    }

    if (active) {
      for (int j = 0; j < P; ++j)
        store_unit<T, VEC>(d0 + j * step + boff, r, full, cnt, spec);
    }
  }
}

// ---- host-side dispatch ----------------------------------------------------
template <class T>
constexpr int vec_of() {
  return (16 / (int)sizeof(T)) > 8 ? 8 : (16 / (int)sizeof(T));
}

template <class T, class Pol>
static int launch(const oscd_geom *g, int lg, uint64_t doff, uint64_t soff,
                  uint64_t n, hipStream_t s) {
  constexpr int VB = vec_of<T>();
  constexpr uint64_t AL = VB * sizeof(T);
  const bool vec_ok = (doff % AL == 0) && (soff % AL == 0) &&
                      (g->slice % AL == 0) && ((uintptr_t)g->base % AL == 0);
  const size_t lds = Pol::SPECIAL ? (size_t)g->npes * sizeof(T) : 0;
  const uint64_t units = vec_ok ? (n + VB - 1) / VB : n;
  uint64_t grid = (units + kBlock - 1) / kBlock;
  if (grid > (1u << 20)) grid = 1u << 20;
  if (grid == 0) return OSC_OK;
  if (vec_ok) {
    if (Pol::TREE && lg >= 8)
      hipLaunchKernelGGL((k_reduce<T, Pol, VB, 13>), dim3((unsigned)grid),
                         dim3(kBlock), lds, s, *g, doff, soff, n, lg);
    else
      hipLaunchKernelGGL((k_reduce<T, Pol, VB, Pol::TREE ? 8 : 1>),
                         dim3((unsigned)grid), dim3(kBlock), lds, s, *g, doff,
                         soff, n, lg);
  } else {
    if (Pol::TREE && lg >= 8)
      hipLaunchKernelGGL((k_reduce<T, Pol, 1, 13>), dim3((unsigned)grid),
                         dim3(kBlock), lds, s, *g, doff, soff, n, lg);
    else
      hipLaunchKernelGGL((k_reduce<T, Pol, 1, Pol::TREE ? 8 : 1>),
                         dim3((unsigned)grid), dim3(kBlock), lds, s, *g, doff,
                         soff, n, lg);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

// Integer dtypes: bitwise/sum/prod use the unsigned type of the same width
// (two's-complement wrap, as the reference's C arithmetic on x86-64);
// MIN/MAX compare with the dtype's own signedness.
template <class S, class U>
static int launch_int(const oscd_geom *g, int op, int lg, uint64_t doff,
                      uint64_t soff, uint64_t n, hipStream_t s) {
  switch (op) {
    case OSC_OP_AND: return launch<U, PolSeq<U, OSC_OP_AND, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_OR: return launch<U, PolSeq<U, OSC_OP_OR, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_XOR: return launch<U, PolSeq<U, OSC_OP_XOR, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_SUM: return launch<U, PolSeq<U, OSC_OP_SUM, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_PROD: return launch<U, PolSeq<U, OSC_OP_PROD, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_MIN: return launch<S, PolSeq<S, OSC_OP_MIN, false>>(g, lg, doff, soff, n, s);
    case OSC_OP_MAX: return launch<S, PolSeq<S, OSC_OP_MAX, false>>(g, lg, doff, soff, n, s);
  }
  return OSC_ERR_ARG;
}

template <class T, template <int> class AccPol>
static int launch_real(const oscd_geom *g, int op, int mode, int lg,
                       uint64_t doff, uint64_t soff, uint64_t n,
                       hipStream_t s) {
  switch (op) {
    case OSC_OP_MIN: return launch<T, PolSeq<T, OSC_OP_MIN, true>>(g, lg, doff, soff, n, s);
    case OSC_OP_MAX: return launch<T, PolSeq<T, OSC_OP_MAX, true>>(g, lg, doff, soff, n, s);
    case OSC_OP_SUM:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_SUM>>(g, lg, doff, soff, n, s)
                 : launch<T, AccPol<OSC_OP_SUM>>(g, lg, doff, soff, n, s);
    case OSC_OP_PROD:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_PROD>>(g, lg, doff, soff, n, s)
                 : launch<T, AccPol<OSC_OP_PROD>>(g, lg, doff, soff, n, s);
  }
  return OSC_ERR_UNSUPPORTED;
}

template <class T, template <int> class AccPol>
static int launch_cplx(const oscd_geom *g, int op, int mode, int lg,
                       uint64_t doff, uint64_t soff, uint64_t n,
                       hipStream_t s) {
  switch (op) {
    case OSC_OP_SUM:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_SUM>>(g, lg, doff, soff, n, s)
                 : launch<T, AccPol<OSC_OP_SUM>>(g, lg, doff, soff, n, s);
    case OSC_OP_PROD:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_PROD>>(g, lg, doff, soff, n, s)
                 : launch<T, AccPol<OSC_OP_PROD>>(g, lg, doff, soff, n, s);
  }
  return OSC_ERR_UNSUPPORTED;
}

}  // namespace osc

extern "C" int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode,
                           uint64_t dest_off, uint64_t src_off, uint64_t n,
                           void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 8191 || op < 0 ||
      op >= OSC_OP_COUNT)
    return OSC_ERR_ARG;
  if (n == 0) return OSC_OK;
  int lg = 0;
  while ((2 << lg) <= g->npes) ++lg;
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case OSC_T_U8: return launch_int<uint8_t, uint8_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_U16: return launch_int<uint16_t, uint16_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_U32: return launch_int<uint32_t, uint32_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_U64: return launch_int<uint64_t, uint64_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_I8: return launch_int<int8_t, uint8_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_I16: return launch_int<int16_t, uint16_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_I32: return launch_int<int32_t, uint32_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_I64: return launch_int<int64_t, uint64_t>(g, op, lg, dest_off, src_off, n, s);
    case OSC_T_F32: return launch_real<float, PolF32Acc>(g, op, fpmode, lg, dest_off, src_off, n, s);
    case OSC_T_F64: return launch_real<double, PolF64Acc>(g, op, fpmode, lg, dest_off, src_off, n, s);
    case OSC_T_C32: return launch_cplx<cf32, PolC32Acc>(g, op, fpmode, lg, dest_off, src_off, n, s);
    case OSC_T_C64: return launch_cplx<cf64, PolC64Acc>(g, op, fpmode, lg, dest_off, src_off, n, s);
  }
  return OSC_ERR_ARG;
}
