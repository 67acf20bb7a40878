// dev_fill.hip - seeded synthetic inputs written straight into the heap.
//
// value(seed, pe, idx) is a pure function of the seed, the GLOBAL PE number
// and the element index, built from the SplitMix64 finaliser (Steele, Lea
// and Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014):
//   mix(z) = z ^= z>>30, z *= 0xBF58476D1CE4E5B9, z ^= z>>27,
//            z *= 0x94D049BB133111EB, z ^= z>>31
//   h(seed, pe, idx) = mix(mix(mix(seed) + G*(pe+1)) + G*(idx+1)),
//   G = 0x9E3779B97F4A7C15.
// The same function is implemented in the host C library and in the
// Python test tooling; integer distributions are bit-identical across all
// three.  Real-valued ones go through FP64 log/sqrt/cos, whose last bits may
// differ between device and host libm, so tests always read the generated
// inputs back from the device rather than regenerating them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

namespace osc {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t pe,
                                          uint64_t idx) {
  const uint64_t G = 0x9E3779B97F4A7C15ull;
  return mix64(mix64(mix64(seed) + G * (pe + 1)) + G * (idx + 1));
}

// Value of one real/integer draw as a double or as raw 64 bits.
struct Draw {
  uint64_t bits;
  double real;
  bool is_real;
};

__device__ Draw draw(int dist, double a, double b, uint64_t h, int i,
                     uint64_t e) {
  Draw d{0, 0.0, false};
  switch (dist) {
    case OSC_DIST_BITS:
      d.bits = h;
      break;
    case OSC_DIST_INT_RANGE: {
      const int64_t lo = (int64_t)a, hi = (int64_t)b;
      const uint64_t range = (uint64_t)hi - (uint64_t)lo + 1ull;
      d.bits = (uint64_t)lo + (range == 0 ? h : __umul64hi(h, range));
      break;
    }
    case OSC_DIST_SET_M1_1_2: {
      const int64_t set[3] = {-1, 1, 2};
      d.bits = (uint64_t)set[__umul64hi(h, 3ull)];
      break;
    }
    case OSC_DIST_NORMAL: {
      const double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;
      const double u2 = (double)(mix64(h ^ 0xD1B54A32D192ED03ull) >> 11) *
                        0x1.0p-53;
      const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      d.real = a + b * z;
      d.is_real = true;
      break;
    }
    case OSC_DIST_REAL_RANGE:
      d.real = a + (b - a) * ((double)(h >> 11) * 0x1.0p-53);
      d.is_real = true;
      break;
    case OSC_DIST_A2A_PATTERN: {
      const uint64_t blk = (uint64_t)a;
      const uint64_t j = e / blk, idx = e - j * blk;
      d.bits = ((uint64_t)i << 40) | (j << 20) | idx;
      break;
    }
    case OSC_DIST_FCOL_PATTERN:
      d.bits = (uint64_t)(uint32_t)(((uint32_t)i << 24) | (uint32_t)e);
      break;
  }
  return d;
}

template <class T>
__device__ __forceinline__ T as_int(const Draw &d) {
  return d.is_real ? (T)(int64_t)d.real : (T)d.bits;
}
template <class T>
__device__ __forceinline__ T as_real(const Draw &d, int dist) {
  if (d.is_real) return (T)d.real;
  if (dist == OSC_DIST_BITS) {
    T r;
    __builtin_memcpy(&r, &d.bits, sizeof(T));
    return r;
  }
  return (T)(int64_t)d.bits;
}

__global__ __launch_bounds__(256) void k_fill(const oscd_geom g,
                                              const uint64_t off,
                                              const uint64_t n,
                                              const int dtype, const int dist,
                                              const double a, const double b,
                                              const uint64_t seed) {
  const int i = blockIdx.y;
  const uint64_t pe = (uint64_t)(int64_t)(g.start + i * g.stride);
  char *base = member_ptr(g, i, off);
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * 256) {
    const bool cplx = dtype == OSC_T_C32 || dtype == OSC_T_C64;
    const uint64_t k0 = cplx ? 2 * e : e;
    const Draw d = draw(dist, a, b, hash3(seed, pe, k0), i, e);
    switch (dtype) {
      case OSC_T_U8: case OSC_T_I8: ((uint8_t *)base)[e] = as_int<uint8_t>(d); break;
      case OSC_T_U16: case OSC_T_I16: ((uint16_t *)base)[e] = as_int<uint16_t>(d); break;
      case OSC_T_U32: case OSC_T_I32: ((uint32_t *)base)[e] = as_int<uint32_t>(d); break;
      case OSC_T_U64: case OSC_T_I64: ((uint64_t *)base)[e] = as_int<uint64_t>(d); break;
      case OSC_T_F32: ((float *)base)[e] = as_real<float>(d, dist); break;
      case OSC_T_F64: ((double *)base)[e] = as_real<double>(d, dist); break;
      case OSC_T_C32: {
        const Draw d2 = draw(dist, a, b, hash3(seed, pe, k0 + 1), i, e);
        ((float *)base)[2 * e] = as_real<float>(d, dist);
        ((float *)base)[2 * e + 1] = as_real<float>(d2, dist);
        break;
      }
      case OSC_T_C64: {
        const Draw d2 = draw(dist, a, b, hash3(seed, pe, k0 + 1), i, e);
        ((double *)base)[2 * e] = as_real<double>(d, dist);
        ((double *)base)[2 * e + 1] = as_real<double>(d2, dist);
        break;
      }
    }
  }
}

}  // namespace osc

extern "C" int oscd_fill(const oscd_geom *g, uint64_t off, uint64_t nelems,
                         int dtype, int dist, double a, double b,
                         uint64_t seed, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535 || dtype < 0 ||
      dtype >= OSC_T_COUNT || dist < 0 || dist >= OSC_DIST_COUNT)
    return OSC_ERR_ARG;
  if (dist == OSC_DIST_A2A_PATTERN && a < 1.0) return OSC_ERR_ARG;
  if (nelems == 0) return OSC_OK;
  uint64_t gx = (nelems + 255) / 256;
  if (gx > 16384) gx = 16384;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)gx, (unsigned)g->npes), dim3(256),
                     0, (hipStream_t)stream, *g, off, nelems, dtype, dist, a,
                     b, seed);
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}
