// dev_batch.hip - batched execution: the persistent tile loop for data
// movement, per-call result digests and hipGraph capture.
//
// A batch is a list of independent calls.  The host planner cuts every
// call into tiles; one launch per kernel family and wave runs a grid of
// workgroups that stride over the tile list, so thousands of calls of
// uneven size share one launch and fill the chip together.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

namespace osc {

constexpr int kBB = 256;  // threads per workgroup

// Data movement of unit width W for fcollect / broadcast / alltoall tiles
// (tile semantics in osc_dev.h).  Each loaded unit stays in registers and
// is stored to every destination it belongs to.
template <int W>
__global__ __launch_bounds__(kBB) void k_batch_move(
    const oscd_bcall *__restrict__ calls, const oscd_tile *__restrict__ tiles,
    const uint32_t ntiles, char *const base, const uint64_t slice) {
  typedef typename Raw<W>::type V;
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
    const int P = c.npes;
    const int64_t step = (int64_t)c.stride * (int64_t)slice;
    const uint64_t nunits = c.count / W;
    if (c.kind == OSC_CALL_FCOLLECT || c.kind == OSC_CALL_BROADCAST) {
      const int i = c.kind == OSC_CALL_FCOLLECT ? (int)tl.aux : c.aux;
      const char *src = member_ptr(g, i, c.src_off);
      char *d0 = member_ptr(g, 0, c.dest_off +
                                      (c.kind == OSC_CALL_FCOLLECT
                                           ? (uint64_t)i * c.count
                                           : 0));
      for (uint64_t u = tl.begin + threadIdx.x; u < tl.end; u += kBB) {
        const V v = ld_nt<W>(src + u * W);
        char *d = d0 + u * W;
        for (int j = 0; j < P; ++j) st_nt<W>(d + j * step, v);
      }
    } else if (nunits >= (uint64_t)kBB) {
      // alltoall, one pair per tile: src_i[j*m + u] -> dest_j[i*m + u]
      const int i = (int)(tl.aux / (uint32_t)P);
      const int j = (int)(tl.aux - (uint32_t)i * (uint32_t)P);
      const char *src = member_ptr(g, i, c.src_off) + (uint64_t)j * c.count;
      char *dst = member_ptr(g, j, c.dest_off) + (uint64_t)i * c.count;
      uint64_t u = tl.begin + threadIdx.x;
      for (; u + 3 * kBB < tl.end; u += 4 * kBB) {
        V v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = ld_nt<W>(src + (u + k * kBB) * W);
#pragma unroll
        for (int k = 0; k < 4; ++k) st_nt<W>(dst + (u + k * kBB) * W, v[k]);
      }
      for (; u < tl.end; u += kBB) st_nt<W>(dst + u * W, ld_nt<W>(src + u * W));
    } else {
      // alltoall, small blocks: pairs [begin, end), flattened with units
      const uint32_t nu = (uint32_t)nunits;
      const uint32_t flat = (uint32_t)(tl.end - tl.begin) * nu;
      for (uint32_t f = threadIdx.x; f < flat; f += kBB) {
        const uint32_t pl = f / nu;
        const uint32_t u = f - pl * nu;
        const uint32_t pair = (uint32_t)tl.begin + pl;
        const int i = (int)(pair / (uint32_t)P);
        const int j = (int)(pair - (uint32_t)i * (uint32_t)P);
        const V v = ld_nt<W>(member_ptr(g, i, c.src_off) +
                             (uint64_t)j * c.count + (uint64_t)u * W);
        st_nt<W>(member_ptr(g, j, c.dest_off) + (uint64_t)i * c.count +
                     (uint64_t)u * W,
                 v);
      }
    }
  }
}

__device__ __forceinline__ uint64_t dmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t elem_size(int dtype) {
  switch (dtype) {
    case OSC_T_U8: case OSC_T_I8: return 1;
    case OSC_T_U16: case OSC_T_I16: return 2;
    case OSC_T_U32: case OSC_T_I32: case OSC_T_F32: return 4;
    case OSC_T_C64: return 16;
  }
  return 8;
}

__device__ __forceinline__ uint64_t dest_bytes_of(const oscd_bcall &c) {
  switch (c.kind) {
    case OSC_CALL_REDUCE: return c.count * elem_size(c.dtype);
    case OSC_CALL_FCOLLECT:
    case OSC_CALL_ALLTOALL: return c.count * (uint64_t)c.npes;
  }
  return c.count;
}

// digest[call] ^= XOR over members j and 8-byte words w of
// mix(word ^ mix(j << 40 ^ w)); tail bytes are zero-padded into a word.
__global__ __launch_bounds__(kBB) void k_digest(const oscd_bcall *calls,
                                                uint64_t *digests,
                                                char *const base,
                                                const uint64_t slice) {
  const oscd_bcall c = calls[blockIdx.y];
  const uint64_t bytes = dest_bytes_of(c);
  const uint64_t nw = (bytes + 7) / 8;
  const uint64_t total = nw * (uint64_t)c.npes;
  uint64_t h = 0;
  for (uint64_t f = (uint64_t)blockIdx.x * kBB + threadIdx.x; f < total;
       f += (uint64_t)gridDim.x * kBB) {
    const uint64_t j = f / nw, w = f - j * nw;
    const char *p = base + (uint64_t)(int64_t)(c.start + (int)j * c.stride) * slice +
                    c.dest_off + w * 8;
    uint64_t word = 0;
    if (w * 8 + 8 <= bytes && (((uintptr_t)p) & 7) == 0) {
      word = *(const uint64_t *)p;
    } else {
      for (uint64_t b = 0; b < 8 && w * 8 + b < bytes; ++b)
        word |= (uint64_t)(uint8_t)p[b] << (8 * b);
    }
    h ^= dmix(word ^ dmix((j << 40) ^ w ^ 0x5bd1e995ull));
  }
  __shared__ uint64_t red[kBB];
  red[threadIdx.x] = h;
  __syncthreads();
  for (int s = kBB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] ^= red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicXor((unsigned long long *)&digests[blockIdx.y],
              (unsigned long long)red[0]);
}

__global__ __launch_bounds__(kBB) void k_fill_u64(uint64_t *p, uint64_t v,
                                                  uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBB + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBB)
    p[i] = v;
}

}  // namespace osc

extern "C" int oscd_batch_move(int unit_width, const oscd_bcall *calls,
                               const oscd_tile *tiles, uint32_t ntiles,
                               uint32_t grid, char *base, uint64_t slice,
                               void *stream) {
  using namespace osc;
  if (ntiles == 0) return OSC_OK;
  if (grid == 0) grid = 1;
  hipStream_t s = (hipStream_t)stream;
  switch (unit_width) {
    case 16: hipLaunchKernelGGL(k_batch_move<16>, dim3(grid), dim3(kBB), 0, s, calls, tiles, ntiles, base, slice); break;
    case 8: hipLaunchKernelGGL(k_batch_move<8>, dim3(grid), dim3(kBB), 0, s, calls, tiles, ntiles, base, slice); break;
    case 4: hipLaunchKernelGGL(k_batch_move<4>, dim3(grid), dim3(kBB), 0, s, calls, tiles, ntiles, base, slice); break;
    case 2: hipLaunchKernelGGL(k_batch_move<2>, dim3(grid), dim3(kBB), 0, s, calls, tiles, ntiles, base, slice); break;
    case 1: hipLaunchKernelGGL(k_batch_move<1>, dim3(grid), dim3(kBB), 0, s, calls, tiles, ntiles, base, slice); break;
    default: return OSC_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

extern "C" int oscd_batch_digest(const oscd_bcall *calls, uint32_t ncalls,
                                 uint64_t *digests, char *base,
                                 uint64_t slice, void *stream) {
  using namespace osc;
  if (ncalls == 0) return OSC_OK;
  if (ncalls > 65535) return OSC_ERR_ARG;
  hipLaunchKernelGGL(k_digest, dim3(64, ncalls), dim3(kBB), 0,
                     (hipStream_t)stream, calls, digests, base, slice);
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

extern "C" int oscd_fill_u64(uint64_t *p, uint64_t value, uint64_t count,
                             void *stream) {
  using namespace osc;
  if (count == 0) return OSC_OK;
  uint64_t grid = (count + kBB - 1) / kBB;
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)grid), dim3(kBB), 0,
                     (hipStream_t)stream, p, value, count);
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

extern "C" int oscd_cu_count(int *n) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return OSC_ERR_HIP;
  if (hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev) !=
      hipSuccess)
    return OSC_ERR_HIP;
  return OSC_OK;
}

extern "C" int oscd_graph_begin(void *stream) {
  return hipStreamBeginCapture((hipStream_t)stream,
                               hipStreamCaptureModeThreadLocal) == hipSuccess
             ? OSC_OK
             : OSC_ERR_HIP;
}

extern "C" int oscd_graph_end(void *stream, void **exec) {
  hipGraph_t graph = nullptr;
  hipGraphExec_t ge = nullptr;
  if (hipStreamEndCapture((hipStream_t)stream, &graph) != hipSuccess)
    return OSC_ERR_HIP;
  const hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return OSC_ERR_HIP;
  *exec = (void *)ge;
  return OSC_OK;
}

extern "C" int oscd_graph_launch(void *exec, void *stream) {
  return hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream) ==
                 hipSuccess
             ? OSC_OK
             : OSC_ERR_HIP;
}

extern "C" int oscd_graph_destroy(void *exec) {
  return hipGraphExecDestroy((hipGraphExec_t)exec) == hipSuccess ? OSC_OK
                                                                  : OSC_ERR_HIP;
}
