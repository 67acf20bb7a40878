// dev_move.hip - data-movement collectives: fcollect, alltoall, broadcast,
// and the plain device copy used by put/get.
//
// These are pure permutations of bytes, so the kernels compute the closed
// form of the reference algorithms directly:
//  * fcollect, default bruck_inplace (src/shcoll/src/fcollect.c:370-403 with
//    the final rotate_inplace of src/shcoll/src/util/rotate.c:22-54):
//    dest_j[i*nbytes .. (i+1)*nbytes) = src_i for every member pair (i, j);
//  * alltoall, default shift_exchange_barrier (src/shcoll/src/alltoall.c:
//    75-110): dest_j[i*nbytes ..) = src_i[j*nbytes ..);
//  * broadcast (src/shcoll/src/broadcast.c:511-553): dest_j = src_root for
//    every member j, the root included.
// Each source byte is loaded once into registers and stored to every
// destination it belongs to; unit width W (16 bytes when alignment allows)
// is picked per call from the alignment of sizes, offsets and slice.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

namespace osc {

constexpr int kMoveBlock = 256;

// Fan-out: one source member's [soff, soff+nbytes) to every member's
// [doff_base + blockIdx.y*doff_step, ...).  blockIdx.y selects the source
// member when src_member < 0 (fcollect), otherwise src_member is used
// (broadcast, gridDim.y == 1).
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_fanout(
    const oscd_geom g, const int src_member, const uint64_t soff,
    const uint64_t doff_base, const uint64_t doff_step, const uint64_t nbytes) {
  const int i = src_member >= 0 ? src_member : (int)blockIdx.y;
  const uint64_t nunits = nbytes / W;
  const char *src = member_ptr(g, i, soff);
  char *d0 = member_ptr(g, 0, doff_base + (uint64_t)blockIdx.y * doff_step);
  const int64_t step = (int64_t)g.stride * (int64_t)g.slice;
  const int P = g.npes;
  for (uint64_t u = (uint64_t)blockIdx.x * kMoveBlock + threadIdx.x;
       u < nunits; u += (uint64_t)gridDim.x * kMoveBlock) {
    const typename Raw<W>::type v = ld<W>(src + u * W);
    char *d = d0 + u * W;
    for (int j = 0; j < P; ++j) st_nt<W>(d + j * step, v);
  }
}

// Alltoall, large blocks: block b covers chunk (b % cpb) of the pair
// (b / cpb); pair p = i*P + j moves src_i[j*nbytes ..] -> dest_j[i*nbytes ..].
// Each thread moves UPT units, all loads issued before the stores.
template <int W, int UPT>
__global__ __launch_bounds__(kMoveBlock) void k_a2a_big(
    const oscd_geom g, const uint64_t doff, const uint64_t soff,
    const uint64_t nbytes, const uint32_t cpb) {
  const uint64_t nunits = nbytes / W;
  const uint32_t b = blockIdx.x;
  const uint32_t pair = b / cpb;
  const uint32_t chunk = b - pair * cpb;
  const int i = (int)(pair / (uint32_t)g.npes);
  const int j = (int)(pair - (uint32_t)i * (uint32_t)g.npes);
  const char *src = member_ptr(g, i, soff) + (uint64_t)j * nbytes;
  char *dst = member_ptr(g, j, doff) + (uint64_t)i * nbytes;
  const uint64_t u0 = (uint64_t)chunk * (kMoveBlock * UPT) + threadIdx.x;
  typename Raw<W>::type v[UPT];
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    const uint64_t u = u0 + (uint64_t)k * kMoveBlock;
    if (u < nunits) v[k] = ld<W>(src + u * W);
  }
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    const uint64_t u = u0 + (uint64_t)k * kMoveBlock;
    if (u < nunits) st_nt<W>(dst + u * W, v[k]);
  }
}

// Alltoall, small blocks (nbytes <= 64): a 16x16 tile of (source i,
// destination j) pairs per workgroup, staged through LDS so the loads run
// along each source row (contiguous j) and the stores along each
// destination row (contiguous i) instead of one tiny store per slice.
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_a2a_small(
    const oscd_geom g, const uint64_t doff, const uint64_t soff,
    const uint64_t nbytes) {
  constexpr int T = 16;
  constexpr int NU = 64 / W;
  __shared__ typename Raw<W>::type tile[T][T][NU + 1];
  const int P = g.npes;
  const int nunits = (int)(nbytes / W);
  const int i0 = blockIdx.y * T;
  const int j0 = blockIdx.x * T;
  const int row_units = T * nunits;
  for (int di = 0; di < T; ++di) {
    const int i = i0 + di;
    if (i >= P) break;
    const char *src = member_ptr(g, i, soff) + (uint64_t)j0 * nbytes;
    for (int f = threadIdx.x; f < row_units; f += kMoveBlock) {
      const int dj = f / nunits;
      const int u = f - dj * nunits;
      if (j0 + dj < P) tile[di][dj][u] = ld<W>(src + (uint64_t)f * W);
    }
  }
  __syncthreads();
  for (int dj = 0; dj < T; ++dj) {
    const int j = j0 + dj;
    if (j >= P) break;
    char *dst = member_ptr(g, j, doff) + (uint64_t)i0 * nbytes;
    for (int f = threadIdx.x; f < row_units; f += kMoveBlock) {
      const int di = f / nunits;
      const int u = f - di * nunits;
      if (i0 + di < P) st_nt<W>(dst + (uint64_t)f * W, tile[di][dj][u]);
    }
  }
}

// Alltoall, medium blocks (fewer than 1024 units per pair): workgroup b
// moves the ppb consecutive pairs [b*ppb, (b+1)*ppb), flattened over
// (pair, unit) so every lane has work; 4 units per thread.
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_a2a_mid(
    const oscd_geom g, const uint64_t doff, const uint64_t soff,
    const uint64_t nbytes, const uint32_t ppb) {
  const uint32_t nunits = (uint32_t)(nbytes / W);
  const uint32_t P = (uint32_t)g.npes;
  const uint64_t npairs = (uint64_t)P * P;
  const uint64_t pair0 = (uint64_t)blockIdx.x * ppb;
  const uint32_t flat = ppb * nunits;
  typename Raw<W>::type v[4];
  char *dp[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t f = threadIdx.x + k * kMoveBlock;
    dp[k] = nullptr;
    if (f < flat) {
      const uint32_t pl = f / nunits;
      const uint32_t u = f - pl * nunits;
      const uint64_t pair = pair0 + pl;
      if (pair < npairs) {
        const int i = (int)(pair / P);
        const int j = (int)(pair - (uint64_t)i * P);
        v[k] = ld<W>(member_ptr(g, i, soff) + (uint64_t)j * nbytes +
                     (uint64_t)u * W);
        dp[k] = member_ptr(g, j, doff) + (uint64_t)i * nbytes + (uint64_t)u * W;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (dp[k] != nullptr) st_nt<W>(dp[k], v[k]);
}

// Plain copy, 16-byte units with a byte tail (alignment checked by host).
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_copy(char *dst,
                                                     const char *src,
                                                     const uint64_t nbytes) {
  const uint64_t nunits = nbytes / W;
  const uint64_t gs = (uint64_t)gridDim.x * kMoveBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kMoveBlock + threadIdx.x;
       u < nunits; u += gs)
    st_nt<W>(dst + u * W, ld<W>(src + u * W));
}

static int unit_width(uint64_t a) {
  // largest power of two <= 16 dividing a (a == 0 -> 16)
  int w = 16;
  while (w > 1 && (a % (uint64_t)w) != 0) w >>= 1;
  return w;
}

static int alignment_of(const oscd_geom *g, uint64_t doff, uint64_t soff,
                        uint64_t nbytes) {
  uint64_t m = doff | soff | nbytes | g->slice | (uint64_t)(uintptr_t)g->base;
  return unit_width(m == 0 ? 16 : m);
}

static unsigned grid_for(uint64_t nunits, uint64_t cap) {
  uint64_t b = (nunits + kMoveBlock - 1) / kMoveBlock;
  if (b > cap) b = cap;
  if (b == 0) b = 1;
  return (unsigned)b;
}

static int check_launch() {
  return hipGetLastError() == hipSuccess ? OSC_OK : OSC_ERR_HIP;
}

#define OSC_W_DISPATCH(W, CALL)                                                \
  switch (W) {                                                                 \
    case 16: { constexpr int WW = 16; CALL; } break;                           \
    case 8: { constexpr int WW = 8; CALL; } break;                             \
    case 4: { constexpr int WW = 4; CALL; } break;                             \
    case 2: { constexpr int WW = 2; CALL; } break;                             \
    default: { constexpr int WW = 1; CALL; } break;                            \
  }

static int fanout(const oscd_geom *g, int src_member, uint64_t soff,
                  uint64_t doff_base, uint64_t doff_step, uint64_t nbytes,
                  unsigned ny, hipStream_t s) {
  const int W = alignment_of(g, doff_base | doff_step, soff, nbytes);
  const uint64_t nunits = nbytes / W;
  // enough workgroups in x to fill the chip when ny is small
  const unsigned gx = grid_for(nunits, ny >= 64 ? 4096 : 65535);
  OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_fanout<WW>), dim3(gx, ny),
                                       dim3(kMoveBlock), 0, s, *g, src_member,
                                       soff, doff_base, doff_step, nbytes));
  return check_launch();
}

}  // namespace osc

extern "C" int oscd_fcollect(const oscd_geom *g, uint64_t dest_off,
                             uint64_t src_off, uint64_t nbytes, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535) return OSC_ERR_ARG;
  if (nbytes == 0) return OSC_OK;
  return fanout(g, -1, src_off, dest_off, nbytes, nbytes, (unsigned)g->npes,
                (hipStream_t)stream);
}

extern "C" int oscd_broadcast(const oscd_geom *g, int root, uint64_t dest_off,
                              uint64_t src_off, uint64_t nbytes, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || root < 0 || root >= g->npes)
    return OSC_ERR_ARG;
  if (nbytes == 0) return OSC_OK;
  return fanout(g, root, src_off, dest_off, 0, nbytes, 1u,
                (hipStream_t)stream);
}

extern "C" int oscd_alltoall(const oscd_geom *g, uint64_t dest_off,
                             uint64_t src_off, uint64_t nbytes, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535) return OSC_ERR_ARG;
  if (nbytes == 0) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int W = alignment_of(g, dest_off, src_off, nbytes);
  const uint64_t nunits = nbytes / W;
  const uint64_t P = (uint64_t)g->npes;
  if (nbytes <= 64) {
    const unsigned t = (unsigned)((P + 15) / 16);
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_a2a_small<WW>), dim3(t, t),
                                         dim3(kMoveBlock), 0, s, *g, dest_off,
                                         src_off, nbytes));
    return check_launch();
  }
  constexpr int UPT = 4;
  constexpr uint64_t kChunk = kMoveBlock * UPT;
  if (nunits < kChunk) {
    const uint64_t ppb = kChunk / nunits;
    const uint64_t blocks = (P * P + ppb - 1) / ppb;
    if (blocks > 0x7fffffffull) return OSC_ERR_ARG;
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_a2a_mid<WW>),
                                         dim3((unsigned)blocks),
                                         dim3(kMoveBlock), 0, s, *g, dest_off,
                                         src_off, nbytes, (uint32_t)ppb));
    return check_launch();
  }
  const uint64_t cpb = (nunits + kChunk - 1) / kChunk;
  const uint64_t blocks = cpb * P * P;
  if (blocks > 0x7fffffffull) return OSC_ERR_ARG;
  OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_a2a_big<WW, UPT>),
                                       dim3((unsigned)blocks),
                                       dim3(kMoveBlock), 0, s, *g, dest_off,
                                       src_off, nbytes, (uint32_t)cpb));
  return check_launch();
}

extern "C" int oscd_copy(void *dst, const void *src, uint64_t bytes,
                         void *stream) {
  using namespace osc;
  if (bytes == 0) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int W = unit_width((uint64_t)(uintptr_t)dst | (uint64_t)(uintptr_t)src |
                           bytes);
  const uint64_t nunits = bytes / W;
  const unsigned grid = grid_for(nunits, 1u << 20);
  OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_copy<WW>), dim3(grid),
                                       dim3(kMoveBlock), 0, s, (char *)dst,
                                       (const char *)src, bytes));
  return check_launch();
}
