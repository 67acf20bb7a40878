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
//    every member j, the root included; the active-set form
//    (broadcast.c:451-470, binomial helper :150-190) skips the root;
//  * collect, default bruck (src/shcoll/src/collect.c:406-492, with the
//    prefix sum of src/shcoll/src/util/scan.c:10-60 and the final rotate):
//    dest_j[off_i .. off_i + n_i) = src_i, off_i = sum of n_k for k < i;
//  * alltoalls, default shift_exchange_barrier (src/shcoll/src/
//    alltoalls.c:54-90): dest_j[(i*n + t)*dst] = src_i[(j*n + t)*sst];
//  * strided put/get (src/api/rma/putget.h:130-186): element-wise copy.
// fcollect, broadcast and collect run destination-major, so the
// workgroups in flight read the same source chunk (served by the caches
// after its first read) and each writes one contiguous run.  Unit width W
// (16 bytes when alignment allows) is picked per call from the alignment
// of sizes, offsets and slice.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/osc.h"
#include "dev_common.h"
#include "osc_dev.h"

namespace osc {

constexpr int kMoveBlock = 256;

// Interleaved write order (k_replicate): HBM write bandwidth on
// MI355X depends on the order in which the workgroups issue their stores.
// 2 GiB of non-temporal 16-byte stores, each workgroup writing its own
// contiguous 32 KiB, run at 5.2-5.4 TB/s; S consecutive workgroups that
// interleave 1 KiB pieces (workgroup x of the group writes pieces x, x+S,
// x+2S, ... of its super-chunk) run at 6.6-7.0 TB/s (S = 64, 8 pieces per
// wave, the fastest fill measured; S = 16: 6.5-6.6), and destinations that
// change every 16-32 KiB stay at 5.5-5.8 TB/s whatever the piece size.
// Reads gain nothing from the interleave.  A piece is one wave's 64 lanes x
// 16 bytes; wave w's k-th piece is k*4 + w, so a workgroup covers 256*U
// units of the super-chunk's S*256*U.  S is a multiple of the 8 XCDs, so
// workgroup x of every interleave group runs on the same XCD.  (The same
// order applied to fcollect at 256 members x 1 MiB, 4, 8 or 16
// destinations per workgroup, was measured against k_gather at 16
// placements of the arena: within +-3% either way, so fcollect keeps
// k_gather.  Both move by up to 30% with the arena's physical placement:
// 9.8-13.0 ms for the same layout and kernel.)
//
// Code shape: a bounds test per unit puts every load and store in its own
// branch, and the compiler then waits for all outstanding memory operations
// before each store (stores count in the same counter as loads), so a wave
// has one store in flight.  Visits that lie wholly in range therefore run a
// branch-free body: all loads issue back to back and each store waits only
// for its own data.  At 64 members x 32 MiB that took the replicate from
// 332-368 us to 308-310 us (6.8 TB/s of writes).
constexpr uint64_t kRepPiece = 64;  // units (1 KiB) per piece
constexpr int kRepUpt = 16;         // replicate: units per thread per visit
constexpr uint64_t kRepWgUnits = (uint64_t)kMoveBlock * kRepUpt;  // 64 KiB
constexpr int kFillS = 64, kFillUpt = 8;  // the write-ceiling fill
constexpr int kReadUpt = 8;  // the read ceiling: 32 KiB contiguous visits
constexpr uint64_t kRepMaxGrid = 1u << 20;  // a multiple of every S

// unit of the region that lane l of wave w stores as its k-th unit, for
// workgroup x of the interleave group (S workgroups, U units per thread) of
// super-chunk sc
template <int S, int U>
__device__ __forceinline__ uint64_t il_unit(uint64_t sc, uint64_t x, int k,
                                            uint32_t w, uint32_t l) {
  const uint64_t q = (uint64_t)(k * 4) + w;
  return sc * (uint64_t)S * (kMoveBlock * U) + (q * S + x) * kRepPiece + l;
}

// Copy of UPT units per thread, units u0 + k*256: FULL (the whole chunk in
// range) takes no per-unit test, so every load is in flight before the
// first store; otherwise each unit is tested, and the compiler then waits
// for all outstanding memory operations before each store.
template <int W, int UPT, bool FULL>
__device__ __forceinline__ void move_chunk(const char *src, char *dst,
                                           uint64_t u0, uint64_t nunits) {
  typename Raw<W>::type v[UPT];
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    const uint64_t u = u0 + (uint64_t)k * kMoveBlock;
    if (FULL || u < nunits) v[k] = ld<W>(src + u * W);
  }
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    const uint64_t u = u0 + (uint64_t)k * kMoveBlock;
    if (FULL || u < nunits) st_nt<W>(dst + u * W, v[k]);
  }
}

// Gather, destination-major: member j's [doff + i*doff_step, +nbytes) =
// source member i's [soff, +nbytes), for every source i (fcollect,
// src_member < 0) or for i = src_member only (broadcast, doff_step = 0).
// Virtual block vb -> destination j = vb % P (fastest), then source i and
// chunk: the workgroups in flight all read the same few source chunks,
// which the caches serve after the first read, and each workgroup writes
// one contiguous run of 256*UPT units.  Member `skip` is left unwritten
// (the legacy broadcast leaves the root's destination alone); -1 for none.
template <int W, int UPT>
__global__ __launch_bounds__(kMoveBlock) void k_gather(
    const oscd_geom g, const int src_member, const uint64_t soff,
    const uint64_t doff, const uint64_t doff_step, const uint64_t nbytes,
    const uint64_t cps, const int skip, const uint64_t total) {
  const uint64_t nunits = nbytes / W;
  const uint64_t P = (uint64_t)g.npes;
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int j = (int)(vb % P);
    if (j == skip) continue;
    const uint64_t r = vb / P;
    const int i = src_member >= 0 ? src_member : (int)(r / cps);
    const uint64_t cc = src_member >= 0 ? r : r - (uint64_t)i * cps;
    const char *src = member_ptr(g, i, soff);
    char *dst = member_ptr(g, j, doff + (uint64_t)i * doff_step);
    const uint64_t u0 = cc * (uint64_t)(kMoveBlock * UPT) + threadIdx.x;
    if ((cc + 1) * (uint64_t)(kMoveBlock * UPT) <= nunits)
      move_chunk<W, UPT, true>(src, dst, u0, nunits);
    else
      move_chunk<W, UPT, false>(src, dst, u0, nunits);
  }
}

// fcollect with small blocks (nunits < 1024 per source): each destination's
// region of P*nunits units is contiguous (block i at i*nunits), so it is cut
// into chunks of 256*UPT units regardless of source boundaries; unit u of
// the region comes from source u / nunits.  Virtual block vb -> destination
// vb % P, chunk vb / P.  This keeps every lane busy where per-source chunks
// would leave most of a workgroup idle.  Requires P*nunits < 2^32.
template <int W, int UPT>
__global__ __launch_bounds__(kMoveBlock) void k_gather_flat(
    const oscd_geom g, const uint64_t soff, const uint64_t doff,
    const uint32_t nunits, const uint64_t total) {
  const uint64_t P = (uint64_t)g.npes;
  const uint32_t region = (uint32_t)P * nunits;
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int j = (int)(vb % P);
    const uint32_t c0 = (uint32_t)(vb / P) * (uint32_t)(kMoveBlock * UPT);
    char *dst = member_ptr(g, j, doff);
    typename Raw<W>::type v[UPT];
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
      const uint32_t u = c0 + threadIdx.x + (uint32_t)k * kMoveBlock;
      if (u < region) {
        const uint32_t i = u / nunits;
        v[k] = ld<W>(member_ptr(g, (int)i, soff) + (uint64_t)(u - i * nunits) * W);
      }
    }
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
      const uint32_t u = c0 + threadIdx.x + (uint32_t)k * kMoveBlock;
      if (u < region) st_nt<W>(dst + (uint64_t)u * W, v[k]);
    }
  }
}

// Collect with per-member sizes: member j's [doff + off_i, + n_i) =
// member i's [soff, + n_i); tab[2i] = off_i, tab[2i+1] = n_i (bytes).
// Same virtual-block order as k_gather, with cps chunks per source sized
// for the largest member; chunks past a smaller member's end exit at once.
template <int W, int UPT>
__global__ __launch_bounds__(kMoveBlock) void k_collect(
    const oscd_geom g, const uint64_t soff, const uint64_t doff,
    const uint64_t *__restrict__ tab, const uint64_t cps,
    const uint64_t total) {
  const uint64_t P = (uint64_t)g.npes;
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int j = (int)(vb % P);
    const uint64_t r = vb / P;
    const int i = (int)(r / cps);
    const uint64_t cc = r - (uint64_t)i * cps;
    const uint64_t nunits = tab[2 * i + 1] / W;
    const uint64_t c0 = cc * (uint64_t)(kMoveBlock * UPT);
    if (c0 >= nunits) continue;
    const char *src = member_ptr(g, i, soff);
    char *dst = member_ptr(g, j, doff + tab[2 * i]);
    const uint64_t u0 = c0 + threadIdx.x;
    if (c0 + (uint64_t)(kMoveBlock * UPT) <= nunits)
      move_chunk<W, UPT, true>(src, dst, u0, nunits);
    else
      move_chunk<W, UPT, false>(src, dst, u0, nunits);
  }
}

// Strided alltoall: for member pair (i source, j destination) and element
// t < n, es bytes (upe units of W) move from src_i at element
// (j*n + t)*sst to dest_j at element (i*n + t)*dst.  Virtual block vb ->
// pair vb / cpp (pair = i*P + j), chunk vb % cpp of the pair's n*upe units.
template <int W, bool ONE>
__global__ __launch_bounds__(kMoveBlock) void k_a2as(
    const oscd_geom g, const uint64_t doff, const uint64_t soff,
    const uint64_t es, const uint64_t dst, const uint64_t sst,
    const uint64_t n, const uint64_t upe, const uint64_t cpp,
    const uint64_t total) {
  const uint64_t P = (uint64_t)g.npes;
  const uint64_t units = n * upe;
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint64_t pair = vb / cpp;
    const uint64_t chunk = vb - pair * cpp;
    const int i = (int)(pair / P);
    const int j = (int)(pair - (uint64_t)i * P);
    const uint64_t u = chunk * kMoveBlock + threadIdx.x;
    if (u >= units) continue;
    const uint64_t t = ONE ? u : u / upe;
    const uint64_t w = ONE ? 0 : u - t * upe;
    const char *s = member_ptr(g, i, soff) +
                    (((uint64_t)j * n + t) * sst) * es + w * W;
    char *d = member_ptr(g, j, doff) + (((uint64_t)i * n + t) * dst) * es +
              w * W;
    st<W>(d, ld<W>(s));
  }
}

// Strided copy of n elements of upe units: dst + t*dstep <- src + t*sstep.
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_icopy(
    char *dst, const char *src, const uint64_t dstep, const uint64_t sstep,
    const uint64_t n, const uint64_t upe) {
  const uint64_t units = n * upe;
  const uint64_t gs = (uint64_t)gridDim.x * kMoveBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kMoveBlock + threadIdx.x;
       u < units; u += gs) {
    const uint64_t t = u / upe;
    const uint64_t w = u - t * upe;
    st<W>(dst + t * dstep + w * W, ld_nt<W>(src + t * sstep + w * W));
  }
}

// One 64-bit atomic signal update (set or add) by a single lane.
__global__ void k_signal(uint64_t *sig, const uint64_t value, const int op) {
  if (threadIdx.x != 0) return;
  if (op == 0)
    __hip_atomic_store(sig, value, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  else
    __hip_atomic_fetch_add(sig, value, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// Alltoall, large blocks: block b covers chunk (b % cpb) of the pair
// (b / cpb); pair p = i*P + j moves src_i[j*nbytes ..] -> dest_j[i*nbytes ..].
// Each thread moves UPT units, all loads issued before the stores.
// Loads stay cached: repeated alltoalls over one 256 MiB working set are
// partly served by the Infinity Cache, and non-temporal loads measured 21%
// slower at 64 PEs x 64 KiB.
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
  if ((uint64_t)(chunk + 1) * (kMoveBlock * UPT) <= nunits)
    move_chunk<W, UPT, true>(src, dst, u0, nunits);
  else
    move_chunk<W, UPT, false>(src, dst, u0, nunits);
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
// Non-temporal loads and stores: a 1 GiB copy measured 6.54 TB/s against
// 6.34 with cached loads on MI355X.
template <int W>
__global__ __launch_bounds__(kMoveBlock) void k_copy(char *dst,
                                                     const char *src,
                                                     const uint64_t nbytes) {
  const uint64_t nunits = nbytes / W;
  const uint64_t gs = (uint64_t)gridDim.x * kMoveBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kMoveBlock + threadIdx.x;
       u < nunits; u += gs)
    st_nt<W>(dst + u * W, ld_nt<W>(src + u * W));
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

static bool gather_il_pays(int W, uint64_t nunits);
static int gather_il(const oscd_geom *g, int src_member, uint64_t soff,
                     uint64_t doff, uint64_t doff_step, uint64_t nunits,
                     uint64_t nsrc, int skip, hipStream_t s);

// nsrc = number of source members gathered (P for fcollect, 1 for
// broadcast); src_member < 0 selects fcollect.
static int gather(const oscd_geom *g, int src_member, uint64_t soff,
                  uint64_t doff, uint64_t doff_step, uint64_t nbytes,
                  uint64_t nsrc, int skip, hipStream_t s) {
  const int W = alignment_of(g, doff | doff_step, soff, nbytes);
  const uint64_t nunits = nbytes / W;
  if (gather_il_pays(W, nunits))
    return gather_il(g, src_member, soff, doff, doff_step, nunits, nsrc, skip,
                     s);
  const bool wide = nunits >= 4 * kMoveBlock;
  if (src_member < 0 && !wide && skip < 0 && doff_step == nbytes) {
    // small fcollect blocks: chunk each destination's contiguous region
    const uint64_t region = (uint64_t)g->npes * nunits;
    const uint64_t chunks = (region + 4 * kMoveBlock - 1) / (4 * kMoveBlock);
    const uint64_t total = (uint64_t)g->npes * chunks;
    const unsigned grid = (unsigned)(total < (1u << 22) ? total : (1u << 22));
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_gather_flat<WW, 4>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g, soff,
                                         doff, (uint32_t)nunits, total));
    return check_launch();
  }
  const uint64_t per = wide ? 4 * kMoveBlock : kMoveBlock;
  const uint64_t cps = (nunits + per - 1) / per;
  const uint64_t total = (uint64_t)g->npes * nsrc * cps;
  const unsigned grid = (unsigned)(total < (1u << 22) ? total : (1u << 22));
  if (wide) {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_gather<WW, 4>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g,
                                         src_member, soff, doff, doff_step,
                                         nbytes, cps, skip, total));
  } else {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_gather<WW, 1>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g,
                                         src_member, soff, doff, doff_step,
                                         nbytes, cps, skip, total));
  }
  return check_launch();
}

}  // namespace osc

extern "C" int oscd_fcollect(const oscd_geom *g, uint64_t dest_off,
                             uint64_t src_off, uint64_t nbytes, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535) return OSC_ERR_ARG;
  if (nbytes == 0) return OSC_OK;
  return gather(g, -1, src_off, dest_off, nbytes, nbytes, (uint64_t)g->npes,
                -1, (hipStream_t)stream);
}

extern "C" int oscd_broadcast(const oscd_geom *g, int root, uint64_t dest_off,
                              uint64_t src_off, uint64_t nbytes, int skip_root,
                              void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || root < 0 || root >= g->npes)
    return OSC_ERR_ARG;
  if (nbytes == 0) return OSC_OK;
  return gather(g, root, src_off, dest_off, 0, nbytes, 1u,
                skip_root ? root : -1, (hipStream_t)stream);
}

extern "C" int oscd_collect(const oscd_geom *g, uint64_t dest_off,
                            uint64_t src_off, const uint64_t *d_tab,
                            uint64_t max_n, int unit_width, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535 || d_tab == nullptr)
    return OSC_ERR_ARG;
  if (unit_width != 1 && unit_width != 2 && unit_width != 4 &&
      unit_width != 8 && unit_width != 16)
    return OSC_ERR_ARG;
  if (max_n == 0) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int W = unit_width;
  const uint64_t nunits = max_n / W;
  const bool wide = nunits >= 4 * kMoveBlock;
  const uint64_t per = wide ? 4 * kMoveBlock : kMoveBlock;
  const uint64_t cps = (nunits + per - 1) / per;
  const uint64_t P = (uint64_t)g->npes;
  const uint64_t total = P * P * cps;
  const unsigned grid = (unsigned)(total < (1u << 22) ? total : (1u << 22));
  if (wide) {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_collect<WW, 4>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g, src_off,
                                         dest_off, d_tab, cps, total));
  } else {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_collect<WW, 1>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g, src_off,
                                         dest_off, d_tab, cps, total));
  }
  return check_launch();
}

// ---- replicate: member 0's bytes to every member -----------------------------
// The second half of a large reduction (the walk stores member 0's result
// only): members 1..P-1 receive member 0's [off, off + nbytes), in the
// interleaved write order described at the top of this file.  The S
// workgroups of an interleave group load each 1 KiB piece of member 0's
// super-chunk once (cached: the next member group re-reads it soon after)
// and store it to a group of G destinations; virtual block vb -> x = vb %
// S, then super-chunk and member group, member group fastest.  Measured on
// MI355X (64 members x 32 MiB / 16 MiB, 128 x 16 MiB, 16 x 32 MiB, 8 x 64
// MiB): 320 / 157 / 300 / 89 / 94 us against 393 / 200 / 402 / 109 / 112
// us for the older order (one 32 KiB chunk per visit, member group
// fastest, 5.3-5.5 TB/s).
namespace osc {

// One visit of k_replicate: the wave's kRepUpt pieces of super-chunk sc,
// loaded from member 0 and stored to members d0 .. d0+G-1 (those below P).
// FULL: the super-chunk lies wholly in range (no unit is tested).  skip
// (or null): one flag per 256 units, set where the walk already stored
// every member; a piece of 64 units lies in one flag's span.  Lane k
// fetches the flag of the wave's k-th piece in the same batch as the data
// loads, and the flags become a wave-uniform mask once the data loads are
// in flight.
template <int G, int S, int U, bool FULL, bool SKIP, bool NTL = false>
__device__ __forceinline__ void rep_visit(const oscd_geom &g, uint64_t off,
                                          const char *src, uint64_t nunits,
                                          uint64_t sc, uint64_t x, int d0,
                                          uint32_t w, uint32_t l,
                                          const uint8_t *skip) {
  // lanes >= U repeat the first U lanes' pieces
  uint8_t f = 0;
  bool fin = true;
  if (SKIP) {
    const uint64_t uf = il_unit<S, U>(sc, x, (int)(l % U), w, 0);
    fin = FULL || uf < nunits;
    f = skip[(fin ? uf : nunits - 1) / 256];
  }
  uint64_t u[U];
  typename Raw<16>::type v[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    u[k] = il_unit<S, U>(sc, x, k, w, l);
    if (FULL || u[k] < nunits)
      v[k] = NTL ? ld_nt<16>(src + u[k] * 16) : ld<16>(src + u[k] * 16);
  }
  __builtin_amdgcn_sched_barrier(0);
  // bit k set: the wave stores its k-th piece
  const uint32_t keep =
      SKIP ? (uint32_t)__ballot(l < U && fin && f == 0) : (1u << U) - 1;
#pragma unroll
  for (int q = 0; q < G; ++q) {
    if (d0 + q >= g.npes) break;  // the last group may be short
    char *dst = member_ptr(g, d0 + q, off);
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (((keep >> k) & 1u) != 0 && (FULL || u[k] < nunits))
        st_nt<16>(dst + u[k] * 16, v[k]);
  }
}

template <int G, int S, bool SKIP>
__global__ __launch_bounds__(kMoveBlock) void k_replicate(
    const oscd_geom g, const uint64_t off, const uint64_t nunits,
    const uint64_t ngroups, const uint64_t total,
    const uint8_t *__restrict__ skip) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const char *src = member_ptr(g, 0, off);
  // gridDim.x is a multiple of S, so every visit of a workgroup keeps its
  // place x in the interleave group
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint64_t x = vb % S, r = vb / S;
    const uint64_t gi = r % ngroups, sc = r / ngroups;
    const int d0 = 1 + (int)gi * G;  // member 0 already holds the result
    if ((sc + 1) * (uint64_t)S * kRepWgUnits <= nunits)
      rep_visit<G, S, kRepUpt, true, SKIP>(g, off, src, nunits, sc, x, d0, w,
                                           l, skip);
    else
      rep_visit<G, S, kRepUpt, false, SKIP>(g, off, src, nunits, sc, x, d0, w,
                                            l, skip);
  }
}

template <int G, int S>
static int replicate_g(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                       const uint8_t *skip, hipStream_t s) {
  const uint64_t nunits = nbytes / 16;
  const uint64_t scu = (uint64_t)S * kRepWgUnits;
  const uint64_t nsc = (nunits + scu - 1) / scu;
  const uint64_t ngroups = ((uint64_t)g->npes - 1 + G - 1) / G;
  const uint64_t total = nsc * ngroups * S;
  const uint64_t grid = total < kRepMaxGrid ? total : kRepMaxGrid;
  if (skip != nullptr)
    hipLaunchKernelGGL((k_replicate<G, S, true>), dim3((unsigned)grid),
                       dim3(kMoveBlock), 0, s, *g, off, nunits, ngroups,
                       total, skip);
  else
    hipLaunchKernelGGL((k_replicate<G, S, false>), dim3((unsigned)grid),
                       dim3(kMoveBlock), 0, s, *g, off, nunits, ngroups,
                       total, skip);
  return check_launch();
}

// Gather in the replicate's interleaved write order: the S workgroups of a
// group share one super-chunk (S*256*U units) of one (source, destination)
// pair.  Virtual block vb -> x = vb % S, then destination (fastest, so the
// source super-chunk is read from HBM once and from the XCD's L2 by the
// other destinations), then super-chunk, then source.  U, the units per
// thread, is the largest of 16, 8, 4, 2 whose super-chunk fits in a block.
// Measured on MI355X against k_gather's order (destination-major chunks,
// which fetch each source chunk once per XCD): fcollect 256 x 1 MiB 11.1-11.3
// vs 11.9-12.1 ms (U = 16; S = 16: 12.6 ms; 2, 4 or 8 destinations per
// workgroup: 11.3-12.0 ms); 256 x 256 KiB 2.69 vs 3.07-3.12 ms; 128 x 256 KiB
// 0.70 vs 0.98-1.05 ms; 128 x 128 KiB 338 vs 361-400 us; 64 x 64 KiB 45 vs
// 45-51 us.
constexpr int kGatherS = 8;

template <int S, int U>
__global__ __launch_bounds__(kMoveBlock) void k_gather_il(
    const oscd_geom g, const int src_member, const uint64_t soff,
    const uint64_t doff, const uint64_t doff_step, const uint64_t nunits,
    const uint64_t nsc, const int skip, const uint64_t total) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint64_t P = (uint64_t)g.npes;
  // gridDim.x is a multiple of S: every visit keeps its place x
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint64_t x = vb % S, r = vb / S;
    const int j = (int)(r % P);
    if (j == skip) continue;
    const uint64_t t = r / P, sc = t % nsc;
    const int i = src_member >= 0 ? src_member : (int)(t / nsc);
    const char *src = member_ptr(g, i, soff);
    const uint64_t off = doff + (uint64_t)i * doff_step;
    if ((sc + 1) * (uint64_t)S * kMoveBlock * U <= nunits)
      rep_visit<1, S, U, true, false>(g, off, src, nunits, sc, x, j, w, l,
                                      nullptr);
    else
      rep_visit<1, S, U, false, false>(g, off, src, nunits, sc, x, j, w, l,
                                       nullptr);
  }
}

// On unless turned off (oscd_set_move_interleave): 16-byte units and a
// block of at least one super-chunk at 2 units per thread (64 KiB).  (One
// unit per thread from 32 KiB measured slower than the destination-major
// chunks: 256 PEs x 40-56 KiB blocks 752-795 vs 533-648 us.)
static bool g_move_il = true;
static bool gather_il_pays(int W, uint64_t nunits) {
  return g_move_il && W == 16 &&
         nunits >= (uint64_t)kGatherS * kMoveBlock * 2;
}

template <int U>
static int gather_il_u(const oscd_geom *g, int src_member, uint64_t soff,
                       uint64_t doff, uint64_t doff_step, uint64_t nunits,
                       uint64_t nsrc, int skip, hipStream_t s) {
  const uint64_t scu = (uint64_t)kGatherS * kMoveBlock * U;
  const uint64_t nsc = (nunits + scu - 1) / scu;
  const uint64_t total = (uint64_t)g->npes * nsrc * nsc * kGatherS;
  const uint64_t grid = total < kRepMaxGrid ? total : kRepMaxGrid;
  hipLaunchKernelGGL((k_gather_il<kGatherS, U>), dim3((unsigned)grid),
                     dim3(kMoveBlock), 0, s, *g, src_member, soff, doff,
                     doff_step, nunits, nsc, skip, total);
  return check_launch();
}

static int gather_il(const oscd_geom *g, int src_member, uint64_t soff,
                     uint64_t doff, uint64_t doff_step, uint64_t nunits,
                     uint64_t nsrc, int skip, hipStream_t s) {
  constexpr uint64_t per = (uint64_t)kGatherS * kMoveBlock;  // units at U=1
  if (nunits >= per * 16)
    return gather_il_u<16>(g, src_member, soff, doff, doff_step, nunits, nsrc,
                           skip, s);
  if (nunits >= per * 8)
    return gather_il_u<8>(g, src_member, soff, doff, doff_step, nunits, nsrc,
                          skip, s);
  if (nunits >= per * 4)
    return gather_il_u<4>(g, src_member, soff, doff, doff_step, nunits, nsrc,
                          skip, s);
  return gather_il_u<2>(g, src_member, soff, doff, doff_step, nunits, nsrc,
                        skip, s);
}

// Members per workgroup for group -1: one.  64 members, width 32, per
// member 1 / 2 / 4 / 8 / 16 / 32 MiB: 13.1 / 22.7 / 43.0 / 83.5 / 162.5 /
// 320.5 us, against 13.0 / 23.2 / 47.3 / 90.8 / 181.3 / 359.5 us with 8
// members per workgroup.  From 8 MiB the reduce path picks the layout by
// measurement instead (replicate_tuned).
static int replicate_auto_group(uint64_t nbytes) {
  (void)nbytes;
  return 1;
}

}  // namespace osc

extern "C" void oscd_set_move_interleave(int on) {
  osc::g_move_il = on != 0;
}

extern "C" int oscd_replicate(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                              int group, int width, const void *skip,
                              void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || nbytes % 16 != 0) return OSC_ERR_ARG;
  if (nbytes == 0 || g->npes == 1) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const uint8_t *sk = (const uint8_t *)skip;
  if (group < 0) group = replicate_auto_group(nbytes);
  if (width == 64) {
    switch (group) {
      case 1: return replicate_g<1, 64>(g, off, nbytes, sk, s);
      case 2: return replicate_g<2, 64>(g, off, nbytes, sk, s);
      case 4: return replicate_g<4, 64>(g, off, nbytes, sk, s);
      case 16: return replicate_g<16, 64>(g, off, nbytes, sk, s);
      default: return replicate_g<8, 64>(g, off, nbytes, sk, s);
    }
  }
  if (width == 32) {
    switch (group) {
      case 1: return replicate_g<1, 32>(g, off, nbytes, sk, s);
      case 2: return replicate_g<2, 32>(g, off, nbytes, sk, s);
      case 4: return replicate_g<4, 32>(g, off, nbytes, sk, s);
      case 16: return replicate_g<16, 32>(g, off, nbytes, sk, s);
      default: return replicate_g<8, 32>(g, off, nbytes, sk, s);
    }
  }
  switch (group) {
    case 1: return replicate_g<1, 16>(g, off, nbytes, sk, s);
    case 2: return replicate_g<2, 16>(g, off, nbytes, sk, s);
    case 4: return replicate_g<4, 16>(g, off, nbytes, sk, s);
    case 16: return replicate_g<16, 16>(g, off, nbytes, sk, s);
    default: return replicate_g<8, 16>(g, off, nbytes, sk, s);
  }
}

// ---- allocation-free collect ------------------------------------------------
// The (offset, size) table of a collect in a file-scope device array
// instead of memory the library allocates.  Each GPU holds its own copy of
// the array (at its own address) until the process ends; oscd_scratch_init
// records that address per device.  Host code must use the recorded
// address, never the array's name: in host code the name is only a
// host-side stand-in.
namespace osc {

constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint64_t scratch[2 * OSCD_STATIC_TAB_MEMBERS];
static uint64_t *g_scratch[kMaxDevices];

// The table arrives by value in the kernel arguments, kTabChunk words per
// launch: no host staging buffer and no copy engine transfer.
constexpr uint32_t kTabChunk = 240;
struct TabChunk {
  uint64_t v[kTabChunk];
};

__global__ __launch_bounds__(kMoveBlock) void k_tab_put(uint64_t *tab,
                                                        const uint32_t first,
                                                        const uint32_t count,
                                                        const TabChunk c) {
  if (threadIdx.x < count) tab[first + threadIdx.x] = c.v[threadIdx.x];
}

}  // namespace osc

extern "C" int oscd_scratch_init(int dev) {
  using namespace osc;
  if (dev < 0 || dev >= kMaxDevices) return OSC_ERR_ARG;
  if (g_scratch[dev] != nullptr) return OSC_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != dev) return OSC_ERR_STATE;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess ||
      p == nullptr) {
    (void)hipGetLastError();
    return OSC_ERR_HIP;
  }
  g_scratch[dev] = (uint64_t *)p;
  return OSC_OK;
}

extern "C" int oscd_collect_static(const oscd_geom *g, uint64_t dest_off,
                                   uint64_t src_off, const uint64_t *nbytes,
                                   uint64_t max_n, int unit_width, int dev,
                                   void *stream) {
  using namespace osc;
  if (g == nullptr || nbytes == nullptr || g->npes <= 0) return OSC_ERR_ARG;
  if (g->npes > OSCD_STATIC_TAB_MEMBERS) return OSC_ERR_ARG;
  if (dev < 0 || dev >= kMaxDevices || g_scratch[dev] == nullptr)
    return OSC_ERR_STATE;
  if (max_n == 0) return OSC_OK;
  uint64_t *tab = g_scratch[dev];
  hipStream_t s = (hipStream_t)stream;
  const uint32_t words = 2u * (uint32_t)g->npes;
  uint64_t off = 0;
  for (uint32_t first = 0; first < words; first += kTabChunk) {
    TabChunk c;
    const uint32_t cnt = words - first < kTabChunk ? words - first : kTabChunk;
    for (uint32_t k = 0; k < cnt; k += 2) {  // first and kTabChunk are even
      const uint64_t n = nbytes[(first + k) / 2];
      c.v[k] = off;
      c.v[k + 1] = n;
      off += n;
    }
    hipLaunchKernelGGL(k_tab_put, dim3(1), dim3(kMoveBlock), 0, s, tab, first,
                       cnt, c);
    if (check_launch() != OSC_OK) return OSC_ERR_HIP;
  }
  return oscd_collect(g, dest_off, src_off, tab, max_n, unit_width, stream);
}

extern "C" int oscd_alltoalls(const oscd_geom *g, uint64_t dest_off,
                              uint64_t src_off, uint64_t es, uint64_t dst,
                              uint64_t sst, uint64_t nelems, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 65535 || es == 0 ||
      dst == 0 || sst == 0)
    return OSC_ERR_ARG;
  if (nelems == 0) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int W = alignment_of(g, dest_off, src_off, es);
  const uint64_t upe = es / W;
  const uint64_t P = (uint64_t)g->npes;
  const uint64_t cpp = (nelems * upe + kMoveBlock - 1) / kMoveBlock;
  const uint64_t total = P * P * cpp;
  const unsigned grid = (unsigned)(total < (1u << 22) ? total : (1u << 22));
  if (upe == 1) {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_a2as<WW, true>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g, dest_off,
                                         src_off, es, dst, sst, nelems, upe,
                                         cpp, total));
  } else {
    OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_a2as<WW, false>), dim3(grid),
                                         dim3(kMoveBlock), 0, s, *g, dest_off,
                                         src_off, es, dst, sst, nelems, upe,
                                         cpp, total));
  }
  return check_launch();
}

extern "C" int oscd_copy_strided(void *dst, const void *src, uint64_t es,
                                 uint64_t dst_step, uint64_t src_step,
                                 uint64_t n, void *stream) {
  using namespace osc;
  if (es == 0) return OSC_ERR_ARG;
  if (n == 0) return OSC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int W = unit_width((uint64_t)(uintptr_t)dst |
                           (uint64_t)(uintptr_t)src | es | dst_step |
                           src_step);
  const uint64_t upe = es / W;
  const unsigned grid = grid_for(n * upe, 1u << 16);
  OSC_W_DISPATCH(W, hipLaunchKernelGGL((k_icopy<WW>), dim3(grid),
                                       dim3(kMoveBlock), 0, s, (char *)dst,
                                       (const char *)src, dst_step, src_step,
                                       n, upe));
  return check_launch();
}

extern "C" int oscd_signal(uint64_t *sig, uint64_t value, int op,
                           void *stream) {
  using namespace osc;
  if (sig == nullptr || ((uintptr_t)sig & 7u) != 0 || (op != 0 && op != 1))
    return OSC_ERR_ARG;
  hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, (hipStream_t)stream, sig,
                     value, op);
  return check_launch();
}

// Alltoall of 16-byte blocks of 64 KiB or more in the interleaved write
// order: the 8 workgroups of a group share one super-chunk (8*256*2 units)
// of one (source i, destination j) pair; virtual block vb -> x = vb % 8,
// then super-chunk, then pair with the source fastest, so consecutive
// groups fill one destination's region.  Streaming loads when the call's
// footprint exceeds the 256 MB Infinity Cache, cached loads when it fits.
// Measured on MI355X against the pair-major chunks of k_a2a_big: 128 PEs x
// 256 KiB 1.28-1.29 vs 1.55-1.59 ms; 128 x 64 KiB 323-327 vs 394-408 us;
// 64 x 1 MiB 1.33 vs 1.56-1.57 ms; 64 x 64 KiB (256 MiB, cached loads)
// 77.6-77.7 vs 78.4-79.4 us.
namespace osc {

constexpr int kA2aS = 8, kA2aU = 2;
// Blocks of 24 KiB to 64 KiB take the interleaved order too when the loads
// stream (the footprint exceeds the Infinity Cache).  Measured on MI355X
// through osc_alltoall, interleaved vs pair-major chunks: 128 PEs x 24, 32,
// 40, 48, 56, 60 KiB 145 / 191 / 243 / 264 / 297 / 335 vs 155 / 215 / 254 /
// 311 / 358 / 384 us; 96 PEs x 40-60 KiB 134-190 vs 144-221 us (but x 32
// KiB 98 vs 87 us).  With cached loads (64 and 96 PEs below 40 KiB) and
// below 24 KiB (one partial super-chunk per pair) the pair-major chunks
// were as fast or faster: 96 PEs x 20-24 KiB 67-89 vs 62-77 us.
constexpr uint64_t kA2aMidUnits = 1536;

static bool a2a_streams(uint64_t P, uint64_t nbytes) {
  // the footprint in double: P^2 x nbytes can exceed 2^64
  return (double)P * (double)P * (double)nbytes > (double)(256ull << 20);
}

template <int S, int U, bool NTL>
__global__ __launch_bounds__(kMoveBlock) void k_a2a_il(
    const oscd_geom g, const uint64_t doff, const uint64_t soff,
    const uint64_t nbytes, const uint64_t nsc, const uint64_t total) {
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  const uint64_t P = (uint64_t)g.npes;
  const uint64_t nunits = nbytes / 16;
  // gridDim.x is a multiple of S: every visit keeps its place x
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint64_t x = vb % S, r = vb / S;
    const uint64_t sc = r % nsc, pair = r / nsc;
    const int i = (int)(pair % P), j = (int)(pair / P);
    const char *src = member_ptr(g, i, soff) + (uint64_t)j * nbytes;
    const uint64_t off = doff + (uint64_t)i * nbytes;
    if ((sc + 1) * (uint64_t)S * kMoveBlock * U <= nunits)
      rep_visit<1, S, U, true, false, NTL>(g, off, src, nunits, sc, x, j, w,
                                           l, nullptr);
    else
      rep_visit<1, S, U, false, false, NTL>(g, off, src, nunits, sc, x, j, w,
                                            l, nullptr);
  }
}

static int a2a_il(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                  uint64_t nbytes, hipStream_t s) {
  const uint64_t P = (uint64_t)g->npes, nunits = nbytes / 16;
  const uint64_t scu = (uint64_t)kA2aS * kMoveBlock * kA2aU;
  const uint64_t nsc = (nunits + scu - 1) / scu;
  const uint64_t total = P * P * nsc * kA2aS;
  const uint64_t grid = total < kRepMaxGrid ? total : kRepMaxGrid;
  if (a2a_streams(P, nbytes))
    hipLaunchKernelGGL((k_a2a_il<kA2aS, kA2aU, true>), dim3((unsigned)grid),
                       dim3(kMoveBlock), 0, s, *g, dest_off, src_off, nbytes,
                       nsc, total);
  else
    hipLaunchKernelGGL((k_a2a_il<kA2aS, kA2aU, false>), dim3((unsigned)grid),
                       dim3(kMoveBlock), 0, s, *g, dest_off, src_off, nbytes,
                       nsc, total);
  return check_launch();
}

}  // namespace osc

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
  if (g_move_il && W == 16 &&
      (nunits >= (uint64_t)kA2aS * kMoveBlock * kA2aU ||
       (nunits >= kA2aMidUnits && a2a_streams(P, nbytes))))
    return a2a_il(g, dest_off, src_off, nbytes, s);
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

// ---- read-only and write-only streams -----------------------------------------
// The two halves of every collective's traffic, alone, for measuring the
// HBM ceilings beside the collectives.  k_xor_read folds every 16-byte unit
// of a range into one value per workgroup (each workgroup reads contiguous
// 32 KiB visits, the walk's read pattern, with non-temporal loads) and
// k_xor_fold folds those into a 16-byte sink.  (Device atomics into the
// sink instead cost 50 us: the 2048 workgroups finish together and their
// atomics on one address queue up, 353-358 us against 302-306 us.)
// k_fill64 stores a 64-bit
// value over a range in the interleaved write order (64 workgroups per
// group, 8 pieces per wave).  Both are the fastest shapes measured on
// MI355X for 2 GiB: reads 300-303 us (7.1 TB/s; 16 units per thread 310 us,
// one load in flight per wave 308-312 us), writes 304-310 us.
namespace osc {

__global__ __launch_bounds__(kMoveBlock) void k_xor_read(
    const char *__restrict__ p, const uint64_t nunits,
    unsigned long long *__restrict__ part) {
  __shared__ unsigned long long red[2][kMoveBlock / 64];
  constexpr uint64_t visit = (uint64_t)kMoveBlock * kReadUpt;
  typename Raw<16>::type a = {0u, 0u, 0u, 0u};
  for (uint64_t c = blockIdx.x; c * visit < nunits; c += gridDim.x) {
    const uint64_t u0 = c * visit + threadIdx.x;
    if ((c + 1) * visit <= nunits) {
      // a whole visit: every load in flight before the first XOR waits
      typename Raw<16>::type t[kReadUpt];
#pragma unroll
      for (int k = 0; k < kReadUpt; ++k)
        t[k] = ld_nt<16>(p + (u0 + (uint64_t)k * kMoveBlock) * 16);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < kReadUpt; ++k) a ^= t[k];
    } else {
#pragma unroll
      for (int k = 0; k < kReadUpt; ++k) {
        const uint64_t u = u0 + (uint64_t)k * kMoveBlock;
        if (u < nunits) a ^= ld_nt<16>(p + u * 16);
      }
    }
  }
  unsigned long long lo = ((unsigned long long)a.y << 32) | a.x;
  unsigned long long hi = ((unsigned long long)a.w << 32) | a.z;
  for (int d = 32; d > 0; d >>= 1) {
    lo ^= __shfl_xor(lo, d);
    hi ^= __shfl_xor(hi, d);
  }
  if (threadIdx.x % 64 == 0) {
    red[0][threadIdx.x / 64] = lo;
    red[1][threadIdx.x / 64] = hi;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long v = 0;
    for (int w = 0; w < kMoveBlock / 64; ++w) v ^= red[threadIdx.x][w];
    part[2 * blockIdx.x + threadIdx.x] = v;
  }
}

// XOR of the nblk workgroups' partial values into the 16-byte sink.
__global__ __launch_bounds__(kMoveBlock) void k_xor_fold(
    const unsigned long long *__restrict__ part, const uint32_t nblk,
    unsigned long long *sink) {
  __shared__ unsigned long long red[2][kMoveBlock / 64];
  unsigned long long lo = 0, hi = 0;
  for (uint32_t b = threadIdx.x; b < nblk; b += kMoveBlock) {
    lo ^= part[2 * b];
    hi ^= part[2 * b + 1];
  }
  for (int d = 32; d > 0; d >>= 1) {
    lo ^= __shfl_xor(lo, d);
    hi ^= __shfl_xor(hi, d);
  }
  if (threadIdx.x % 64 == 0) {
    red[0][threadIdx.x / 64] = lo;
    red[1][threadIdx.x / 64] = hi;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long v = 0;
    for (int w = 0; w < kMoveBlock / 64; ++w) v ^= red[threadIdx.x][w];
    sink[threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(kMoveBlock) void k_fill64(
    char *__restrict__ p, const uint64_t nunits, const uint64_t value,
    const uint64_t total) {
  constexpr uint64_t sc_units = (uint64_t)kFillS * kMoveBlock * kFillUpt;
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  typename Raw<16>::type v;
  v.x = (uint32_t)value;
  v.y = (uint32_t)(value >> 32);
  v.z = v.x;
  v.w = v.y;
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const uint64_t x = vb % kFillS, sc = vb / kFillS;
    const bool full = (sc + 1) * sc_units <= nunits;
#pragma unroll
    for (int k = 0; k < kFillUpt; ++k) {
      const uint64_t u = il_unit<kFillS, kFillUpt>(sc, x, k, w, l);
      if (full || u < nunits) st_nt<16>(p + u * 16, v);
    }
  }
}

}  // namespace osc

// Per-device buffer of the read ceiling's per-workgroup partial values,
// allocated on first use and kept until the process ends.  Calls that
// overlap on one device (on different streams) share it and race.
namespace osc {
constexpr unsigned kReadGrid = 2048;
static std::mutex g_part_mu;
static unsigned long long *g_part[kMaxDevices];
static unsigned long long *read_partials() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_part_mu);
  if (g_part[dev] == nullptr) {
    void *p = nullptr;
    if (hipMalloc(&p, 2 * sizeof(unsigned long long) * kReadGrid) !=
        hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_part[dev] = (unsigned long long *)p;
  }
  return g_part[dev];
}
}  // namespace osc

extern "C" int oscd_xor_read(const void *p, uint64_t bytes, void *sink16,
                             void *stream) {
  using namespace osc;
  if (p == nullptr || sink16 == nullptr || bytes % 16 != 0 ||
      (uintptr_t)p % 16 != 0 || (uintptr_t)sink16 % 8 != 0)
    return OSC_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long *part = read_partials();
  if (part == nullptr) return OSC_ERR_NOMEM;
  const uint64_t nunits = bytes / 16;
  // 2048 workgroups (8 per CU) loop over the 32 KiB visits
  const uint64_t visit = (uint64_t)kMoveBlock * kReadUpt;
  const uint64_t visits = (nunits + visit - 1) / visit;
  const unsigned grid = (unsigned)(visits < kReadGrid ? visits : kReadGrid);
  if (grid > 0)
    hipLaunchKernelGGL(k_xor_read, dim3(grid), dim3(kMoveBlock), 0, s,
                       (const char *)p, nunits, part);
  hipLaunchKernelGGL(k_xor_fold, dim3(1), dim3(kMoveBlock), 0, s,
                     (const unsigned long long *)part, grid,
                     (unsigned long long *)sink16);
  return check_launch();
}

extern "C" int oscd_fill64(void *p, uint64_t bytes, uint64_t value,
                           void *stream) {
  using namespace osc;
  if (p == nullptr || bytes % 16 != 0 || (uintptr_t)p % 16 != 0)
    return OSC_ERR_ARG;
  if (bytes == 0) return OSC_OK;
  const uint64_t nunits = bytes / 16;
  constexpr uint64_t sc_units = (uint64_t)kFillS * kMoveBlock * kFillUpt;
  const uint64_t total = (nunits + sc_units - 1) / sc_units * kFillS;
  const uint64_t grid = total < kRepMaxGrid ? total : kRepMaxGrid;
  hipLaunchKernelGGL(k_fill64, dim3((unsigned)grid), dim3(kMoveBlock), 0,
                     (hipStream_t)stream, (char *)p, nunits, value, total);
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
