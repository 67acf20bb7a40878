// dev_amo.hip - atomic memory operations on symmetric variables.
//
// One lane of a one-wave kernel performs the operation on the target's
// device address (32 or 64 bits).  A fetching operation clears its slot's
// ticket, stores the old value, then the ticket with a system-scope release,
// into its own slot of coherent pinned host memory (a sequence lock: the
// host re-reads the ticket after the value); the host spins on the ticket
// for a bounded time and then falls back to hipStreamSynchronize, as the
// completion wait of blocking calls does.  A non-fetching operation is only
// enqueued: stream order puts it before every later operation of the
// library.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <chrono>

#include "../../include/osc.h"
#include "osc_dev.h"

namespace osc {

template <class W>
__global__ void k_amo(W *addr, const int op, const W v, const W cond,
                      uint64_t *slot, const uint64_t ticket) {
  if (threadIdx.x != 0) return;
  constexpr int kScope = __HIP_MEMORY_SCOPE_AGENT;
  W old = 0;
  switch (op) {
    case OSCD_AMO_FETCH:
      old = __hip_atomic_load(addr, __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_SET:
      __hip_atomic_store(addr, v, __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_SWAP:
      old = __hip_atomic_exchange(addr, v, __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_CSWAP:
      // old = cond on success; the value found otherwise
      old = cond;
      __hip_atomic_compare_exchange_strong(addr, &old, v, __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_ADD:
      old = __hip_atomic_fetch_add(addr, v, __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_AND:
      old = __hip_atomic_fetch_and(addr, v, __ATOMIC_RELAXED, kScope);
      break;
    case OSCD_AMO_OR:
      old = __hip_atomic_fetch_or(addr, v, __ATOMIC_RELAXED, kScope);
      break;
    default:  // OSCD_AMO_XOR
      old = __hip_atomic_fetch_xor(addr, v, __ATOMIC_RELAXED, kScope);
      break;
  }
  if (slot != nullptr) {
    // Sequence lock: the slot's ticket is cleared before the value is
    // replaced, so a host reader that finds the same ticket before and
    // after reading the value knows the value is this operation's.
    __hip_atomic_store(slot + 1, (uint64_t)0, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __hip_atomic_store(slot, (uint64_t)old, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(slot + 1, ticket, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// (old value, ticket) pairs; each fetching operation owns the pair
// ticket % kAmoSlots
constexpr uint64_t kAmoSlots = 1024;
static uint64_t *g_amo_slots;
static std::atomic<uint64_t> g_amo_ticket{0};

}  // namespace osc

extern "C" int oscd_amo_init(void) {
  using namespace osc;
  if (g_amo_slots != nullptr) return OSC_OK;
  if (hipHostMalloc((void **)&g_amo_slots, 2 * kAmoSlots * sizeof(uint64_t),
                    hipHostMallocCoherent) != hipSuccess)
    return OSC_ERR_HIP;
  memset(g_amo_slots, 0, 2 * kAmoSlots * sizeof(uint64_t));
  return OSC_OK;
}

extern "C" void oscd_amo_fini(void) {
  using namespace osc;
  if (g_amo_slots != nullptr) (void)hipHostFree(g_amo_slots);
  g_amo_slots = nullptr;
}

extern "C" int oscd_amo(void *addr, int width, int op, uint64_t value,
                        uint64_t cond, uint64_t *old, void *stream,
                        unsigned spin_us) {
  using namespace osc;
  hipStream_t s = (hipStream_t)stream;
  uint64_t *slot = nullptr;
  uint64_t ticket = 0;
  if (addr == nullptr || (width != 4 && width != 8) ||
      ((uintptr_t)addr % (uintptr_t)width) != 0 || op < OSCD_AMO_FETCH ||
      op > OSCD_AMO_XOR)
    return OSC_ERR_ARG;
  if (old != nullptr) {
    if (g_amo_slots == nullptr) return OSC_ERR_STATE;
    ticket = g_amo_ticket.fetch_add(1, std::memory_order_relaxed) + 1;
    slot = g_amo_slots + 2 * (ticket % kAmoSlots);
  }
  if (width == 4)
    hipLaunchKernelGGL(k_amo<uint32_t>, dim3(1), dim3(64), 0, s,
                       (uint32_t *)addr, op, (uint32_t)value, (uint32_t)cond,
                       slot, ticket);
  else
    hipLaunchKernelGGL(k_amo<uint64_t>, dim3(1), dim3(64), 0, s,
                       (uint64_t *)addr, op, value, cond, slot, ticket);
  if (hipGetLastError() != hipSuccess) return OSC_ERR_HIP;
  if (old == nullptr) return OSC_OK;
  const auto end =
      std::chrono::steady_clock::now() + std::chrono::microseconds(spin_us);
  while (__atomic_load_n(slot + 1, __ATOMIC_ACQUIRE) != ticket) {
    if (std::chrono::steady_clock::now() >= end) {
      if (hipStreamSynchronize(s) != hipSuccess) return OSC_ERR_HIP;
      if (__atomic_load_n(slot + 1, __ATOMIC_ACQUIRE) != ticket)
        return OSC_ERR_STATE;  // slot reused by 1024 later operations
      break;
    }
  }
  *old = __atomic_load_n(slot, __ATOMIC_RELAXED);
  // the slot was reused (1024 later operations) while it was being read
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  if (__atomic_load_n(slot + 1, __ATOMIC_RELAXED) != ticket) return OSC_ERR_STATE;
  return OSC_OK;
}
