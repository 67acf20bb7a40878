// dev_reduce_int_wide.hip - reduction kernels for 32/64-bit integers.
#include "dev_reduce.h"

namespace osc {

// Bitwise/sum/prod use the unsigned type of the same width (two's-complement
// wrap, as the reference's C arithmetic on x86-64); MIN/MAX compare with
// the element type's own signedness.
template <class S, class U>
static int launch_int(const oscd_geom *g, int op, int lg, uint64_t doff,
                      uint64_t soff, uint64_t n, int shape, hipStream_t s) {
  switch (op) {
    case OSC_OP_AND: return launch<U, PolSeq<U, OSC_OP_AND, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_OR: return launch<U, PolSeq<U, OSC_OP_OR, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_XOR: return launch<U, PolSeq<U, OSC_OP_XOR, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_SUM: return launch<U, PolSeq<U, OSC_OP_SUM, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_PROD: return launch<U, PolSeq<U, OSC_OP_PROD, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_MIN: return launch<S, PolSeq<S, OSC_OP_MIN, false>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_MAX: return launch<S, PolSeq<S, OSC_OP_MAX, false>>(g, lg, doff, soff, n, shape, s);
  }
  return OSC_ERR_ARG;
}

int launch_reduce_int_narrow(const oscd_geom *g, int op, int dtype, int lg,
                             uint64_t doff, uint64_t soff, uint64_t n,
                             int shape, hipStream_t s);

int launch_reduce_int(const oscd_geom *g, int op, int dtype, int lg,
                      uint64_t doff, uint64_t soff, uint64_t n, int shape,
                      hipStream_t s) {
  switch (dtype) {
    case OSC_T_U32: return launch_int<uint32_t, uint32_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_U64: return launch_int<uint64_t, uint64_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_I32: return launch_int<int32_t, uint32_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_I64: return launch_int<int64_t, uint64_t>(g, op, lg, doff, soff, n, shape, s);
  }
  return launch_reduce_int_narrow(g, op, dtype, lg, doff, soff, n, shape, s);
}

}  // namespace osc
