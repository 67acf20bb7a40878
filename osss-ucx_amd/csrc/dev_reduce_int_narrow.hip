// dev_reduce_int_narrow.hip - reduction kernels for 8/16-bit integers.
#include "dev_reduce.h"

namespace osc {

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
                             int shape, hipStream_t s) {
  switch (dtype) {
    case OSC_T_U8: return launch_int<uint8_t, uint8_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_U16: return launch_int<uint16_t, uint16_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_I8: return launch_int<int8_t, uint8_t>(g, op, lg, doff, soff, n, shape, s);
    case OSC_T_I16: return launch_int<int16_t, uint16_t>(g, op, lg, doff, soff, n, shape, s);
  }
  return OSC_ERR_ARG;
}

}  // namespace osc
