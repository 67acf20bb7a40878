// dev_reduce_f64.hip - reduction kernels for double.
#include "dev_reduce.h"

namespace osc {

// MIN/MAX: exact selection, per-member results on NaN / signed-zero ties.
// SUM/PROD: reference-order tree in double (mode 1) or accurate (mode 0).
int launch_reduce_f64(const oscd_geom *g, int op, int mode, int lg,
                      uint64_t doff, uint64_t soff, uint64_t n, int shape,
                      hipStream_t s) {
  typedef double T;
  switch (op) {
    case OSC_OP_MIN: return launch<T, PolSeq<T, OSC_OP_MIN, true>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_MAX: return launch<T, PolSeq<T, OSC_OP_MAX, true>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_SUM:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_SUM>>(g, lg, doff, soff, n, shape, s)
                 : launch<T, PolF64Acc<OSC_OP_SUM>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_PROD:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_PROD>>(g, lg, doff, soff, n, shape, s)
                 : launch<T, PolF64Acc<OSC_OP_PROD>>(g, lg, doff, soff, n, shape, s);
  }
  return OSC_ERR_UNSUPPORTED;
}

}  // namespace osc
