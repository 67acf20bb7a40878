// dev_reduce_cplx.hip - reduction kernels for float/double _Complex.
#include "dev_reduce.h"

namespace osc {

template <class T, template <int> class AccPol>
static int launch_c(const oscd_geom *g, int op, int mode, int lg,
                    uint64_t doff, uint64_t soff, uint64_t n, int shape,
                    hipStream_t s) {
  switch (op) {
    case OSC_OP_SUM:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_SUM>>(g, lg, doff, soff, n, shape, s)
                 : launch<T, AccPol<OSC_OP_SUM>>(g, lg, doff, soff, n, shape, s);
    case OSC_OP_PROD:
      return mode == OSC_FP_REFERENCE
                 ? launch<T, PolTree<T, OSC_OP_PROD>>(g, lg, doff, soff, n, shape, s)
                 : launch<T, AccPol<OSC_OP_PROD>>(g, lg, doff, soff, n, shape, s);
  }
  return OSC_ERR_UNSUPPORTED;
}

int launch_reduce_cplx(const oscd_geom *g, int op, int dtype, int mode,
                       int lg, uint64_t doff, uint64_t soff, uint64_t n,
                       int shape, hipStream_t s) {
  if (dtype == OSC_T_C32)
    return launch_c<cf32, PolC32Acc>(g, op, mode, lg, doff, soff, n, shape, s);
  return launch_c<cf64, PolC64Acc>(g, op, mode, lg, doff, soff, n, shape, s);
}

}  // namespace osc
