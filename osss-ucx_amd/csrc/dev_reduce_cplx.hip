// dev_reduce_cplx.hip - reduction kernels for float/double _Complex.
#include "dev_reduce.h"

namespace osc {

OSC_VISIT_FAMILY(visit_reduce_cplx, {
  if (dtype == OSC_T_C32) return visit_cplx_t<cf32, PolC32Acc>(op, mode, v);
  return visit_cplx_t<cf64, PolC64Acc>(op, mode, v);
})

}  // namespace osc
