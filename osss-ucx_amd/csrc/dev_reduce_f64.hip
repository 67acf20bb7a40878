// dev_reduce_f64.hip - reduction kernels for double.
#include "dev_reduce.h"

namespace osc {

OSC_VISIT_FAMILY(visit_reduce_f64,
                 { return visit_real_t<double, PolF64Acc>(op, mode, v); })

}  // namespace osc
