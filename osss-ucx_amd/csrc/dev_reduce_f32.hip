// dev_reduce_f32.hip - reduction kernels for float.
#include "dev_reduce.h"

namespace osc {

OSC_VISIT_FAMILY(visit_reduce_f32,
                 { return visit_real_t<float, PolF32Acc>(op, mode, v); })

}  // namespace osc
