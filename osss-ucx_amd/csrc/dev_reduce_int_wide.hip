// dev_reduce_int_wide.hip - reduction kernels for 32/64-bit integers.
#include "dev_reduce.h"

namespace osc {

OSC_VISIT_FAMILY(visit_reduce_int_wide, {
  switch (dtype) {
    case OSC_T_U32: return visit_int_t<uint32_t, uint32_t>(op, v);
    case OSC_T_U64: return visit_int_t<uint64_t, uint64_t>(op, v);
    case OSC_T_I32: return visit_int_t<int32_t, uint32_t>(op, v);
    case OSC_T_I64: return visit_int_t<int64_t, uint64_t>(op, v);
  }
  return OSC_ERR_ARG;
})

}  // namespace osc
