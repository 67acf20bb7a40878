// dev_reduce_int_narrow.hip - reduction kernels for 8/16-bit integers.
#include "dev_reduce.h"

namespace osc {

OSC_VISIT_FAMILY(visit_reduce_int_narrow, {
  switch (dtype) {
    case OSC_T_U8: return visit_int_t<uint8_t, uint8_t>(op, v);
    case OSC_T_U16: return visit_int_t<uint16_t, uint16_t>(op, v);
    case OSC_T_I8: return visit_int_t<int8_t, uint8_t>(op, v);
    case OSC_T_I16: return visit_int_t<int16_t, uint16_t>(op, v);
  }
  return OSC_ERR_ARG;
})

}  // namespace osc
