// dev_reduce.hip - C-ABI entry of the team reduction (see dev_reduce.h).
#include "dev_reduce.h"

extern "C" int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode,
                           int shape, uint64_t dest_off, uint64_t src_off,
                           uint64_t n, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 8191 || op < 0 ||
      op >= OSC_OP_COUNT)
    return OSC_ERR_ARG;
  if (n == 0) return OSC_OK;
  int lg = 0;
  while ((2 << lg) <= g->npes) ++lg;
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case OSC_T_U8: case OSC_T_U16: case OSC_T_U32: case OSC_T_U64:
    case OSC_T_I8: case OSC_T_I16: case OSC_T_I32: case OSC_T_I64:
      return launch_reduce_int(g, op, dtype, lg, dest_off, src_off, n, shape, s);
    case OSC_T_F32:
      return launch_reduce_f32(g, op, fpmode, lg, dest_off, src_off, n, shape, s);
    case OSC_T_F64:
      return launch_reduce_f64(g, op, fpmode, lg, dest_off, src_off, n, shape, s);
    case OSC_T_C32: case OSC_T_C64:
      return launch_reduce_cplx(g, op, dtype, fpmode, lg, dest_off, src_off, n, shape, s);
  }
  return OSC_ERR_ARG;
}
