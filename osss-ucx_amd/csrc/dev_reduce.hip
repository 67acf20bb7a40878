// dev_reduce.hip - C-ABI entries of the team reduction (see dev_reduce.h):
// the per-call launch, the batched tile loop and the tile-span query.
#include "dev_reduce.h"

namespace osc {

int visit_reduce(int dtype, int op, int mode, PerCallV *pc, BatchV *bv,
                 SpanV *sv) {
  switch (dtype) {
    case OSC_T_U32: case OSC_T_U64: case OSC_T_I32: case OSC_T_I64:
      return visit_reduce_int_wide(dtype, op, mode, pc, bv, sv);
    case OSC_T_U8: case OSC_T_U16: case OSC_T_I8: case OSC_T_I16:
      return visit_reduce_int_narrow(dtype, op, mode, pc, bv, sv);
    case OSC_T_F32: return visit_reduce_f32(dtype, op, mode, pc, bv, sv);
    case OSC_T_F64: return visit_reduce_f64(dtype, op, mode, pc, bv, sv);
    case OSC_T_C32: case OSC_T_C64:
      return visit_reduce_cplx(dtype, op, mode, pc, bv, sv);
  }
  return OSC_ERR_ARG;
}

}  // namespace osc

namespace osc {
int g_split_group = -1;
}  // namespace osc

extern "C" void oscd_set_reduce_split(int enabled, int group) {
  osc::g_split_group =
      !enabled ? 0
      : (group == 1 || group == 2 || group == 4 || group == 8 || group == 16)
          ? group
          : -1;
}

static int log2_floor(int p) {
  int lg = 0;
  while ((2 << lg) <= p) ++lg;
  return lg;
}

extern "C" int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode,
                           int shape, uint64_t dest_off, uint64_t src_off,
                           uint64_t n, void *stream) {
  using namespace osc;
  if (g == nullptr || g->npes <= 0 || g->npes > 8191 || op < 0 ||
      op >= OSC_OP_COUNT)
    return OSC_ERR_ARG;
  if (n == 0) return OSC_OK;
  PerCallV v{g, log2_floor(g->npes), dest_off, src_off, n, shape,
             (hipStream_t)stream};
  return visit_reduce(dtype, op, fpmode, &v, nullptr, nullptr);
}

extern "C" int oscd_batch_reduce(int dtype, int op, int fpmode, int shape,
                                 int maxlg, const oscd_bcall *calls,
                                 const oscd_tile *tiles, uint32_t ntiles,
                                 uint32_t grid, size_t lds_bytes, char *base,
                                 uint64_t slice, void *stream) {
  using namespace osc;
  if (grid == 0) grid = 1;
  BatchV v{{calls, tiles, ntiles, grid, lds_bytes, base, slice, maxlg,
            (hipStream_t)stream},
           shape};
  return visit_reduce(dtype, op, fpmode, nullptr, &v, nullptr);
}

extern "C" int oscd_reduce_span_units(int dtype, int op, int fpmode,
                                      int shape) {
  using namespace osc;
  SpanV v{shape};
  return visit_reduce(dtype, op, fpmode, nullptr, nullptr, &v);
}

extern "C" int oscd_reduce_vec(int dtype) {
  static const int vec[OSC_T_COUNT] = {8, 8, 4, 2, 8, 8, 4, 2, 4, 2, 2, 1};
  return (dtype < 0 || dtype >= OSC_T_COUNT) ? 0 : vec[dtype];
}
