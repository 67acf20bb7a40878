/*
 * osc_registry.c - collective algorithm selection.
 *
 * Keeps the reference's configuration surface: SHMEM_<OP>_REDUCE_ALGO,
 * SHMEM_<OP>_TO_ALL_ALGO, SHMEM_FCOLLECT_ALGO, SHMEM_ALLTOALL_ALGO,
 * SHMEM_BROADCAST_ALGO and the other per-collective variables hold "algo"
 * or "algo:type" (src/shmemc/env.c:187-312), and an unknown name is an
 * error (src/api/collectives/shcoll-shim.c:26-35).  The difference is
 * when it is resolved: the reference formats "algo:type" and scans its
 * registry table on every call (shcoll-shim.c:43-59, table.c:642-670);
 * here every (op, type) cell is resolved once, at init, into a table the
 * entry points index directly.
 *
 * Reduction algorithms on the GPU (src/api/collectives/table.c:441-519
 * registers the same five reference names for every reduction):
 *   "hip"          one-pass kernel; float SUM/PROD accumulate in FP64
 *                  (FP32) or double-double (FP64) and round once.  The
 *                  default.
 *   "rec_dbl"      the same kernel in the reference's recursive-doubling
 *                  combine order and the type's own arithmetic:
 *                  bit-identical to the reference default
 *                  (src/shcoll/src/reduce.c:182-293).
 *   "linear"       team-order fold (reduce.c:55-91).
 *   "binomial"     binomial tree (reduce.c:102-171), children folded in
 *                  ascending order.
 *   "rabenseifner", "rabenseifner2"
 *                  reduce-scatter + allgather (reduce.c:306-724).
 * Integer and bitwise results are the same under all of them; float
 * results follow each algorithm's combine order (osc.h, osc_fpmode_t).
 * Data movement (fcollect, alltoall, broadcast) is a permutation of bytes,
 * identical under every reference algorithm, so the reference's algorithm
 * names are accepted as aliases of the one GPU kernel.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_internal.h"

static const char *const k_op_env[OSC_OP_COUNT] = {
    "SHMEM_AND_REDUCE_ALGO", "SHMEM_OR_REDUCE_ALGO",  "SHMEM_XOR_REDUCE_ALGO",
    "SHMEM_MIN_REDUCE_ALGO", "SHMEM_MAX_REDUCE_ALGO", "SHMEM_SUM_REDUCE_ALGO",
    "SHMEM_PROD_REDUCE_ALGO"};

/* type names of the API (the part after "algo:") -> dtype */
static const struct {
  const char *name;
  int dtype;
} k_types[] = {
    {"char", OSC_T_I8},       {"schar", OSC_T_I8},       {"short", OSC_T_I16},
    {"int", OSC_T_I32},       {"long", OSC_T_I64},       {"longlong", OSC_T_I64},
    {"ptrdiff", OSC_T_I64},   {"uchar", OSC_T_U8},       {"ushort", OSC_T_U16},
    {"uint", OSC_T_U32},      {"ulong", OSC_T_U64},      {"ulonglong", OSC_T_U64},
    {"int8", OSC_T_I8},       {"int16", OSC_T_I16},      {"int32", OSC_T_I32},
    {"int64", OSC_T_I64},     {"uint8", OSC_T_U8},       {"uint16", OSC_T_U16},
    {"uint32", OSC_T_U32},    {"uint64", OSC_T_U64},     {"size", OSC_T_U64},
    {"float", OSC_T_F32},     {"double", OSC_T_F64},     {"complexf", OSC_T_C32},
    {"complexd", OSC_T_C64},
};

static const char *const k_move_aliases[] = {
    /* fcollect (src/shcoll/src/fcollect.c:32-484) */
    "linear", "all_linear", "all_linear1", "rec_dbl", "ring", "bruck",
    "bruck_no_rotate", "bruck_signal", "bruck_inplace", "neighbor_exchange",
    /* alltoall (src/shcoll/src/alltoall.c:336-346) */
    "shift_exchange_barrier", "shift_exchange_counter", "shift_exchange_signal",
    "xor_pairwise_exchange_barrier", "xor_pairwise_exchange_counter",
    "xor_pairwise_exchange_signal", "color_pairwise_exchange_barrier",
    "color_pairwise_exchange_counter", "color_pairwise_exchange_signal",
    /* broadcast (src/shcoll/src/broadcast.c:545-551) */
    "complete_tree", "binomial_tree", "knomial_tree", "knomial_tree_signal",
    "scatter_collect",
    /* collect (src/shcoll/src/collect.c:682-738) */
    "rec_dbl_signal", "simple",
    /* barrier / sync (src/shcoll/src/barrier.c:328-366) */
    "dissemination", NULL};

/* the reference's other per-collective algorithm variables
 * (src/shmemc/env.c:187-252): permutations or synchronisation, the same
 * under every algorithm, so they are only checked for a known name */
static const char *const k_other_env[] = {
    "SHMEM_BARRIER_ALGO",        "SHMEM_BARRIER_ALL_ALGO",
    "SHMEM_SYNC_ALGO",           "SHMEM_TEAM_SYNC_ALGO",
    "SHMEM_SYNC_ALL_ALGO",       "SHMEM_BROADCASTMEM_ALGO",
    "SHMEM_COLLECT_ALGO",        "SHMEM_COLLECTMEM_ALGO",
    "SHMEM_FCOLLECTMEM_ALGO",    "SHMEM_ALLTOALLMEM_ALGO",
    "SHMEM_ALLTOALLS_ALGO",      "SHMEM_ALLTOALLSMEM_ALGO",
    "SHMEM_ALLTOALL_SIZE_ALGO",  "SHMEM_ALLTOALLS_SIZE_ALGO",
    "SHMEM_COLLECT_SIZE_ALGO",   "SHMEM_FCOLLECT_SIZE_ALGO",
    "SHMEM_BROADCAST_SIZE_ALGO", NULL};

/* the active-set reductions have their own variables (env.c:257-281) */
static const char *const k_to_all_env[OSC_OP_COUNT] = {
    "SHMEM_AND_TO_ALL_ALGO", "SHMEM_OR_TO_ALL_ALGO",  "SHMEM_XOR_TO_ALL_ALGO",
    "SHMEM_MIN_TO_ALL_ALGO", "SHMEM_MAX_TO_ALL_ALGO", "SHMEM_SUM_TO_ALL_ALGO",
    "SHMEM_PROD_TO_ALL_ALGO"};

static int type_of(const char *name) {
  size_t i;
  for (i = 0; i < sizeof(k_types) / sizeof(k_types[0]); ++i)
    if (strcmp(k_types[i].name, name) == 0) return k_types[i].dtype;
  return -1;
}

/* Resolve one "algo[:type]" spec.  kind: OSC_KIND_*.  For reductions
 * *fpmode receives the float mode.  Returns OSC_OK or OSC_ERR_ALGO. */
int osc_registry_resolve(int kind, int op, const char *spec, int *fpmode) {
  char algo[64];
  const char *colon;
  size_t n;
  (void)op;
  if (spec == NULL) return OSC_ERR_ALGO;
  colon = strchr(spec, ':');
  n = colon ? (size_t)(colon - spec) : strlen(spec);
  if (n == 0 || n >= sizeof(algo)) return OSC_ERR_ALGO;
  memcpy(algo, spec, n);
  algo[n] = '\0';
  if (colon != NULL && type_of(colon + 1) < 0) return OSC_ERR_ALGO;
  if (kind == OSC_KIND_REDUCE) {
    if (strcmp(algo, "hip") == 0) {
      if (fpmode) *fpmode = OSC_FP_ACCURATE;
      return OSC_OK;
    }
    if (strcmp(algo, "rec_dbl") == 0) {
      if (fpmode) *fpmode = OSC_FP_REFERENCE;
      return OSC_OK;
    }
    if (strcmp(algo, "linear") == 0) {
      if (fpmode) *fpmode = OSC_FP_LINEAR;
      return OSC_OK;
    }
    if (strcmp(algo, "binomial") == 0) {
      if (fpmode) *fpmode = OSC_FP_BINOMIAL;
      return OSC_OK;
    }
    if (strcmp(algo, "rabenseifner") == 0 ||
        strcmp(algo, "rabenseifner2") == 0) {
      if (fpmode) *fpmode = OSC_FP_RABENSEIFNER;
      return OSC_OK;
    }
    return OSC_ERR_ALGO;
  }
  if (strcmp(algo, "hip") == 0) return OSC_OK;
  for (n = 0; k_move_aliases[n] != NULL; ++n)
    if (strcmp(algo, k_move_aliases[n]) == 0) return OSC_OK;
  return OSC_ERR_ALGO;
}

int osc_registry_init_from_env(void) {
  int op, t, mode;
  const char *fp = getenv("OSC_FP_MODE");
  const char *mv[] = {"SHMEM_FCOLLECT_ALGO", "SHMEM_ALLTOALL_ALGO",
                      "SHMEM_BROADCAST_ALGO"};
  const int kinds[] = {OSC_KIND_FCOLLECT, OSC_KIND_ALLTOALL, OSC_KIND_BROADCAST};
  osc_g.fp_default = OSC_FP_ACCURATE;
  if (fp != NULL && fp[0] != '\0') {
    if (strcmp(fp, "reference") == 0)
      osc_g.fp_default = OSC_FP_REFERENCE;
    else if (strcmp(fp, "accurate") != 0)
      return osc_set_error(OSC_ERR_ALGO, "OSC_FP_MODE=%s (want accurate|reference)", fp);
  }
  for (op = 0; op < OSC_OP_COUNT; ++op)
    for (t = 0; t < OSC_T_COUNT; ++t)
      osc_g.fpmode[op][t] = osc_g.fpmode_to_all[op][t] =
          (unsigned char)osc_g.fp_default;
  for (op = 0; op < OSC_OP_COUNT; ++op) {
    int pass;
    for (pass = 0; pass < 2; ++pass) {
      const char *name = pass == 0 ? k_op_env[op] : k_to_all_env[op];
      unsigned char(*tab)[OSC_T_COUNT] = pass == 0 ? osc_g.fpmode : osc_g.fpmode_to_all;
      const char *spec = getenv(name);
      const char *colon;
      if (spec == NULL || spec[0] == '\0') continue;
      if (osc_registry_resolve(OSC_KIND_REDUCE, op, spec, &mode) != OSC_OK)
        return osc_set_error(OSC_ERR_ALGO, "couldn't register collective \"%s\"=%s",
                             name, spec);
      colon = strchr(spec, ':');
      if (colon != NULL) {
        tab[op][type_of(colon + 1)] = (unsigned char)mode;
      } else {
        for (t = 0; t < OSC_T_COUNT; ++t) tab[op][t] = (unsigned char)mode;
      }
    }
  }
  for (t = 0; k_other_env[t] != NULL; ++t) {
    const char *spec = getenv(k_other_env[t]);
    if (spec != NULL && spec[0] != '\0' &&
        osc_registry_resolve(OSC_KIND_FCOLLECT, 0, spec, NULL) != OSC_OK)
      return osc_set_error(OSC_ERR_ALGO, "couldn't register collective \"%s\"=%s",
                           k_other_env[t], spec);
  }
  for (t = 0; t < 3; ++t) {
    const char *spec = getenv(mv[t]);
    if (spec != NULL && spec[0] != '\0' &&
        osc_registry_resolve(kinds[t], 0, spec, NULL) != OSC_OK)
      return osc_set_error(OSC_ERR_ALGO, "couldn't register collective \"%s\"=%s",
                           mv[t], spec);
  }
  return OSC_OK;
}

int osc_set_fp_mode(int mode) {
  int op, t;
  if (mode < OSC_FP_ACCURATE || mode > OSC_FP_RABENSEIFNER)
    return osc_set_error(OSC_ERR_ARG, "fp mode %d", mode);
  osc_g.fp_default = mode;
  for (op = 0; op < OSC_OP_COUNT; ++op)
    for (t = 0; t < OSC_T_COUNT; ++t)
      osc_g.fpmode[op][t] = osc_g.fpmode_to_all[op][t] = (unsigned char)mode;
  return OSC_OK;
}

int osc_get_fp_mode(void) { return osc_g.fp_default; }

int osc_algo_resolve(const char *kind, const char *spec, int *fpmode) {
  int k;
  if (kind == NULL) return OSC_ERR_ARG;
  if (strcmp(kind, "reduce") == 0)
    k = OSC_KIND_REDUCE;
  else if (strcmp(kind, "fcollect") == 0)
    k = OSC_KIND_FCOLLECT;
  else if (strcmp(kind, "alltoall") == 0)
    k = OSC_KIND_ALLTOALL;
  else if (strcmp(kind, "broadcast") == 0)
    k = OSC_KIND_BROADCAST;
  else
    return OSC_ERR_ARG;
  return osc_registry_resolve(k, 0, spec, fpmode);
}
