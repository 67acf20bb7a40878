/*
 * shmem.h - the OpenSHMEM 1.5 entry points this library exports.
 *
 * Only the collective hot path and what callers need to place data for it:
 * init/finalize, PE queries, the symmetric heap, teams, the typed team
 * reductions, fcollect, alltoall and broadcast, plus put/get between PEs.
 * Names, argument order and meaning match the OpenSHMEM specification and
 * the reference's declarations (include/shmem/api.h:3079-3260 reductions,
 * :3449-3467 fcollect, :3513-3530 alltoall, :3325-3344 broadcast;
 * include/shmem/teams.h:16-129 teams), so a program written against the
 * reference links against libosc.so unchanged for these calls.
 *
 * Execution model (differs from a multi-process SHMEM, same results): all
 * PEs are emulated in one process on one GPU.  A collective is called ONCE
 * by the host for the whole team; dest/source are symmetric addresses (any
 * PE's copy) and every member's dest is written.  Calls are blocking: they
 * return after the device work is complete.  Return value 0 = success,
 * nonzero = failure (an OSC_ERR_* code from <osc.h>), never abort.
 */
#ifndef OSC_SHMEM_H
#define OSC_SHMEM_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- teams (reference include/shmem/teams.h:16-30) ---------------------- */
typedef void *shmem_team_t;
extern shmem_team_t SHMEM_TEAM_WORLD;
extern shmem_team_t SHMEM_TEAM_SHARED;
extern shmem_team_t SHMEM_TEAM_INVALID;

typedef struct shmem_team_config {
  int num_contexts;
} shmem_team_config_t;

#define SHMEM_TEAM_NUM_CONTEXTS (1 << 0)

/* ---- library setup (reference src/api/init.c:100-211) ------------------- */
void shmem_init(void);
void shmem_finalize(void);
int shmem_my_pe(void);
int shmem_n_pes(void);
void shmem_barrier_all(void);
void shmem_quiet(void);
void shmem_fence(void);
void shmem_info_get_version(int *major, int *minor);
void shmem_info_get_name(char *name);

/* ---- symmetric heap (reference src/api/shmalloc.c:47-195) --------------- */
void *shmem_malloc(size_t size);
void *shmem_calloc(size_t count, size_t size);
void *shmem_align(size_t alignment, size_t size);
void *shmem_realloc(void *ptr, size_t size);
void shmem_free(void *ptr);
void *shmem_ptr(const void *dest, int pe);
int shmem_addr_accessible(const void *addr, int pe);

/* ---- teams (reference src/api/teams/teams.c, src/shmemc/teams.c) -------- */
int shmem_team_my_pe(shmem_team_t team);
int shmem_team_n_pes(shmem_team_t team);
int shmem_team_get_config(shmem_team_t team, long config_mask,
                          shmem_team_config_t *config);
int shmem_team_translate_pe(shmem_team_t src_team, int src_pe,
                            shmem_team_t dest_team);
int shmem_team_split_strided(shmem_team_t parent_team, int start, int stride,
                             int size, const shmem_team_config_t *config,
                             long config_mask, shmem_team_t *new_team);
int shmem_team_split_2d(shmem_team_t parent_team, int xrange,
                        const shmem_team_config_t *xaxis_config,
                        long xaxis_mask, shmem_team_t *xaxis_team,
                        const shmem_team_config_t *yaxis_config,
                        long yaxis_mask, shmem_team_t *yaxis_team);
void shmem_team_destroy(shmem_team_t team);
int shmem_team_sync(shmem_team_t team);

/* ---- RMA between emulated PEs (reference src/api/rma/putget.h:96-556) --- */
void shmem_putmem(void *dest, const void *source, size_t nbytes, int pe);
void shmem_getmem(void *dest, const void *source, size_t nbytes, int pe);
void shmem_putmem_nbi(void *dest, const void *source, size_t nbytes, int pe);
void shmem_getmem_nbi(void *dest, const void *source, size_t nbytes, int pe);

/* ---- type tables -------------------------------------------------------
 * The sets of <T> each collective is defined for.  The lists are those of
 * the OpenSHMEM 1.5 specification (Tables 5 and 10), as in the reference's
 * include/shmem/api_types.h:37-61, :212-226, :235-259, :268-294.  Each X()
 * entry is X(c_type, typename).                                           */
#define OSC_RMA_TYPES(X)                                                       \
  X(float, float)                                                              \
  X(double, double)                                                            \
  X(long double, longdouble)                                                   \
  X(char, char)                                                                \
  X(signed char, schar)                                                        \
  X(short, short)                                                              \
  X(int, int)                                                                  \
  X(long, long)                                                                \
  X(long long, longlong)                                                       \
  X(unsigned char, uchar)                                                      \
  X(unsigned short, ushort)                                                    \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int8_t, int8)                                                              \
  X(int16_t, int16)                                                            \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint8_t, uint8)                                                            \
  X(uint16_t, uint16)                                                          \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)                                                          \
  X(size_t, size)                                                              \
  X(ptrdiff_t, ptrdiff)

#define OSC_BITWISE_TYPES(X)                                                   \
  X(unsigned char, uchar)                                                      \
  X(unsigned short, ushort)                                                    \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int8_t, int8)                                                              \
  X(int16_t, int16)                                                            \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint8_t, uint8)                                                            \
  X(uint16_t, uint16)                                                          \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)                                                          \
  X(size_t, size)

#define OSC_INTEGER_MINMAX_ARITH_TYPES(X)                                      \
  X(char, char)                                                                \
  X(signed char, schar)                                                        \
  X(short, short)                                                              \
  X(int, int)                                                                  \
  X(long, long)                                                                \
  X(long long, longlong)                                                       \
  X(ptrdiff_t, ptrdiff)                                                        \
  X(unsigned char, uchar)                                                      \
  X(unsigned short, ushort)                                                    \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int8_t, int8)                                                              \
  X(int16_t, int16)                                                            \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint8_t, uint8)                                                            \
  X(uint16_t, uint16)                                                          \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)                                                          \
  X(size_t, size)

#define OSC_MINMAX_TYPES(X)                                                    \
  OSC_INTEGER_MINMAX_ARITH_TYPES(X)                                            \
  X(float, float)                                                              \
  X(double, double)                                                            \
  X(long double, longdouble)

#define OSC_ARITH_TYPES(X)                                                     \
  OSC_INTEGER_MINMAX_ARITH_TYPES(X)                                            \
  X(float, float)                                                              \
  X(double, double)                                                            \
  X(long double, longdouble)                                                   \
  X(double _Complex, complexd)                                                 \
  X(float _Complex, complexf)

/* ---- team reductions (reference include/shmem/api.h:3079-3260) ---------
 * long double is declared for link compatibility; on gfx950 (no 80-bit
 * type) it returns OSC_ERR_UNSUPPORTED. */
#define OSC_DECL_REDUCE(_type, _typename, _op)                                 \
  int shmem_##_typename##_##_op##_reduce(shmem_team_t team, _type *dest,       \
                                         const _type *source, size_t nreduce);
#define OSC_DECL_AND(_t, _n) OSC_DECL_REDUCE(_t, _n, and)
#define OSC_DECL_OR(_t, _n) OSC_DECL_REDUCE(_t, _n, or)
#define OSC_DECL_XOR(_t, _n) OSC_DECL_REDUCE(_t, _n, xor)
#define OSC_DECL_MIN(_t, _n) OSC_DECL_REDUCE(_t, _n, min)
#define OSC_DECL_MAX(_t, _n) OSC_DECL_REDUCE(_t, _n, max)
#define OSC_DECL_SUM(_t, _n) OSC_DECL_REDUCE(_t, _n, sum)
#define OSC_DECL_PROD(_t, _n) OSC_DECL_REDUCE(_t, _n, prod)
OSC_BITWISE_TYPES(OSC_DECL_AND)
OSC_BITWISE_TYPES(OSC_DECL_OR)
OSC_BITWISE_TYPES(OSC_DECL_XOR)
OSC_MINMAX_TYPES(OSC_DECL_MIN)
OSC_MINMAX_TYPES(OSC_DECL_MAX)
OSC_ARITH_TYPES(OSC_DECL_SUM)
OSC_ARITH_TYPES(OSC_DECL_PROD)

/* ---- data movement (reference include/shmem/api.h:3325-3530) ----------- */
#define OSC_DECL_MOVE(_type, _typename)                                        \
  int shmem_##_typename##_fcollect(shmem_team_t team, _type *dest,             \
                                   const _type *source, size_t nelems);        \
  int shmem_##_typename##_alltoall(shmem_team_t team, _type *dest,             \
                                   const _type *source, size_t nelems);        \
  int shmem_##_typename##_broadcast(shmem_team_t team, _type *dest,            \
                                    const _type *source, size_t nelems,        \
                                    int PE_root);
OSC_RMA_TYPES(OSC_DECL_MOVE)

int shmem_fcollectmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes);
int shmem_alltoallmem(shmem_team_t team, void *dest, const void *source,
                      size_t nbytes);
int shmem_broadcastmem(shmem_team_t team, void *dest, const void *source,
                       size_t nbytes, int PE_root);

#ifdef __cplusplus
}
#endif
#endif /* OSC_SHMEM_H */
