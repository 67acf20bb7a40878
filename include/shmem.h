/*
 * shmem.h - the OpenSHMEM 1.5 entry points this library exports.
 *
 * The collective hot path and what callers need around it: init/finalize,
 * PE queries, the symmetric heap, teams, the typed team reductions,
 * fcollect, alltoall, broadcast, collect and alltoalls, the pre-1.5
 * active-set forms, and RMA between PEs (typed, sized and strided put/get,
 * put-with-signal).  Names, argument order and meaning match the OpenSHMEM
 * specification and the reference's declarations (include/shmem/api.h:
 * 3079-3260 reductions, :3449-3467 fcollect, :3513-3530 alltoall,
 * :3325-3344 broadcast, :3395-3415 collect, :3565-3587 alltoalls,
 * :3262-3609 active-set forms, :403-1021 RMA and signals;
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

/* ---- communication contexts (reference include/shmem/defs.h:112-126,
 * api.h:3615-3700) ----------------------------------------------------------
 * A context is an independent ordering of RMA and atomic operations: here
 * its own device stream, so work on different contexts can overlap, and
 * shmem_ctx_quiet completes only that context's operations.
 * SHMEM_CTX_DEFAULT is the library stream that the routines without a
 * context use.  On a context made by shmem_team_create_ctx, PE numbers
 * are indices into that team.  The options are accepted and have no
 * further effect. */
typedef void *shmem_ctx_t;
extern shmem_ctx_t SHMEM_CTX_DEFAULT;
#define SHMEM_CTX_INVALID NULL
enum {
  SHMEM_CTX_SERIALIZED = 1 << 0,
  SHMEM_CTX_PRIVATE = 1 << 1,
  SHMEM_CTX_NOSTORE = 1 << 2
};
int shmem_ctx_create(long options, shmem_ctx_t *ctx);
int shmem_team_create_ctx(shmem_team_t team, long options, shmem_ctx_t *ctx);
void shmem_ctx_destroy(shmem_ctx_t ctx);
int shmem_ctx_get_team(shmem_ctx_t ctx, shmem_team_t *team);
void shmem_ctx_fence(shmem_ctx_t ctx);
void shmem_ctx_quiet(shmem_ctx_t ctx);

/* ---- library setup (reference src/api/init.c:100-211) -------------------
 * Thread levels, malloc hints and the name length take the reference's
 * values (include/shmem/defs.h:93-98, :136-139; configure.ac:18). */
enum {
  SHMEM_THREAD_SINGLE = 0,
  SHMEM_THREAD_FUNNELED,
  SHMEM_THREAD_SERIALIZED,
  SHMEM_THREAD_MULTIPLE
};
enum { SHMEM_MALLOC_ATOMICS_REMOTE = 0, SHMEM_MALLOC_SIGNAL_REMOTE };
#define SHMEM_MAJOR_VERSION 1
#define SHMEM_MINOR_VERSION 5
#define SHMEM_MAX_NAME_LEN 64

void shmem_init(void);
/* 0 on success, nonzero (no abort) when the library cannot start; the
 * level provided is SHMEM_THREAD_MULTIPLE at most in team-call mode and
 * SHMEM_THREAD_FUNNELED at most in the SPMD mode (each PE's own thread) */
int shmem_init_thread(int requested, int *provided);
void shmem_query_thread(int *provided);
void shmem_finalize(void);
int shmem_initialized(void);
int shmem_finalized(void);
/* flushes stdio and ends the process (every PE) with status */
void shmem_global_exit(int status);
int shmem_my_pe(void);
int shmem_n_pes(void);
int shmem_pe_accessible(int pe);
void shmem_barrier_all(void);
void shmem_quiet(void);
void shmem_fence(void);
void shmem_info_get_version(int *major, int *minor);
void shmem_info_get_name(char *name);
/* proposed additions the reference also has (api.h:321-331), declared
 * under PR463 as there (the library always exports them): the spec
 * version as major * 100 + minor; this library's own version 1.0.0, as a
 * number major * 10000 + minor * 100 + patch */
#ifdef PR463
void shmem_info_get_version_number(int *version);
void shmem_info_get_vendor_version(int *major, int *minor, int *patch);
void shmem_info_get_vendor_version_number(int *version);
#endif
void shmem_pcontrol(const int level, ...); /* profiling control: no effect */
/* names deprecated since OpenSHMEM 1.2 (reference api.h:58, :203, :232):
 * start_pes(n) is shmem_init (n is ignored, as the reference ignores it),
 * _my_pe/_num_pes are shmem_my_pe/shmem_n_pes */
void start_pes(int npes);
int _my_pe(void);
int _num_pes(void);

/* ---- symmetric heap (reference src/api/shmalloc.c:47-195) --------------- */
void *shmem_malloc(size_t size);
void *shmem_calloc(size_t count, size_t size);
void *shmem_align(size_t alignment, size_t size);
void *shmem_realloc(void *ptr, size_t size);
void shmem_free(void *ptr);
/* hints only describe expected use; the block is a plain shmem_malloc */
void *shmem_malloc_with_hints(size_t size, long hints);
/* names deprecated since OpenSHMEM 1.2 (reference api.h:1290-1318) */
void *shmalloc(size_t size);
void shfree(void *ptr);
void *shrealloc(void *ptr, size_t size);
void *shmemalign(size_t alignment, size_t size);
/* the device address of dest's copy on PE pe (a global PE for shmem_ptr,
 * a team index for shmem_team_ptr), NULL if pe is not one */
void *shmem_ptr(const void *dest, int pe);
void *shmem_team_ptr(shmem_team_t team, const void *dest, int pe);
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

/* ---- constants (reference include/shmem/defs.h:61-103) ------------------ */
enum shmem_cmp_constants {
  SHMEM_CMP_EQ = 0,
  SHMEM_CMP_NE,
  SHMEM_CMP_GT,
  SHMEM_CMP_LE,
  SHMEM_CMP_LT,
  SHMEM_CMP_GE
};
enum shmem_signal_ops { SHMEM_SIGNAL_SET = 0, SHMEM_SIGNAL_ADD };

/* pSync / pWrk sizes the active-set forms take.  The arrays are accepted
 * and never touched: stream order replaces the pSync protocol. */
#define SHMEM_BCAST_SYNC_SIZE 64
#define SHMEM_BARRIER_SYNC_SIZE 2
#define SHMEM_REDUCE_SYNC_SIZE 128
#define SHMEM_REDUCE_MIN_WRKDATA_SIZE 64
#define SHMEM_COLLECT_SYNC_SIZE 128
#define SHMEM_ALLTOALL_SYNC_SIZE 64
#define SHMEM_ALLTOALLS_SYNC_SIZE 64
#define SHMEM_SYNC_SIZE SHMEM_REDUCE_SYNC_SIZE
#define SHMEM_SYNC_VALUE 0L

/* ---- RMA between emulated PEs (reference src/api/rma/putget.h:96-600,
 * put_signal.h:28-290) ------------------------------------------------------
 * The local side of a put (source) or get (dest) may be host memory or a
 * symmetric address; the remote side is the symmetric object on PE pe.
 * Blocking forms return when the data has landed; _nbi forms are complete
 * at the next shmem_quiet / shmem_barrier_all.  Strides (tst, sst) are in
 * elements. */
void shmem_putmem(void *dest, const void *source, size_t nbytes, int pe);
void shmem_getmem(void *dest, const void *source, size_t nbytes, int pe);
void shmem_putmem_nbi(void *dest, const void *source, size_t nbytes, int pe);
void shmem_getmem_nbi(void *dest, const void *source, size_t nbytes, int pe);

#define OSC_DECL_RMA(_type, _typename)                                         \
  void shmem_##_typename##_put(_type *dest, const _type *src, size_t nelems,   \
                               int pe);                                        \
  void shmem_##_typename##_put_nbi(_type *dest, const _type *src,              \
                                   size_t nelems, int pe);                     \
  void shmem_##_typename##_get(_type *dest, const _type *src, size_t nelems,   \
                               int pe);                                        \
  void shmem_##_typename##_get_nbi(_type *dest, const _type *src,              \
                                   size_t nelems, int pe);                     \
  void shmem_##_typename##_p(_type *dest, _type value, int pe);                \
  _type shmem_##_typename##_g(const _type *src, int pe);                       \
  void shmem_##_typename##_iput(_type *dest, const _type *src, ptrdiff_t tst,  \
                                ptrdiff_t sst, size_t nelems, int pe);         \
  void shmem_##_typename##_iget(_type *dest, const _type *src, ptrdiff_t tst,  \
                                ptrdiff_t sst, size_t nelems, int pe);         \
  void shmem_##_typename##_put_signal(_type *dest, const _type *src,           \
                                      size_t nelems, uint64_t *sig_addr,       \
                                      uint64_t signal, int sig_op, int pe);    \
  void shmem_##_typename##_put_signal_nbi(_type *dest, const _type *src,       \
                                          size_t nelems, uint64_t *sig_addr,   \
                                          uint64_t signal, int sig_op,         \
                                          int pe);
#define OSC_DECL_SIZED_RMA(_size)                                              \
  void shmem_put##_size(void *dest, const void *src, size_t nelems, int pe);   \
  void shmem_put##_size##_nbi(void *dest, const void *src, size_t nelems,      \
                              int pe);                                         \
  void shmem_get##_size(void *dest, const void *src, size_t nelems, int pe);   \
  void shmem_get##_size##_nbi(void *dest, const void *src, size_t nelems,      \
                              int pe);                                         \
  void shmem_iput##_size(void *dest, const void *src, ptrdiff_t tst,           \
                         ptrdiff_t sst, size_t nelems, int pe);                \
  void shmem_iget##_size(void *dest, const void *src, ptrdiff_t tst,           \
                         ptrdiff_t sst, size_t nelems, int pe);                \
  void shmem_put##_size##_signal(void *dest, const void *src, size_t nelems,   \
                                 uint64_t *sig_addr, uint64_t signal,          \
                                 int sig_op, int pe);                          \
  void shmem_put##_size##_signal_nbi(void *dest, const void *src,              \
                                     size_t nelems, uint64_t *sig_addr,        \
                                     uint64_t signal, int sig_op, int pe);
#define OSC_SIZED_RMA(X) X(8) X(16) X(32) X(64) X(128)
#define OSC_DECL_CTX_RMA(_type, _typename)                                     \
  void shmem_ctx_##_typename##_put(shmem_ctx_t ctx, _type *dest,               \
      const _type *src, size_t nelems, int pe);                                \
  void shmem_ctx_##_typename##_put_nbi(shmem_ctx_t ctx, _type *dest,           \
      const _type *src, size_t nelems, int pe);                                \
  void shmem_ctx_##_typename##_get(shmem_ctx_t ctx, _type *dest,               \
      const _type *src, size_t nelems, int pe);                                \
  void shmem_ctx_##_typename##_get_nbi(shmem_ctx_t ctx, _type *dest,           \
      const _type *src, size_t nelems, int pe);                                \
  void shmem_ctx_##_typename##_p(shmem_ctx_t ctx, _type *dest, _type value,    \
      int pe);                                                                 \
  _type shmem_ctx_##_typename##_g(shmem_ctx_t ctx, const _type *src,           \
      int pe);                                                                 \
  void shmem_ctx_##_typename##_iput(shmem_ctx_t ctx, _type *dest,              \
      const _type *src, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,           \
      int pe);                                                                 \
  void shmem_ctx_##_typename##_iget(shmem_ctx_t ctx, _type *dest,              \
      const _type *src, ptrdiff_t tst, ptrdiff_t sst, size_t nelems,           \
      int pe);                                                                 \
  void shmem_ctx_##_typename##_put_signal(shmem_ctx_t ctx, _type *dest,        \
      const _type *src, size_t nelems, uint64_t *sig_addr, uint64_t signal,    \
      int sig_op, int pe);                                                     \
  void shmem_ctx_##_typename##_put_signal_nbi(shmem_ctx_t ctx, _type *dest,    \
      const _type *src, size_t nelems, uint64_t *sig_addr, uint64_t signal,    \
      int sig_op, int pe);
#define OSC_DECL_CTX_SIZED_RMA(_size)                                          \
  void shmem_ctx_put##_size(shmem_ctx_t ctx, void *dest, const void *src,      \
      size_t nelems, int pe);                                                  \
  void shmem_ctx_put##_size##_nbi(shmem_ctx_t ctx, void *dest,                 \
      const void *src, size_t nelems, int pe);                                 \
  void shmem_ctx_get##_size(shmem_ctx_t ctx, void *dest, const void *src,      \
      size_t nelems, int pe);                                                  \
  void shmem_ctx_get##_size##_nbi(shmem_ctx_t ctx, void *dest,                 \
      const void *src, size_t nelems, int pe);                                 \
  void shmem_ctx_iput##_size(shmem_ctx_t ctx, void *dest, const void *src,     \
      ptrdiff_t tst, ptrdiff_t sst, size_t nelems, int pe);                    \
  void shmem_ctx_iget##_size(shmem_ctx_t ctx, void *dest, const void *src,     \
      ptrdiff_t tst, ptrdiff_t sst, size_t nelems, int pe);                    \
  void shmem_ctx_put##_size##_signal(shmem_ctx_t ctx, void *dest,              \
      const void *src, size_t nelems, uint64_t *sig_addr, uint64_t signal,     \
      int sig_op, int pe);                                                     \
  void shmem_ctx_put##_size##_signal_nbi(shmem_ctx_t ctx, void *dest,          \
      const void *src, size_t nelems, uint64_t *sig_addr, uint64_t signal,     \
      int sig_op, int pe);
void shmem_ctx_putmem(shmem_ctx_t ctx, void *dest, const void *source,
                      size_t nbytes, int pe);
void shmem_ctx_getmem(shmem_ctx_t ctx, void *dest, const void *source,
                      size_t nbytes, int pe);
void shmem_ctx_putmem_nbi(shmem_ctx_t ctx, void *dest, const void *source,
                          size_t nbytes, int pe);
void shmem_ctx_getmem_nbi(shmem_ctx_t ctx, void *dest, const void *source,
                          size_t nbytes, int pe);
void shmem_ctx_putmem_signal(shmem_ctx_t ctx, void *dest, const void *source,
                             size_t nbytes, uint64_t *sig_addr,
                             uint64_t signal, int sig_op, int pe);
void shmem_ctx_putmem_signal_nbi(shmem_ctx_t ctx, void *dest,
                                 const void *source, size_t nbytes,
                                 uint64_t *sig_addr, uint64_t signal,
                                 int sig_op, int pe);

void shmem_putmem_signal(void *dest, const void *source, size_t nbytes,
                         uint64_t *sig_addr, uint64_t signal, int sig_op,
                         int pe);
void shmem_putmem_signal_nbi(void *dest, const void *source, size_t nbytes,
                             uint64_t *sig_addr, uint64_t signal, int sig_op,
                             int pe);
/* the calling PE's signal word (reference src/api/rma/put_signal.c:351) */
uint64_t shmem_signal_fetch(const uint64_t *sig_addr);
/* waits until the calling PE's signal word satisfies cmp (reference
 * src/api/pt2pt_synch/waituntil.c:121-127) */
uint64_t shmem_signal_wait_until(uint64_t *sig_addr, int cmp,
                                 uint64_t cmp_value);

/* ---- point-to-point synchronization (reference include/shmem/api.h:
 * 1472-2020) ---------------------------------------------------------------
 * On symmetric variables of the calling PE's copy (device memory, read by
 * the host).  Elements whose status entry is nonzero are excluded; the
 * _any forms return SIZE_MAX and the _some forms 0 when nothing is
 * included.  The wait forms block until another thread's put satisfies
 * them.  Types: the standard AMO table (api_types.h:93-105), plus short
 * and unsigned short for the deprecated test/wait_until/wait. */
#define OSC_P2P_SYNC_TYPES(X)                                                  \
  X(int, int)                                                                  \
  X(long, long)                                                                \
  X(long long, longlong)                                                       \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)                                                          \
  X(size_t, size)                                                              \
  X(ptrdiff_t, ptrdiff)
#define OSC_P2P_DEPR_TYPES(X) X(short, short) X(unsigned short, ushort)
#define OSC_DECL_TEST_WAIT(_type, _typename)                                   \
  int shmem_##_typename##_test(_type *ivar, int cmp, _type cmp_value);         \
  void shmem_##_typename##_wait_until(_type *ivar, int cmp, _type cmp_value);  \
  void shmem_##_typename##_wait(_type *ivar, _type cmp_value);
#define OSC_DECL_P2P_SETS(_type, _typename)                                    \
  int shmem_##_typename##_test_all(_type *ivars, size_t nelems,                \
                                   const int *status, int cmp,                 \
                                   _type cmp_value);                           \
  size_t shmem_##_typename##_test_any(_type *ivars, size_t nelems,             \
                                      const int *status, int cmp,              \
                                      _type cmp_value);                        \
  size_t shmem_##_typename##_test_some(_type *ivars, size_t nelems,            \
                                       size_t *indices, const int *status,     \
                                       int cmp, _type cmp_value);              \
  void shmem_##_typename##_wait_until_all(_type *ivars, size_t nelems,         \
                                          const int *status, int cmp,          \
                                          _type cmp_value);                    \
  size_t shmem_##_typename##_wait_until_any(_type *ivars, size_t nelems,       \
                                            const int *status, int cmp,        \
                                            _type cmp_value);                  \
  size_t shmem_##_typename##_wait_until_some(_type *ivars, size_t nelems,      \
                                             size_t *indices,                  \
                                             const int *status, int cmp,       \
                                             _type cmp_value);                 \
  int shmem_##_typename##_test_all_vector(_type *ivars, size_t nelems,         \
                                          const int *status, int cmp,          \
                                          _type *cmp_values);                  \
  size_t shmem_##_typename##_test_any_vector(_type *ivars, size_t nelems,      \
                                             const int *status, int cmp,       \
                                             _type *cmp_values);               \
  size_t shmem_##_typename##_test_some_vector(                                 \
      _type *ivars, size_t nelems, size_t *indices, const int *status,         \
      int cmp, _type *cmp_values);                                             \
  void shmem_##_typename##_wait_until_all_vector(_type *ivars, size_t nelems,  \
                                                 const int *status, int cmp,   \
                                                 _type *cmp_values);           \
  size_t shmem_##_typename##_wait_until_any_vector(                            \
      _type *ivars, size_t nelems, const int *status, int cmp,                 \
      _type *cmp_values);                                                      \
  size_t shmem_##_typename##_wait_until_some_vector(                           \
      _type *ivars, size_t nelems, size_t *indices, const int *status,         \
      int cmp, _type *cmp_values);
OSC_P2P_SYNC_TYPES(OSC_DECL_TEST_WAIT)
OSC_P2P_DEPR_TYPES(OSC_DECL_TEST_WAIT)
OSC_P2P_SYNC_TYPES(OSC_DECL_P2P_SETS)

/* ---- atomic memory operations (reference include/shmem/api.h:2030-2990,
 * type tables api_types.h:93-154) ------------------------------------------
 * On the 4- or 8-byte symmetric word dest (source) of PE pe.  The fetching
 * forms return the value found and wait for it; the others are ordered
 * before every later operation of the library (shmem_quiet completes
 * them).  The _nbi forms store the fetched value to *fetch (host memory or
 * a symmetric address) before they return.  Errors are recorded
 * (osc_last_error) and the fetching forms then return 0. */
#define OSC_AMO_STD_TYPES(X)                                                   \
  X(int, int)                                                                  \
  X(long, long)                                                                \
  X(long long, longlong)                                                       \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)                                                          \
  X(size_t, size)                                                              \
  X(ptrdiff_t, ptrdiff)
#define OSC_AMO_EXT_TYPES(X) X(float, float) X(double,                         \
      double) OSC_AMO_STD_TYPES(X)
#define OSC_AMO_BITWISE_TYPES(X)                                               \
  X(unsigned int, uint)                                                        \
  X(unsigned long, ulong)                                                      \
  X(unsigned long long, ulonglong)                                             \
  X(int32_t, int32)                                                            \
  X(int64_t, int64)                                                            \
  X(uint32_t, uint32)                                                          \
  X(uint64_t, uint64)
#define OSC_DECL_AMO_EXT(_type, _typename)                                     \
  _type shmem_##_typename##_atomic_fetch(const _type *source, int pe);         \
  void shmem_##_typename##_atomic_set(_type *dest, _type value, int pe);       \
  _type shmem_##_typename##_atomic_swap(_type *dest, _type value, int pe);     \
  void shmem_##_typename##_atomic_fetch_nbi(_type *fetch, const _type *source, \
                                            int pe);                           \
  void shmem_##_typename##_atomic_swap_nbi(_type *fetch, _type *dest,          \
                                           _type value, int pe);
#define OSC_DECL_AMO_STD(_type, _typename)                                     \
  _type shmem_##_typename##_atomic_compare_swap(_type *dest, _type cond,       \
                                                _type value, int pe);          \
  _type shmem_##_typename##_atomic_fetch_inc(_type *dest, int pe);             \
  void shmem_##_typename##_atomic_inc(_type *dest, int pe);                    \
  _type shmem_##_typename##_atomic_fetch_add(_type *dest, _type value,         \
                                             int pe);                          \
  void shmem_##_typename##_atomic_add(_type *dest, _type value, int pe);       \
  void shmem_##_typename##_atomic_compare_swap_nbi(                            \
      _type *fetch, _type *dest, _type cond, _type value, int pe);             \
  void shmem_##_typename##_atomic_fetch_inc_nbi(_type *fetch, _type *dest,     \
                                                int pe);                       \
  void shmem_##_typename##_atomic_fetch_add_nbi(_type *fetch,                  \
                                                const _type *dest,             \
                                                _type value, int pe);
#define OSC_DECL_AMO_BITWISE(_type, _typename)                                 \
  _type shmem_##_typename##_atomic_fetch_and(_type *dest, _type value,         \
                                             int pe);                          \
  _type shmem_##_typename##_atomic_fetch_or(_type *dest, _type value, int pe); \
  _type shmem_##_typename##_atomic_fetch_xor(_type *dest, _type value,         \
                                             int pe);                          \
  void shmem_##_typename##_atomic_and(_type *dest, _type value, int pe);       \
  void shmem_##_typename##_atomic_or(_type *dest, _type value, int pe);        \
  void shmem_##_typename##_atomic_xor(_type *dest, _type value, int pe);       \
  void shmem_##_typename##_atomic_fetch_and_nbi(_type *fetch, _type *dest,     \
                                                _type value, int pe);          \
  void shmem_##_typename##_atomic_fetch_or_nbi(_type *fetch, _type *dest,      \
                                               _type value, int pe);           \
  void shmem_##_typename##_atomic_fetch_xor_nbi(_type *fetch, _type *dest,     \
                                                _type value, int pe);          \
  /* the reference also declares this name, with a fetch argument */          \
  void shmem_##_typename##_atomic_xor_nbi(_type *fetch, _type *dest,           \
                                          _type value, int pe);
OSC_AMO_EXT_TYPES(OSC_DECL_AMO_EXT)
OSC_AMO_STD_TYPES(OSC_DECL_AMO_STD)
OSC_AMO_BITWISE_TYPES(OSC_DECL_AMO_BITWISE)
#define OSC_DECL_CTX_AMO_EXT(_type, _typename)                                 \
  _type shmem_ctx_##_typename##_atomic_fetch(shmem_ctx_t ctx,                  \
      const _type *source, int pe);                                            \
  void shmem_ctx_##_typename##_atomic_set(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe);                                                    \
  _type shmem_ctx_##_typename##_atomic_swap(shmem_ctx_t ctx, _type *dest,      \
      _type value, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_fetch_nbi(shmem_ctx_t ctx,               \
      _type *fetch, const _type *source, int pe);                              \
  void shmem_ctx_##_typename##_atomic_swap_nbi(shmem_ctx_t ctx,                \
      _type *fetch, _type *dest, _type value, int pe);
#define OSC_DECL_CTX_AMO_STD(_type, _typename)                                 \
  _type shmem_ctx_##_typename##_atomic_compare_swap(shmem_ctx_t ctx,           \
      _type *dest, _type cond, _type value, int pe);                           \
  _type shmem_ctx_##_typename##_atomic_fetch_inc(shmem_ctx_t ctx,              \
      _type *dest, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_inc(shmem_ctx_t ctx, _type *dest,        \
      int pe);                                                                 \
  _type shmem_ctx_##_typename##_atomic_fetch_add(shmem_ctx_t ctx,              \
      _type *dest, _type value, int pe);                                       \
  void shmem_ctx_##_typename##_atomic_add(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_compare_swap_nbi(shmem_ctx_t ctx,        \
       _type *fetch, _type *dest, _type cond, _type value, int pe);            \
  void shmem_ctx_##_typename##_atomic_fetch_inc_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, int pe);                                      \
  void shmem_ctx_##_typename##_atomic_fetch_add_nbi(shmem_ctx_t ctx,           \
      _type *fetch, const _type *dest, _type value, int pe);
#define OSC_DECL_CTX_AMO_BITWISE(_type, _typename)                             \
  _type shmem_ctx_##_typename##_atomic_fetch_and(shmem_ctx_t ctx,              \
      _type *dest, _type value, int pe);                                       \
  _type shmem_ctx_##_typename##_atomic_fetch_or(shmem_ctx_t ctx,               \
      _type *dest, _type value, int pe);                                       \
  _type shmem_ctx_##_typename##_atomic_fetch_xor(shmem_ctx_t ctx,              \
      _type *dest, _type value, int pe);                                       \
  void shmem_ctx_##_typename##_atomic_and(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_or(shmem_ctx_t ctx, _type *dest,         \
      _type value, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_xor(shmem_ctx_t ctx, _type *dest,        \
      _type value, int pe);                                                    \
  void shmem_ctx_##_typename##_atomic_fetch_and_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, _type value, int pe);                         \
  void shmem_ctx_##_typename##_atomic_fetch_or_nbi(shmem_ctx_t ctx,            \
      _type *fetch, _type *dest, _type value, int pe);                         \
  void shmem_ctx_##_typename##_atomic_fetch_xor_nbi(shmem_ctx_t ctx,           \
      _type *fetch, _type *dest, _type value, int pe);                         \
  void shmem_ctx_##_typename##_atomic_xor_nbi(shmem_ctx_t ctx, _type *fetch,   \
      _type *dest, _type value, int pe);
OSC_AMO_EXT_TYPES(OSC_DECL_CTX_AMO_EXT)
OSC_AMO_STD_TYPES(OSC_DECL_CTX_AMO_STD)
OSC_AMO_BITWISE_TYPES(OSC_DECL_CTX_AMO_BITWISE)
/* deprecated names (OpenSHMEM 1.3): int, long, long long; float and double
 * for fetch/set/swap */
#define OSC_AMO_DEPR_INT_TYPES(X) X(int, int) X(long, long) X(long long,       \
      longlong)
#define OSC_DECL_AMO_DEPR_EXT(_type, _typename)                                \
  _type shmem_##_typename##_fetch(const _type *target, int pe);                \
  void shmem_##_typename##_set(_type *target, _type value, int pe);            \
  _type shmem_##_typename##_swap(_type *target, _type value, int pe);
#define OSC_DECL_AMO_DEPR_INT(_type, _typename)                                \
  _type shmem_##_typename##_cswap(_type *target, _type cond, _type value,      \
                                  int pe);                                     \
  _type shmem_##_typename##_finc(_type *target, int pe);                       \
  void shmem_##_typename##_inc(_type *target, int pe);                         \
  _type shmem_##_typename##_fadd(_type *target, _type value, int pe);          \
  void shmem_##_typename##_add(_type *target, _type value, int pe);
OSC_AMO_DEPR_INT_TYPES(OSC_DECL_AMO_DEPR_EXT)
OSC_AMO_DEPR_INT_TYPES(OSC_DECL_AMO_DEPR_INT)
OSC_DECL_AMO_DEPR_EXT(float, float)
OSC_DECL_AMO_DEPR_EXT(double, double)

/* ---- distributed locks (reference include/shmem/api.h:3007-3069) --------
 * lock is a symmetric long, zero when free; the lock word is PE 0's copy.
 * shmem_test_lock returns 0 when it took the lock and 1 when another PE
 * holds it; shmem_clear_lock completes the critical section's stores
 * first. */
void shmem_set_lock(long *lock);
void shmem_clear_lock(long *lock);
int shmem_test_lock(long *lock);

/* ---- deprecated cache routines: no effect (reference
 * include/shmem/cache.h) ---------------------------------------------------- */
void shmem_clear_cache_inv(void);
void shmem_set_cache_inv(void);
void shmem_clear_cache_line_inv(void *dest);
void shmem_set_cache_line_inv(void *dest);
void shmem_udcflush(void);
void shmem_udcflush_line(void *dest);

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

OSC_RMA_TYPES(OSC_DECL_RMA)
OSC_SIZED_RMA(OSC_DECL_SIZED_RMA)
OSC_RMA_TYPES(OSC_DECL_CTX_RMA)
OSC_SIZED_RMA(OSC_DECL_CTX_SIZED_RMA)

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

/* collect (reference include/shmem/api.h:3395-3415): every member's dest
 * receives the members' sources concatenated in team order.  In team-call
 * mode one call stands for all members, so nelems applies to each of them;
 * osc_collect() in <osc.h> takes one size per member (the variable-size
 * case).  alltoalls (api.h:3565-3587): strided alltoall, dst/sst in
 * elements; alltoallsmem moves one element of nelems bytes per pair, as
 * the reference's front end does (src/shcoll/src/alltoalls.c:391-428). */
#define OSC_DECL_MOVE2(_type, _typename)                                       \
  int shmem_##_typename##_collect(shmem_team_t team, _type *dest,              \
                                  const _type *source, size_t nelems);         \
  int shmem_##_typename##_alltoalls(shmem_team_t team, _type *dest,            \
                                    const _type *source, ptrdiff_t dst,        \
                                    ptrdiff_t sst, size_t nelems);
OSC_RMA_TYPES(OSC_DECL_MOVE2)
int shmem_collectmem(shmem_team_t team, void *dest, const void *source,
                     size_t nelems);
int shmem_alltoallsmem(shmem_team_t team, void *dest, const void *source,
                       ptrdiff_t dst, ptrdiff_t sst, size_t nelems);

/* ---- active-set forms, deprecated in OpenSHMEM 1.5 ---------------------
 * (reference include/shmem/api.h:3262-3607, :1040-1140).  The active set
 * is global PEs PE_start + k * 2^logPE_stride, k < PE_size; PE_root of
 * broadcast is an index into it and the root's target is not written.
 * These return void, as declared; a failure is reported by
 * osc_last_error() and leaves target untouched.  pWrk and pSync are not
 * used. */
#define OSC_DECL_TO_ALL(_type, _typename, _op)                                 \
  void shmem_##_typename##_##_op##_to_all(                                     \
      _type *target, const _type *source, int nreduce, int PE_start,           \
      int logPE_stride, int PE_size, _type *pWrk, long *pSync);
/* type tables of the to_all forms (reference api_types.h:164-205) */
#define OSC_TO_ALL_BITWISE_TYPES(X)                                            \
  X(short, short) X(int, int) X(long, long) X(long long, longlong)
#define OSC_TO_ALL_MINMAX_TYPES(X)                                             \
  OSC_TO_ALL_BITWISE_TYPES(X)                                                  \
  X(float, float) X(double, double) X(long double, longdouble)
#define OSC_TO_ALL_ARITH_TYPES(X)                                              \
  OSC_TO_ALL_MINMAX_TYPES(X)                                                   \
  X(double _Complex, complexd) X(float _Complex, complexf)
#define OSC_DECL_AND_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, and)
#define OSC_DECL_OR_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, or)
#define OSC_DECL_XOR_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, xor)
#define OSC_DECL_MIN_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, min)
#define OSC_DECL_MAX_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, max)
#define OSC_DECL_SUM_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, sum)
#define OSC_DECL_PROD_TO_ALL(_t, _n) OSC_DECL_TO_ALL(_t, _n, prod)
OSC_TO_ALL_BITWISE_TYPES(OSC_DECL_AND_TO_ALL)
OSC_TO_ALL_BITWISE_TYPES(OSC_DECL_OR_TO_ALL)
OSC_TO_ALL_BITWISE_TYPES(OSC_DECL_XOR_TO_ALL)
OSC_TO_ALL_MINMAX_TYPES(OSC_DECL_MIN_TO_ALL)
OSC_TO_ALL_MINMAX_TYPES(OSC_DECL_MAX_TO_ALL)
OSC_TO_ALL_ARITH_TYPES(OSC_DECL_SUM_TO_ALL)
OSC_TO_ALL_ARITH_TYPES(OSC_DECL_PROD_TO_ALL)

#define OSC_DECL_SIZED_COLL(_size)                                             \
  void shmem_fcollect##_size(void *target, const void *source, size_t nelems, \
                             int PE_start, int logPE_stride, int PE_size,     \
                             long *pSync);                                     \
  void shmem_collect##_size(void *target, const void *source, size_t nelems,  \
                            int PE_start, int logPE_stride, int PE_size,      \
                            long *pSync);                                      \
  void shmem_alltoall##_size(void *target, const void *source, size_t nelems, \
                             int PE_start, int logPE_stride, int PE_size,     \
                             long *pSync);                                     \
  void shmem_alltoalls##_size(void *target, const void *source,               \
                              ptrdiff_t dst, ptrdiff_t sst, size_t nelems,    \
                              int PE_start, int logPE_stride, int PE_size,    \
                              long *pSync);                                    \
  void shmem_broadcast##_size(void *target, const void *source,               \
                              size_t nelems, int PE_root, int PE_start,       \
                              int logPE_stride, int PE_size, long *pSync);
OSC_DECL_SIZED_COLL(32)
OSC_DECL_SIZED_COLL(64)

void shmem_barrier(int PE_start, int logPE_stride, int PE_size, long *pSync);
void shmem_sync_deprecated(int PE_start, int logPE_stride, int PE_size,
                           long *pSync);
void shmem_sync_all(void);
/* shmem_sync(team) or shmem_sync(PE_start, logPE_stride, PE_size, pSync),
 * chosen by argument count as in the reference (api.h:3711-3729) */
#define OSC_SYNC_PICK(a, b, c, d, e, ...) e
#define shmem_sync(...)                                                        \
  OSC_SYNC_PICK(__VA_ARGS__, shmem_sync_deprecated, shmem_sync_deprecated,     \
                shmem_sync_deprecated, shmem_team_sync)                        \
  (__VA_ARGS__)

/* ---- C11 type-generic routines -----------------------------------------
 * The OpenSHMEM 1.5 type-generic forms (reference include/shmem/generics.h)
 * over the typed routines above.  The routine is chosen by the element type
 * of the destination (of the source for shmem_g): the controlling
 * expression *(p) is not evaluated, and its lvalue conversion drops
 * qualifiers.  Only distinct C types are listed.  On LP64 Linux every
 * <stdint.h> and <stddef.h> type is a typedef of one of them (int64_t of
 * long, size_t of unsigned long, ...), so it selects the routine of that C
 * type, which has the same element size and arithmetic.  An element type
 * outside a routine's table does not compile.  The RMA, atomic and
 * signal forms also take a leading shmem_ctx_t, which selects the
 * routine's shmem_ctx_ form (as the reference's generics do). */
#if defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L && !defined(__cplusplus)
#define OSC_GEN_INTS_P(_pre, _op)                                              \
  char: _pre##char_##_op, signed char: _pre##schar_##_op,                      \
  short: _pre##short_##_op, int: _pre##int_##_op, long: _pre##long_##_op,      \
  long long: _pre##longlong_##_op, unsigned char: _pre##uchar_##_op,           \
  unsigned short: _pre##ushort_##_op, unsigned int: _pre##uint_##_op,          \
  unsigned long: _pre##ulong_##_op, unsigned long long: _pre##ulonglong_##_op
#define OSC_GEN_REALS_P(_pre, _op)                                             \
  float: _pre##float_##_op, double: _pre##double_##_op,                        \
  long double: _pre##longdouble_##_op
#define OSC_GEN_INTS(_op) OSC_GEN_INTS_P(shmem_, _op)
#define OSC_GEN_REALS(_op) OSC_GEN_REALS_P(shmem_, _op)
/* RMA and data-movement table; min/max table */
#define OSC_GEN_STD_P(_pre, _p, _op)                                           \
  _Generic(*(_p), OSC_GEN_INTS_P(_pre, _op), OSC_GEN_REALS_P(_pre, _op))
#define OSC_GEN_STD(_p, _op) OSC_GEN_STD_P(shmem_, _p, _op)
/* and/or/xor table: unsigned types, int8..int64 */
#define OSC_GEN_BITWISE(_p, _op)                                               \
  _Generic(*(_p), unsigned char: shmem_uchar_##_op,                            \
           unsigned short: shmem_ushort_##_op, unsigned int: shmem_uint_##_op, \
           unsigned long: shmem_ulong_##_op,                                   \
           unsigned long long: shmem_ulonglong_##_op,                          \
           signed char: shmem_int8_##_op, short: shmem_int16_##_op,            \
           int: shmem_int32_##_op, long: shmem_int64_##_op)
/* sum/prod table: min/max plus the complex types */
#define OSC_GEN_ARITH(_p, _op)                                                 \
  _Generic(*(_p), OSC_GEN_INTS(_op), OSC_GEN_REALS(_op),                       \
           float _Complex: shmem_complexf_##_op,                               \
           double _Complex: shmem_complexd_##_op)
/* the active-set (to_all) tables */
#define OSC_GEN_TO_ALL_INTS(_op)                                               \
  short: shmem_short_##_op, int: shmem_int_##_op, long: shmem_long_##_op,      \
  long long: shmem_longlong_##_op
#define OSC_GEN_TO_ALL_BITWISE(_p, _op) _Generic(*(_p),                        \
      OSC_GEN_TO_ALL_INTS(_op))
#define OSC_GEN_TO_ALL_MINMAX(_p, _op)                                         \
  _Generic(*(_p), OSC_GEN_TO_ALL_INTS(_op), OSC_GEN_REALS(_op))
#define OSC_GEN_TO_ALL_ARITH(_p, _op)                                          \
  _Generic(*(_p), OSC_GEN_TO_ALL_INTS(_op), OSC_GEN_REALS(_op),                \
           float _Complex: shmem_complexf_##_op,                               \
           double _Complex: shmem_complexd_##_op)

/* RMA and atomics take an optional leading context, told apart by the
 * number of arguments: OSC_SEL<k+1> picks the context form (k + 1
 * arguments) or the plain one (k arguments). */
#define OSC_SEL3(_1, _2, _3, N, ...) N
#define OSC_SEL4(_1, _2, _3, _4, N, ...) N
#define OSC_SEL5(_1, _2, _3, _4, _5, N, ...) N
#define OSC_SEL6(_1, _2, _3, _4, _5, _6, N, ...) N
#define OSC_SEL7(_1, _2, _3, _4, _5, _6, _7, N, ...) N
#define OSC_SEL8(_1, _2, _3, _4, _5, _6, _7, _8, N, ...) N
#define OSC_GP_put(dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, put)(dest, source, nelems, pe)
#define OSC_GC_put(ctx, dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, put)(ctx, dest, source, nelems, pe)
#define shmem_put(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_put, OSC_GP_put, 0)(__VA_ARGS__)
#define OSC_GP_get(dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, get)(dest, source, nelems, pe)
#define OSC_GC_get(ctx, dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, get)(ctx, dest, source, nelems, pe)
#define shmem_get(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_get, OSC_GP_get, 0)(__VA_ARGS__)
#define OSC_GP_put_nbi(dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, put_nbi)(dest, source, nelems, pe)
#define OSC_GC_put_nbi(ctx, dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, put_nbi)(ctx, dest, source, nelems, pe)
#define shmem_put_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_put_nbi, OSC_GP_put_nbi, 0)(__VA_ARGS__)
#define OSC_GP_get_nbi(dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, get_nbi)(dest, source, nelems, pe)
#define OSC_GC_get_nbi(ctx, dest, source, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, get_nbi)(ctx, dest, source, nelems, pe)
#define shmem_get_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_get_nbi, OSC_GP_get_nbi, 0)(__VA_ARGS__)
#define OSC_GP_p(dest, value, pe) \
  OSC_GEN_STD_P(shmem_, dest, p)(dest, value, pe)
#define OSC_GC_p(ctx, dest, value, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, p)(ctx, dest, value, pe)
#define shmem_p(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_p, OSC_GP_p, 0)(__VA_ARGS__)
#define OSC_GP_g(source, pe) \
  OSC_GEN_STD_P(shmem_, source, g)(source, pe)
#define OSC_GC_g(ctx, source, pe) \
  OSC_GEN_STD_P(shmem_ctx_, source, g)(ctx, source, pe)
#define shmem_g(...) \
  OSC_SEL3(__VA_ARGS__, OSC_GC_g, OSC_GP_g, 0)(__VA_ARGS__)
#define OSC_GP_iput(dest, source, dst, sst, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, iput)(dest, source, dst, sst, nelems, pe)
#define OSC_GC_iput(ctx, dest, source, dst, sst, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, iput)(ctx, dest, source, dst, sst, nelems, pe)
#define shmem_iput(...) \
  OSC_SEL7(__VA_ARGS__, OSC_GC_iput, OSC_GP_iput, 0)(__VA_ARGS__)
#define OSC_GP_iget(dest, source, dst, sst, nelems, pe) \
  OSC_GEN_STD_P(shmem_, dest, iget)(dest, source, dst, sst, nelems, pe)
#define OSC_GC_iget(ctx, dest, source, dst, sst, nelems, pe) \
  OSC_GEN_STD_P(shmem_ctx_, dest, iget)(ctx, dest, source, dst, sst, nelems, pe)
#define shmem_iget(...) \
  OSC_SEL7(__VA_ARGS__, OSC_GC_iget, OSC_GP_iget, 0)(__VA_ARGS__)
#define OSC_GP_put_signal(dest, source, nelems, sig_addr, signal, sig_op, pe) \
  OSC_GEN_STD_P(shmem_, dest, put_signal)(dest, source, nelems, sig_addr,      \
      signal, sig_op, pe)
#define OSC_GC_put_signal(ctx, dest, source, nelems, sig_addr, signal,         \
      sig_op, pe)                                                              \
  OSC_GEN_STD_P(shmem_ctx_, dest, put_signal)(ctx, dest, source, nelems,       \
      sig_addr, signal, sig_op, pe)
#define shmem_put_signal(...) \
  OSC_SEL8(__VA_ARGS__, OSC_GC_put_signal, OSC_GP_put_signal, 0)(__VA_ARGS__)
#define OSC_GP_put_signal_nbi(dest, source, nelems, sig_addr, signal, sig_op,  \
      pe)                                                                      \
  OSC_GEN_STD_P(shmem_, dest, put_signal_nbi)(dest, source, nelems, sig_addr,  \
      signal, sig_op, pe)
#define OSC_GC_put_signal_nbi(ctx, dest, source, nelems, sig_addr, signal,     \
      sig_op, pe)                                                              \
  OSC_GEN_STD_P(shmem_ctx_, dest, put_signal_nbi)(ctx, dest, source, nelems,   \
      sig_addr, signal, sig_op, pe)
#define shmem_put_signal_nbi(...) \
  OSC_SEL8(__VA_ARGS__, OSC_GC_put_signal_nbi, OSC_GP_put_signal_nbi,          \
      0)(__VA_ARGS__)

#define shmem_and_reduce(team, dest, source, nreduce)                          \
  OSC_GEN_BITWISE(dest, and_reduce)(team, dest, source, nreduce)
#define shmem_or_reduce(team, dest, source, nreduce)                           \
  OSC_GEN_BITWISE(dest, or_reduce)(team, dest, source, nreduce)
#define shmem_xor_reduce(team, dest, source, nreduce)                          \
  OSC_GEN_BITWISE(dest, xor_reduce)(team, dest, source, nreduce)
#define shmem_min_reduce(team, dest, source, nreduce)                          \
  OSC_GEN_STD(dest, min_reduce)(team, dest, source, nreduce)
#define shmem_max_reduce(team, dest, source, nreduce)                          \
  OSC_GEN_STD(dest, max_reduce)(team, dest, source, nreduce)
#define shmem_sum_reduce(team, dest, source, nreduce)                          \
  OSC_GEN_ARITH(dest, sum_reduce)(team, dest, source, nreduce)
#define shmem_prod_reduce(team, dest, source, nreduce)                         \
  OSC_GEN_ARITH(dest, prod_reduce)(team, dest, source, nreduce)

#define shmem_fcollect(team, dest, source, nelems)                             \
  OSC_GEN_STD(dest, fcollect)(team, dest, source, nelems)
#define shmem_collect(team, dest, source, nelems)                              \
  OSC_GEN_STD(dest, collect)(team, dest, source, nelems)
#define shmem_alltoall(team, dest, source, nelems)                             \
  OSC_GEN_STD(dest, alltoall)(team, dest, source, nelems)
#define shmem_alltoalls(team, dest, source, dst, sst, nelems)                  \
  OSC_GEN_STD(dest, alltoalls)(team, dest, source, dst, sst, nelems)
#define shmem_broadcast(team, dest, source, nelems, PE_root)                   \
  OSC_GEN_STD(dest, broadcast)(team, dest, source, nelems, PE_root)

#define shmem_and_to_all(target, source, nreduce, PE_start, logPE_stride,     \
                         PE_size, pWrk, pSync)                                 \
  OSC_GEN_TO_ALL_BITWISE(target, and_to_all)(target, source, nreduce,          \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_or_to_all(target, source, nreduce, PE_start, logPE_stride,      \
                        PE_size, pWrk, pSync)                                  \
  OSC_GEN_TO_ALL_BITWISE(target, or_to_all)(target, source, nreduce,           \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_xor_to_all(target, source, nreduce, PE_start, logPE_stride,     \
                         PE_size, pWrk, pSync)                                 \
  OSC_GEN_TO_ALL_BITWISE(target, xor_to_all)(target, source, nreduce,          \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_min_to_all(target, source, nreduce, PE_start, logPE_stride,     \
                         PE_size, pWrk, pSync)                                 \
  OSC_GEN_TO_ALL_MINMAX(target, min_to_all)(target, source, nreduce,           \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_max_to_all(target, source, nreduce, PE_start, logPE_stride,     \
                         PE_size, pWrk, pSync)                                 \
  OSC_GEN_TO_ALL_MINMAX(target, max_to_all)(target, source, nreduce,           \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_sum_to_all(target, source, nreduce, PE_start, logPE_stride,     \
                         PE_size, pWrk, pSync)                                 \
  OSC_GEN_TO_ALL_ARITH(target, sum_to_all)(target, source, nreduce, PE_start,  \
      logPE_stride, PE_size, pWrk, pSync)
#define shmem_prod_to_all(target, source, nreduce, PE_start, logPE_stride,    \
                          PE_size, pWrk, pSync)                                \
  OSC_GEN_TO_ALL_ARITH(target, prod_to_all)(target, source, nreduce,           \
      PE_start,                                                                \
      logPE_stride, PE_size, pWrk, pSync)
/* point-to-point synchronization: the AMO table; short and unsigned short
 * also for the deprecated single-variable forms */
#define OSC_GEN_SYNC_INTS(_op)                                                 \
  int: shmem_int_##_op, long: shmem_long_##_op,                                \
  long long: shmem_longlong_##_op, unsigned int: shmem_uint_##_op,             \
  unsigned long: shmem_ulong_##_op, unsigned long long: shmem_ulonglong_##_op
#define OSC_GEN_SYNC(_p, _op) _Generic(*(_p), OSC_GEN_SYNC_INTS(_op))
#define OSC_GEN_SYNC1(_p, _op)                                                 \
  _Generic(*(_p), OSC_GEN_SYNC_INTS(_op), short: shmem_short_##_op,            \
           unsigned short: shmem_ushort_##_op)
#define shmem_test(ivar, cmp, cmp_value)                                       \
  OSC_GEN_SYNC1(ivar, test)(ivar, cmp, cmp_value)
#define shmem_wait_until(ivar, cmp, cmp_value)                                 \
  OSC_GEN_SYNC1(ivar, wait_until)(ivar, cmp, cmp_value)
#define shmem_wait(ivar, cmp_value) OSC_GEN_SYNC1(ivar, wait)(ivar, cmp_value)
#define shmem_test_all(ivars, nelems, status, cmp, cmp_value)                  \
  OSC_GEN_SYNC(ivars, test_all)(ivars, nelems, status, cmp, cmp_value)
#define shmem_test_any(ivars, nelems, status, cmp, cmp_value)                  \
  OSC_GEN_SYNC(ivars, test_any)(ivars, nelems, status, cmp, cmp_value)
#define shmem_test_some(ivars, nelems, indices, status, cmp, cmp_value)        \
  OSC_GEN_SYNC(ivars, test_some)(ivars, nelems, indices, status, cmp,          \
                                 cmp_value)
#define shmem_wait_until_all(ivars, nelems, status, cmp, cmp_value)            \
  OSC_GEN_SYNC(ivars, wait_until_all)(ivars, nelems, status, cmp, cmp_value)
#define shmem_wait_until_any(ivars, nelems, status, cmp, cmp_value)            \
  OSC_GEN_SYNC(ivars, wait_until_any)(ivars, nelems, status, cmp, cmp_value)
#define shmem_wait_until_some(ivars, nelems, indices, status, cmp, cmp_value)  \
  OSC_GEN_SYNC(ivars, wait_until_some)(ivars, nelems, indices, status, cmp,    \
                                       cmp_value)
#define shmem_test_all_vector(ivars, nelems, status, cmp, cmp_values)          \
  OSC_GEN_SYNC(ivars, test_all_vector)(ivars, nelems, status, cmp, cmp_values)
#define shmem_test_any_vector(ivars, nelems, status, cmp, cmp_values)          \
  OSC_GEN_SYNC(ivars, test_any_vector)(ivars, nelems, status, cmp, cmp_values)
#define shmem_test_some_vector(ivars, nelems, indices, status, cmp,            \
                               cmp_values)                                     \
  OSC_GEN_SYNC(ivars, test_some_vector)(ivars, nelems, indices, status, cmp,   \
                                        cmp_values)
#define shmem_wait_until_all_vector(ivars, nelems, status, cmp, cmp_values)    \
  OSC_GEN_SYNC(ivars, wait_until_all_vector)(ivars, nelems, status, cmp,       \
                                             cmp_values)
#define shmem_wait_until_any_vector(ivars, nelems, status, cmp, cmp_values)    \
  OSC_GEN_SYNC(ivars, wait_until_any_vector)(ivars, nelems, status, cmp,       \
                                             cmp_values)
#define shmem_wait_until_some_vector(ivars, nelems, indices, status, cmp,      \
                                     cmp_values)                               \
  OSC_GEN_SYNC(ivars, wait_until_some_vector)(ivars, nelems, indices, status,  \
                                              cmp, cmp_values)
/* atomics: extended, standard and bitwise AMO tables */
#define OSC_GEN_AMO_STD_INTS_P(_pre, _op)                                      \
  int: _pre##int_##_op, long: _pre##long_##_op,                                \
  long long: _pre##longlong_##_op, unsigned int: _pre##uint_##_op,             \
  unsigned long: _pre##ulong_##_op,                                            \
  unsigned long long: _pre##ulonglong_##_op
#define OSC_GEN_AMO_EXT_P(_pre, _p, _op)                                       \
  _Generic(*(_p), OSC_GEN_AMO_STD_INTS_P(_pre, _op),                           \
           float: _pre##float_##_op, double: _pre##double_##_op)
#define OSC_GEN_AMO_STD_P(_pre, _p, _op)                                       \
  _Generic(*(_p), OSC_GEN_AMO_STD_INTS_P(_pre, _op))
#define OSC_GEN_AMO_BITWISE_P(_pre, _p, _op)                                   \
  _Generic(*(_p), unsigned int: _pre##uint_##_op,                              \
           unsigned long: _pre##ulong_##_op,                                   \
           unsigned long long: _pre##ulonglong_##_op,                          \
           int: _pre##int32_##_op, long: _pre##int64_##_op)
#define OSC_GEN_AMO_EXT(_p, _op) OSC_GEN_AMO_EXT_P(shmem_, _p, _op)
#define OSC_GEN_AMO_STD(_p, _op) OSC_GEN_AMO_STD_P(shmem_, _p, _op)
#define OSC_GEN_AMO_BITWISE(_p, _op) OSC_GEN_AMO_BITWISE_P(shmem_, _p, _op)
#define OSC_GP_atomic_fetch(source, pe) \
  OSC_GEN_AMO_EXT_P(shmem_, source, atomic_fetch)(source, pe)
#define OSC_GC_atomic_fetch(ctx, source, pe) \
  OSC_GEN_AMO_EXT_P(shmem_ctx_, source, atomic_fetch)(ctx, source, pe)
#define shmem_atomic_fetch(...) \
  OSC_SEL3(__VA_ARGS__, OSC_GC_atomic_fetch, OSC_GP_atomic_fetch,              \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_set(dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_, dest, atomic_set)(dest, value, pe)
#define OSC_GC_atomic_set(ctx, dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_ctx_, dest, atomic_set)(ctx, dest, value, pe)
#define shmem_atomic_set(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_set, OSC_GP_atomic_set, 0)(__VA_ARGS__)
#define OSC_GP_atomic_swap(dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_, dest, atomic_swap)(dest, value, pe)
#define OSC_GC_atomic_swap(ctx, dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_ctx_, dest, atomic_swap)(ctx, dest, value, pe)
#define shmem_atomic_swap(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_swap, OSC_GP_atomic_swap, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_nbi(fetch, source, pe) \
  OSC_GEN_AMO_EXT_P(shmem_, source, atomic_fetch_nbi)(fetch, source, pe)
#define OSC_GC_atomic_fetch_nbi(ctx, fetch, source, pe) \
  OSC_GEN_AMO_EXT_P(shmem_ctx_, source, atomic_fetch_nbi)(ctx, fetch, source,  \
      pe)
#define shmem_atomic_fetch_nbi(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_nbi, OSC_GP_atomic_fetch_nbi,      \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_swap_nbi(fetch, dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_, dest, atomic_swap_nbi)(fetch, dest, value, pe)
#define OSC_GC_atomic_swap_nbi(ctx, fetch, dest, value, pe) \
  OSC_GEN_AMO_EXT_P(shmem_ctx_, dest, atomic_swap_nbi)(ctx, fetch, dest,       \
      value, pe)
#define shmem_atomic_swap_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_swap_nbi, OSC_GP_atomic_swap_nbi,        \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_compare_swap(dest, cond, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_compare_swap)(dest, cond, value, pe)
#define OSC_GC_atomic_compare_swap(ctx, dest, cond, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_compare_swap)(ctx, dest, cond,    \
      value, pe)
#define shmem_atomic_compare_swap(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_compare_swap,                            \
      OSC_GP_atomic_compare_swap, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_inc(dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_fetch_inc)(dest, pe)
#define OSC_GC_atomic_fetch_inc(ctx, dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_fetch_inc)(ctx, dest, pe)
#define shmem_atomic_fetch_inc(...) \
  OSC_SEL3(__VA_ARGS__, OSC_GC_atomic_fetch_inc, OSC_GP_atomic_fetch_inc,      \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_inc(dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_inc)(dest, pe)
#define OSC_GC_atomic_inc(ctx, dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_inc)(ctx, dest, pe)
#define shmem_atomic_inc(...) \
  OSC_SEL3(__VA_ARGS__, OSC_GC_atomic_inc, OSC_GP_atomic_inc, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_add(dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_fetch_add)(dest, value, pe)
#define OSC_GC_atomic_fetch_add(ctx, dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_fetch_add)(ctx, dest, value, pe)
#define shmem_atomic_fetch_add(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_add, OSC_GP_atomic_fetch_add,      \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_add(dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_add)(dest, value, pe)
#define OSC_GC_atomic_add(ctx, dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_add)(ctx, dest, value, pe)
#define shmem_atomic_add(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_add, OSC_GP_atomic_add, 0)(__VA_ARGS__)
#define OSC_GP_atomic_compare_swap_nbi(fetch, dest, cond, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_compare_swap_nbi)(fetch, dest, cond,  \
      value, pe)
#define OSC_GC_atomic_compare_swap_nbi(ctx, fetch, dest, cond, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_compare_swap_nbi)(ctx, fetch,     \
      dest, cond, value, pe)
#define shmem_atomic_compare_swap_nbi(...) \
  OSC_SEL6(__VA_ARGS__, OSC_GC_atomic_compare_swap_nbi,                        \
      OSC_GP_atomic_compare_swap_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_inc_nbi(fetch, dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_fetch_inc_nbi)(fetch, dest, pe)
#define OSC_GC_atomic_fetch_inc_nbi(ctx, fetch, dest, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_fetch_inc_nbi)(ctx, fetch, dest,  \
      pe)
#define shmem_atomic_fetch_inc_nbi(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_inc_nbi,                           \
      OSC_GP_atomic_fetch_inc_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_add_nbi(fetch, dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_, dest, atomic_fetch_add_nbi)(fetch, dest, value, pe)
#define OSC_GC_atomic_fetch_add_nbi(ctx, fetch, dest, value, pe) \
  OSC_GEN_AMO_STD_P(shmem_ctx_, dest, atomic_fetch_add_nbi)(ctx, fetch, dest,  \
      value, pe)
#define shmem_atomic_fetch_add_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_fetch_add_nbi,                           \
      OSC_GP_atomic_fetch_add_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_and(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_and)(dest, value, pe)
#define OSC_GC_atomic_fetch_and(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_and)(ctx, dest, value,  \
      pe)
#define shmem_atomic_fetch_and(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_and, OSC_GP_atomic_fetch_and,      \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_and(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_and)(dest, value, pe)
#define OSC_GC_atomic_and(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_and)(ctx, dest, value, pe)
#define shmem_atomic_and(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_and, OSC_GP_atomic_and, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_and_nbi(fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_and_nbi)(fetch, dest,       \
      value, pe)
#define OSC_GC_atomic_fetch_and_nbi(ctx, fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_and_nbi)(ctx, fetch,    \
      dest, value, pe)
#define shmem_atomic_fetch_and_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_fetch_and_nbi,                           \
      OSC_GP_atomic_fetch_and_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_or(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_or)(dest, value, pe)
#define OSC_GC_atomic_fetch_or(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_or)(ctx, dest, value, pe)
#define shmem_atomic_fetch_or(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_or, OSC_GP_atomic_fetch_or,        \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_or(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_or)(dest, value, pe)
#define OSC_GC_atomic_or(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_or)(ctx, dest, value, pe)
#define shmem_atomic_or(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_or, OSC_GP_atomic_or, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_or_nbi(fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_or_nbi)(fetch, dest,        \
      value, pe)
#define OSC_GC_atomic_fetch_or_nbi(ctx, fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_or_nbi)(ctx, fetch,     \
      dest, value, pe)
#define shmem_atomic_fetch_or_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_fetch_or_nbi,                            \
      OSC_GP_atomic_fetch_or_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_xor(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_xor)(dest, value, pe)
#define OSC_GC_atomic_fetch_xor(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_xor)(ctx, dest, value,  \
      pe)
#define shmem_atomic_fetch_xor(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_fetch_xor, OSC_GP_atomic_fetch_xor,      \
      0)(__VA_ARGS__)
#define OSC_GP_atomic_xor(dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_xor)(dest, value, pe)
#define OSC_GC_atomic_xor(ctx, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_xor)(ctx, dest, value, pe)
#define shmem_atomic_xor(...) \
  OSC_SEL4(__VA_ARGS__, OSC_GC_atomic_xor, OSC_GP_atomic_xor, 0)(__VA_ARGS__)
#define OSC_GP_atomic_fetch_xor_nbi(fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_fetch_xor_nbi)(fetch, dest,       \
      value, pe)
#define OSC_GC_atomic_fetch_xor_nbi(ctx, fetch, dest, value, pe) \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_fetch_xor_nbi)(ctx, fetch,    \
      dest, value, pe)
#define shmem_atomic_fetch_xor_nbi(...) \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_fetch_xor_nbi,                           \
      OSC_GP_atomic_fetch_xor_nbi, 0)(__VA_ARGS__)
#define OSC_GP_atomic_xor_nbi(fetch, dest, value, pe)                          \
  OSC_GEN_AMO_BITWISE_P(shmem_, dest, atomic_xor_nbi)(fetch, dest, value, pe)
#define OSC_GC_atomic_xor_nbi(ctx, fetch, dest, value, pe)                     \
  OSC_GEN_AMO_BITWISE_P(shmem_ctx_, dest, atomic_xor_nbi)(ctx, fetch, dest,    \
                                                           value, pe)
#define shmem_atomic_xor_nbi(...)                                              \
  OSC_SEL5(__VA_ARGS__, OSC_GC_atomic_xor_nbi, OSC_GP_atomic_xor_nbi,          \
           0)(__VA_ARGS__)
#endif /* C11 */

#ifdef __cplusplus
}
#endif
#endif /* OSC_SHMEM_H */
