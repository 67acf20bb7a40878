/*
 * osc_internal.h - state shared by the C host files of libosc.
 */
#ifndef OSC_INTERNAL_H
#define OSC_INTERNAL_H 1

#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

#include "../../../include/osc.h"
#include "../osc_dev.h"

#define OSC_TEAM_MAGIC 0x05C7EA11
#define OSC_MAX_NPES 8191

typedef struct osc_team {
  int magic;
  int start;  /* global PE of member 0 */
  int stride; /* global PE distance between members (>= 1) */
  int npes;
  int predefined;
  int num_contexts;
  struct osc_team *parent;
} osc_team;

/* one block of the symmetric offset space */
typedef struct osc_block {
  size_t off;
  size_t size;
  int used;
  struct osc_block *next;
} osc_block;

typedef struct osc_state {
  int initialized;
  int device;
  int npes;
  size_t heap_bytes; /* usable symmetric bytes per PE */
  size_t slice;      /* distance between PE slices (>= heap_bytes) */
  char *base;        /* device arena: npes * slice bytes */
  void *stream;      /* library stream for blocking calls */
  osc_block *blocks; /* sorted by offset, covers [0, heap_bytes) */
  pthread_mutex_t lock;
  int log_calls;
  int spmd; /* 1 while osc_spmd_run drives one host thread per PE */
  /* resolved algorithm choice: float mode per (op, dtype) */
  unsigned char fpmode[OSC_OP_COUNT][OSC_T_COUNT];
  unsigned char fpmode_to_all[OSC_OP_COUNT][OSC_T_COUNT]; /* active-set forms */
  int fp_default;
  int reduce_shape; /* -1 = per-policy default; OSC_REDUCE_SHAPE overrides */
  int batch_branches; /* parallel branches per wave of a batch's graph */
  int batch_phases;   /* waves in two phases: split walks, then replicates
                         and data movement (osc_set_batch_phases) */
  unsigned wait_spin_us; /* blocking calls: bounded spin on a completion
                            slot before hipStreamSynchronize (0: never) */
  void *aux_dev;    /* 256 device bytes for single-call descriptors */
  /* per-member (offset, size) table of the variable-size collect:
   * pinned staging, its device copy, and the event of its last use */
  uint64_t *tab_host;
  uint64_t *tab_dev;
  void *tab_event;
  int tab_used;
  /* event pair of the per-call records (OSC_LOG) */
  void *log_ev0, *log_ev1;
} osc_state;

extern osc_state osc_g;
extern osc_team osc_team_world;
extern osc_team osc_team_shared;

/* completion of everything enqueued on the library stream so far */
int osc_wait(void);
/* destroys the contexts the program did not (shmem_api.c) */
void osc_ctx_release_all(void);
/* a context's stream and the global PE of its PE number pe; -1 with the
 * error recorded if the context or the PE is invalid (shmem_api.c) */
int osc_ctx_target(void *ctx, int pe, void **stream);

/* errors */
int osc_set_error(int code, const char *fmt, ...);
int osc_hip_error(int code, const char *what);

/* heap (osc_heap.c) */
int osc_heap_init(size_t heap_bytes);
void osc_heap_fini(void);
int osc_heap_alloc(size_t size, size_t align, size_t *off);
int osc_heap_free(size_t off);
int osc_heap_block_size(size_t off, size_t *size);
size_t osc_heap_used(void);

/* teams (osc_team.c) */
const osc_team *osc_team_get(void *team);
void osc_team_geom(const osc_team *t, oscd_geom *g);
int osc_my_pe_get(void);

/* algorithm registry (osc_registry.c) */
int osc_registry_init_from_env(void);
int osc_registry_resolve(int kind, int op, const char *spec, int *fpmode);
size_t osc_dtype_size(int dtype);

enum { OSC_KIND_REDUCE = 0, OSC_KIND_FCOLLECT, OSC_KIND_ALLTOALL,
       OSC_KIND_BROADCAST };

/* tuning table (osc_tune.c) */
void osc_tune_init_from_env(void);
int osc_tune_bucket(uint64_t bytes_per_pe);

/* symmetric-argument validation (osc_coll.c) */
int osc_sym_range(const void *p, size_t bytes, size_t *off);

/* symmetric-heap operations on offsets, for every PE at once
 * (shmem_api.c); zero: clear the new block on every PE */
int osc_heap_call_alloc(size_t size, size_t align, int zero, size_t *off);
int osc_heap_call_free(size_t off);
int osc_heap_call_realloc(size_t off, size_t size, size_t *new_off);

/* SPMD calling mode (osc_spmd.c): every member thread calls; the last to
 * arrive runs the team-call path once for all and the others wait for it.
 * Kinds for spmd_move: */
enum { OSC_SPMD_FCOLLECT = 1, OSC_SPMD_ALLTOALL, OSC_SPMD_BROADCAST };
/* fpmode < 0: the team reductions' table */
int osc_spmd_reduce(void *team, int op, int dtype, void *dest,
                    const void *src, size_t n, int fpmode);
/* osc_reduce with an explicit float mode (osc_coll.c) */
int osc_reduce_mode(void *team, int op, int dtype, void *dest, const void *src,
                    size_t nreduce, int fpmode, void *stream);
int osc_spmd_move(int kind, void *team, void *dest, const void *src,
                  size_t nbytes, int root, int skip_root);
int osc_spmd_collect(void *team, void *dest, const void *src,
                     size_t my_nbytes);
int osc_spmd_alltoalls(void *team, void *dest, const void *src, ptrdiff_t dst,
                       ptrdiff_t sst, size_t elem_bytes, size_t nelems);
int osc_spmd_sync(void *team);
void *osc_spmd_heap(int kind, void *ptr, size_t a, size_t b);
enum { OSC_SPMD_MALLOC = 1, OSC_SPMD_CALLOC, OSC_SPMD_ALIGN, OSC_SPMD_REALLOC,
       OSC_SPMD_FREE };
int osc_spmd_split_strided(void *parent, int start, int stride, int size,
                           const void *config, long config_mask,
                           void **new_team);
int osc_spmd_split_2d(void *parent, int xrange, const void *xconfig,
                      long xmask, void **xteam, const void *yconfig,
                      long ymask, void **yteam);
void osc_spmd_team_destroy(void *team);
/* team-call bodies the SPMD leader runs (osc_team.c) */
int osc_team_split_strided_tc(void *parent, int start, int stride, int size,
                              const void *config, long config_mask,
                              void **new_team);
int osc_team_rank(const osc_team *t, int pe);
int osc_team_split_2d_all(const osc_team *p, int xrange, const void *xc,
                          long xm, const void *yc, long ym, osc_team **rows,
                          osc_team **cols);

/* Active set (PE_start, 2^logPE_stride, PE_size) of the pre-1.5 API as a
 * transient team (not registered; valid while `t` lives). */
int osc_active_set(int PE_start, int logPE_stride, int PE_size, osc_team *t);
/* broadcast with the root's dest left alone when skip_root (active-set
 * form) */
int osc_broadcast_impl(void *team, void *dest, const void *src, size_t nbytes,
                       int root, int skip_root, void *stream);

#endif
