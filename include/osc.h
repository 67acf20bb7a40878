/*
 * osc.h - MI355X (gfx950) collective library: the thin launch API.
 *
 * The OpenSHMEM entry points in <shmem.h> are the drop-in boundary.  This
 * header is the lower, stream-aware layer those entry points are built on,
 * plus the pieces a benchmark, a test harness or a serving front end needs:
 * explicit initialisation, float-accuracy mode, asynchronous team
 * collectives on a HIP stream, host<->PE copies, a seeded synthetic-input
 * generator, HIP-event timing, the batched (persistent, graph-captured)
 * executor and the served request loop.
 *
 * Model.  All PEs of the job are emulated inside one GPU: PE p owns the
 * symmetric-heap slice [base + p*slice, base + (p+1)*slice) of one device
 * arena.  A symmetric address is any address inside PE p's slice; its
 * offset from the slice start is the same on every PE.  One host call of a
 * collective acts for every member of the team ("team-call" mode).
 *
 * Every function returns 0 (OSC_OK) on success or a negative OSC_ERR_* code;
 * osc_strerror() names the code and osc_last_error() gives the last message.
 */
#ifndef OSC_H
#define OSC_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
enum {
  OSC_OK = 0,
  OSC_ERR_ARG = -1,         /* bad argument (size, op, type, overlap)      */
  OSC_ERR_NOTINIT = -2,     /* library not initialised                     */
  OSC_ERR_TEAM = -3,        /* invalid team handle                         */
  OSC_ERR_PTR = -4,         /* pointer is not inside the symmetric heap    */
  OSC_ERR_NOMEM = -5,       /* symmetric heap or device memory exhausted   */
  OSC_ERR_HIP = -6,         /* a HIP runtime call failed                   */
  OSC_ERR_UNSUPPORTED = -7, /* type/op not available on gfx950             */
  OSC_ERR_ALGO = -8,        /* unknown collective algorithm name           */
  OSC_ERR_BUDGET = -9,      /* request exceeds the HBM admission budget    */
  OSC_ERR_STATE = -10       /* call not valid in the current state         */
};

/* ---- reduction operators and element types ----------------------------- */
typedef enum {
  OSC_OP_AND = 0,
  OSC_OP_OR,
  OSC_OP_XOR,
  OSC_OP_MIN,
  OSC_OP_MAX,
  OSC_OP_SUM,
  OSC_OP_PROD,
  OSC_OP_COUNT
} osc_op_t;

/* Element types as the kernels see them.  Signedness only matters for
 * MIN/MAX; SUM/PROD on integers wrap modulo 2^bits as two's complement. */
typedef enum {
  OSC_T_U8 = 0,
  OSC_T_U16,
  OSC_T_U32,
  OSC_T_U64,
  OSC_T_I8,
  OSC_T_I16,
  OSC_T_I32,
  OSC_T_I64,
  OSC_T_F32,
  OSC_T_F64,
  OSC_T_C32, /* float _Complex  */
  OSC_T_C64, /* double _Complex */
  OSC_T_COUNT
} osc_dtype_t;

/* Float mode: the reduction algorithm of float and complex types.
 *  ACCURATE : FP32 accumulates in FP64, FP64 in double-double; one final
 *             rounding.  Order fixed (team index 0..P-1), so results are
 *             reproducible run to run and call to call.
 *  REFERENCE: the type's own arithmetic in exactly the combine order of the
 *             recursive-doubling reduction (fold of non-power-of-two
 *             members, then the XOR butterfly): bit-identical to it.
 * MIN/MAX reproduce the butterfly's per-PE selection semantics in both;
 * integer ops are exact in every mode. */
typedef enum {
  OSC_FP_ACCURATE = 0,
  OSC_FP_REFERENCE = 1,
  /* the reference's other reduction algorithms, each in its own combine
   * order and the type's own arithmetic (integer results are the same
   * under every mode):
   *  LINEAR: team-order fold op(acc, x_i), acc = x_0 (reduce.c:55-91);
   *  BINOMIAL: binomial tree, a member folding its children's results in
   *    ascending order as op(child, own) (reduce.c:102-171; the reference
   *    folds them in arrival order, so this is its result when every child
   *    has reported before its parent folds);
   *  RABENSEIFNER: reduce-scatter + allgather (reduce.c:306-724, both
   *    variants): REFERENCE's bits for + and *; float MIN/MAX elements with
   *    a NaN or signed-zero extremum take their block owner's tree. */
  OSC_FP_LINEAR = 2,
  OSC_FP_BINOMIAL = 3,
  OSC_FP_RABENSEIFNER = 4
} osc_fpmode_t;

/* Synthetic-input distributions for osc_fill().  Every value is a pure
 * function of (seed, global PE, element index) through the SplitMix64
 * finaliser, so inputs do not depend on team size or launch shape. */
typedef enum {
  OSC_DIST_BITS = 0,     /* raw 64-bit hash, truncated to the type          */
  OSC_DIST_INT_RANGE,    /* integer uniform in [a, b] (a, b integers)       */
  OSC_DIST_SET_M1_1_2,   /* integer uniform over {-1, 1, 2}                 */
  OSC_DIST_NORMAL,       /* N(a, b^2) via Box-Muller, computed in FP64      */
  OSC_DIST_REAL_RANGE,   /* real uniform in [a, b), computed in FP64        */
  OSC_DIST_A2A_PATTERN,  /* (src_pe<<40 | dst_pe<<20 | idx), a = block elems */
  OSC_DIST_FCOL_PATTERN, /* (uint32)(pe<<24 | idx)                          */
  OSC_DIST_COUNT
} osc_dist_t;

/* ---- lifecycle ---------------------------------------------------------- */
/* Create npes emulated PEs with heap_bytes_per_pe of symmetric heap each on
 * HIP device `device`.  shmem_init() calls this with OSC_NPES (default 8)
 * and SHMEM_SYMMETRIC_SIZE (default 32M). */
int osc_init(int npes, size_t heap_bytes_per_pe, int device);
int osc_finalize(void);
int osc_initialized(void);
int osc_npes(void);
size_t osc_slice_bytes(void); /* distance between consecutive PE slices */
size_t osc_heap_bytes(void);  /* usable symmetric bytes per PE          */
void *osc_heap_base(void);    /* PE 0's slice                            */
void *osc_default_stream(void);
int osc_set_fp_mode(int mode);
int osc_get_fp_mode(void);
/* Reduce kernel shape for per-call launches: -1 (default) picks by
 * accumulation policy and call size; 0..3 force 8x2, 8x1, 2x4 or 1x8 units
 * per thread x slices in flight (OSC_REDUCE_SHAPE sets it at init). */
int osc_set_reduce_shape(int shape);
int osc_get_reduce_shape(void);
/* Parallel branches per wave in a batch's captured graph, 1..8 (default 4,
 * OSC_BATCH_BRANCHES sets it at init); takes effect at the next capture. */
int osc_set_batch_branches(int n);
/* 1 (default): each wave of a batch's graph runs in two phases, the walks
 * of split reductions (and the reductions that stay one kernel) first,
 * then the replicates of the split ones with the data movement, so each
 * phase's HBM traffic runs mostly one way; 0: one phase per wave.  Results
 * are identical either way; takes effect at the next capture.
 * OSC_BATCH_PHASES sets it at init. */
int osc_set_batch_phases(int on);
/* Large reductions (1 MiB or more per member on 32+ members, 3 MiB or more
 * on 8 to 31; elements of 2, 4, 8 or 16 bytes) run as a walk that stores
 * member 0's result and a replicate kernel that copies it to the other
 * members in whole 16-byte units (a partial last unit is stored to every
 * member by the walk), storing each loaded
 * piece to `group` members, the pieces interleaved over `width` workgroups.
 * -1 (default): one member per workgroup, width 32, below 8 MiB per
 * member; from 8 MiB the fastest of (1, 16), (1, 32), (1, 64) and (2, 32)
 * per byte written, timed between events (without waiting) on the first
 * eight calls of each power-of-two size class and fixed by a later one.  1, 2, 4, 8 or 16 fix
 * the group (width 16); 0 = one kernel that stores to every member.
 * Results are identical either way.  OSC_REDUCE_SPLIT sets it at init. */
int osc_set_reduce_split(int group);
/* The group the split uses for npes members and nbytes per member (0: not
 * timed yet; the team size does not enter the choice); *width (may be
 * NULL) receives the interleave width. */
int osc_reduce_split_group(int npes, size_t nbytes, int *width);
/* fcollect, broadcast and alltoall blocks of 16-byte units from 64 KiB per
 * member (alltoall from 24 KiB when P^2 x block exceeds 256 MiB): 1
 * (default) = the replicate's interleaved write order, 8
 * workgroups per super-chunk; 0 = the chunked orders smaller blocks take
 * (destination-major for fcollect/broadcast, pair-major for alltoall).
 * Results are identical either way.  OSC_MOVE_INTERLEAVE sets it at
 * init. */
int osc_set_move_interleave(int on);
const char *osc_strerror(int code);
const char *osc_last_error(void);
int osc_last_error_code(void); /* code of the calling thread's last error */
void osc_clear_error(void);
int osc_set_my_pe(int pe); /* which emulated PE the calling thread is */

/* ---- SPMD calling mode ----------------------------------------------------
 * osc_spmd_run initialises the library (npes PEs of heap_bytes_per_pe on
 * `device`), runs pe_main(pe, arg) on one host thread per PE, waits for all
 * of them and finalises.  While it runs, the OpenSHMEM entry points follow
 * SPMD semantics: every member calls each collective, heap call, barrier
 * and team split with its own copy of the arguments; the last member to
 * arrive launches once for all, and each member's nelems of collect is its
 * own contribution.  Returns 0, the first nonzero pe_main result, or an
 * OSC_ERR_* code.  Symmetric memory is device memory: move data between
 * host variables and symmetric objects with shmem_putmem/getmem. */
int osc_spmd_run(int npes, size_t heap_bytes_per_pe, int device,
                 int (*pe_main)(int pe, void *arg), void *arg);
int osc_spmd_active(void); /* 1 inside osc_spmd_run */

/* ---- symmetric addresses ------------------------------------------------ */
/* Split a symmetric address into (pe, offset).  Returns OSC_ERR_PTR if the
 * address is outside the heap. */
int osc_sym_offset(const void *sym, int *pe, size_t *offset);
void *osc_pe_ptr(const void *sym, int pe); /* same object on PE pe */
/* Blocking host <-> symmetric copies on PE pe, ordered after everything
 * enqueued on the library stream before the call. */
int osc_put_host(void *sym_dest, const void *host_src, size_t bytes, int pe);
int osc_get_host(void *host_dest, const void *sym_src, size_t bytes, int pe);

/* ---- asynchronous team collectives (enqueue on `stream`; NULL = library
 * stream).  dest/src are symmetric addresses; nbytes counts one block. --- */
int osc_reduce(void *team, int op, int dtype, void *dest, const void *src,
               size_t nreduce, void *stream);
int osc_fcollect(void *team, void *dest, const void *src, size_t nbytes,
                 void *stream);
int osc_alltoall(void *team, void *dest, const void *src, size_t nbytes,
                 void *stream);
int osc_broadcast(void *team, void *dest, const void *src, size_t nbytes,
                  int root, void *stream);
/* collect with one size per member: nbytes[i] for team member i (npes
 * entries); every member's dest receives the sources concatenated in team
 * order (reference src/shcoll/src/collect.c:406-492 closed form). */
int osc_collect(void *team, void *dest, const void *src, const size_t *nbytes,
                void *stream);
/* osc_collect without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device or pinned memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-member (offset, size) table goes into a static
 * device array of the library (4096 members, 64 KiB per GPU, held until
 * the process ends) through kernel arguments on `stream`; teams of more
 * than 4096 members are rejected with OSC_ERR_ARG.  Every call on a GPU
 * shares that array and rewrites the entries it uses, so calls that
 * overlap in time on one GPU (other streams, unordered host threads)
 * race; calls ordered on one stream do not.  Results equal osc_collect's. */
int osc_collect_noalloc(void *team, void *dest, const void *src,
                        const size_t *nbytes, void *stream);
/* strided alltoall of nelems elements of elem_bytes per member pair;
 * dst/sst >= 1 in elements (reference src/shcoll/src/alltoalls.c:54-90). */
int osc_alltoalls(void *team, void *dest, const void *src, ptrdiff_t dst,
                  ptrdiff_t sst, size_t elem_bytes, size_t nelems,
                  void *stream);
int osc_stream_sync(void *stream);

/* ---- synthetic inputs --------------------------------------------------- */
int osc_fill(void *team, void *sym, size_t nelems, int dtype, int dist,
             double a, double b, uint64_t seed, void *stream);

/* ---- timing ------------------------------------------------------------- */
int osc_event_create(void **ev);
int osc_event_destroy(void *ev);
int osc_event_record(void *ev, void *stream);
int osc_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- batched execution ---------------------------------------------------
 * Many independent collective calls (uneven sizes, teams, types, ops) run
 * as one captured hipGraph: calls are grouped in waves whose destinations
 * do not overlap; each wave runs one persistent tile-loop kernel per kernel
 * family over all of the wave's calls.  Results equal the same calls
 * issued one by one. */
typedef enum {
  OSC_CALL_REDUCE = 0,
  OSC_CALL_FCOLLECT,
  OSC_CALL_ALLTOALL,
  OSC_CALL_BROADCAST
} osc_call_kind_t;

typedef struct osc_call {
  int32_t kind;       /* OSC_CALL_* */
  int32_t op, dtype;  /* reduce only */
  int32_t root;       /* broadcast: team rank of the root */
  int32_t team_start, team_stride, team_npes; /* global PEs of the team */
  int32_t wave;       /* set by osc_batch_plan (or by the caller) */
  uint64_t count;     /* reduce: elements; others: bytes per block */
  uint64_t src_off;   /* symmetric offsets, 16-byte aligned */
  uint64_t dest_off;
} osc_call;

typedef struct osc_batch osc_batch;

typedef struct osc_batch_info {
  int ncalls, nwaves, nlaunches;
  uint64_t ntiles;
  double algorithmic_bytes; /* sum over calls of the bytes moved */
} osc_batch_info;

/* Place every call's destination inside [dest_base, dest_base+dest_bytes)
 * of each PE of its team span and assign waves: within a wave, no two
 * calls write the same byte of the same PE (skyline packing in call
 * order).  *nwaves receives the wave count. */
int osc_batch_plan(osc_call *calls, int ncalls, uint64_t dest_base,
                   uint64_t dest_bytes, int *nwaves);
/* The same for a world of world_npes PEs; pure host code, no GPU needed. */
int osc_batch_plan_world(osc_call *calls, int ncalls, int world_npes,
                         uint64_t dest_base, uint64_t dest_bytes,
                         int *nwaves);
/* Validate, tile and upload; tiles follow the tuning table. */
int osc_batch_create(const osc_call *calls, int ncalls, osc_batch **out);
/* Run all waves.  The first call captures the launches into a hipGraph
 * on `stream`; later calls replay it. */
int osc_batch_launch(osc_batch *b, void *stream);
/* The same calls, one per-call kernel each, in call order (baseline). */
int osc_batch_launch_serial(osc_batch *b, void *stream);
/* Per-call digest of all members' destination bytes (host array of
 * ncalls); equal digests mean bit-identical results. */
int osc_batch_digests(osc_batch *b, uint64_t *digests, void *stream);
int osc_batch_get_info(const osc_batch *b, osc_batch_info *info);
int osc_batch_destroy(osc_batch *b);
/* Tile parameters: per (class, size bucket); see osc_tune.c. */
int osc_tune_load(const char *path);
int osc_tune_save(const char *path);
int osc_tune_set(int cls, int bucket, int shape, int tile_iters, int grid);
int osc_tune_get(int cls, int bucket, int *shape, int *tile_iters, int *grid);

/* ---- served mode -----------------------------------------------------------
 * A request loop for many concurrent teams: requests arrive over time
 * (arrival_s), are admitted when a slot and destination memory inside the
 * HBM budget are free, and run on the device as soon as admitted.  The
 * arrived requests wait in a queue ordered by cfg->order: arrival order,
 * or size-aware (shortest first with aging: key = algorithmic bytes +
 * age_bytes_per_s * arrival_s, so a request that has waited longer may be
 * proportionally larger and still go first; age_bytes_per_s = 0 is pure
 * shortest-first).  OSC_SERVE_BATCHED is continuous batching: every admitted
 * request that is not yet running joins the next launch group (one tiled
 * persistent launch per kernel family), groups run concurrently on several
 * streams, and memory is released as each group completes.
 * OSC_SERVE_PER_CALL is the baseline: each request is its own per-call
 * kernel launch, in arrival order on one stream. */
typedef enum { OSC_SERVE_BATCHED = 0, OSC_SERVE_PER_CALL = 1 } osc_serve_mode_t;
typedef enum { OSC_SERVE_ORDER_ARRIVAL = 0, OSC_SERVE_ORDER_SIZE = 1 } osc_serve_order_t;

typedef struct osc_request {
  osc_call call;     /* dest_off is assigned by the server */
  double arrival_s;  /* seconds after the start of the run (non-decreasing) */
  double admit_s;    /* filled: memory and slot granted */
  double launch_s;   /* filled: enqueued on a device stream */
  double done_s;     /* filled: device work observed complete */
  uint64_t digest;   /* filled when cfg->digests */
  int32_t group;     /* filled: launch group id */
  int32_t status;    /* filled: 0, or an OSC_ERR_* code if rejected */
} osc_request;

typedef struct osc_serve_cfg {
  int32_t mode;         /* OSC_SERVE_* */
  int32_t max_inflight; /* admitted, unfinished requests */
  int32_t nstreams;     /* device streams for launch groups (batched) */
  int32_t max_group;    /* requests per launch group (batched) */
  uint64_t dest_base;   /* per-PE destination arena [base, base + bytes) */
  uint64_t dest_bytes;
  uint64_t page_bytes;  /* allocation granule inside the arena */
  double budget_bytes;  /* admission budget: in-flight destination bytes
                           summed over every team PE */
  int32_t digests;      /* digest each result before its memory is reused */
  int32_t order;        /* OSC_SERVE_ORDER_*: admission order of the queue */
  double age_bytes_per_s; /* OSC_SERVE_ORDER_SIZE: priority gained per second
                             of earlier arrival, in algorithmic bytes (>= 0) */
} osc_serve_cfg;

typedef struct osc_serve_stats {
  int32_t completed, rejected, groups, max_inflight_seen;
  double wall_s, mean_ms, p50_ms, p90_ms, p99_ms, max_ms;
  double algorithmic_bytes, gbps, peak_inflight_bytes;
} osc_serve_stats;

int osc_serve_run(osc_request *reqs, int nreq, const osc_serve_cfg *cfg,
                  osc_serve_stats *stats);
/* One call alone through the per-call kernels, writing at dest_off; its
 * digest (same function as the served and batched digests). */
int osc_request_alone(const osc_call *call, uint64_t dest_off,
                      uint64_t *digest);

/* ---- configuration helpers (pure host code, usable without a GPU) ------- */
/* Parse a heap size such as "32M", "1G", "4096" (binary K/M/G/T). */
int osc_parse_size(const char *s, size_t *out);
/* Check a collective algorithm spec "algo" or "algo:type" for one of
 * "reduce", "fcollect", "alltoall", "broadcast"; for reductions *fpmode
 * (may be NULL) receives the float mode the name selects. */
int osc_algo_resolve(const char *kind, const char *spec, int *fpmode);

/* ---- raw device copy (the roofline calibration kernel) ------------------ */
int osc_device_copy(void *dst, const void *src, size_t bytes, void *stream);
/* The two halves of a collective's traffic alone, for measuring HBM
 * ceilings: the XOR of every 16-byte unit of [p, p + bytes) stored to the
 * 16 bytes at sink16 (device memory); and a fill of [p, p + bytes) with a
 * 64-bit value.  p 16-byte aligned, bytes a multiple of 16.  The XOR read
 * keeps per-workgroup partial values in a 32 KiB device buffer allocated
 * on its first call on a device: calls that overlap on one device race. */
int osc_device_xor_read(const void *p, size_t bytes, void *sink16,
                        void *stream);
int osc_device_fill64(void *p, size_t bytes, uint64_t value, void *stream);
/* The kernels the calling thread's last reduction launched (the per-call
 * path), named as rocprof names them and separated by ';' (e.g. the slice
 * walk and the replicate of a split call); returns how many. */
int osc_last_reduce_kernels(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* OSC_H */
