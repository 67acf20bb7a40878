/*
 * osc_dev.h - the thin C-ABI between the C host layer and the HIP kernels.
 *
 * Everything below is implemented in the .hip files of this directory and
 * called from the C files.  Plain pointers, sizes and PODs only; no HIP or
 * C++ types cross this line.  Return codes are the OSC_ERR_* values.
 */
#ifndef OSC_DEV_H
#define OSC_DEV_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Team geometry over the symmetric arena: team member i lives in global
 * slice (start + i*stride), i.e. at base + (start + i*stride) * slice. */
typedef struct oscd_geom {
  char *base;     /* device address of global PE 0's slice */
  uint64_t slice; /* bytes between consecutive global slices */
  int32_t start;  /* global PE of team member 0            */
  int32_t stride; /* global PE distance between members    */
  int32_t npes;   /* team size P                           */
  int32_t pad_;
} oscd_geom;

/* ---- runtime ------------------------------------------------------------ */
int oscd_set_device(int dev);
int oscd_get_device_count(int *n);
int oscd_malloc(void **p, size_t bytes);
int oscd_free(void *p);
int oscd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int oscd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int oscd_memcpy_d2d_async(void *dst, const void *src, size_t bytes,
                          void *stream);
int oscd_memcpy_h2d_async(void *dst, const void *src, size_t bytes,
                          void *stream);
int oscd_memset_async(void *dst, int value, size_t bytes, void *stream);
int oscd_stream_create(void **s);
int oscd_stream_destroy(void *s);
int oscd_stream_sync(void *s);
/* Completion wait: a one-wave kernel on stream s stores a ticket into its
 * own slot of coherent pinned host memory with a system-scope release; the
 * host spins on the slot for at most spin_us microseconds, then falls back
 * to hipStreamSynchronize (spin_us = 0: synchronize at once).
 * oscd_wait_init allocates the slots; oscd_wait_fini frees them once no
 * wait is pending. */
int oscd_wait_init(void);
void oscd_wait_fini(void);
int oscd_stream_wait(void *s, unsigned spin_us);
int oscd_stream_query(void *s); /* 1 = idle, 0 = busy, <0 error */
int oscd_event_create(void **e);
int oscd_event_destroy(void *e);
int oscd_event_record(void *e, void *s);
int oscd_event_sync(void *e);
int oscd_stream_wait_event(void *s, void *e); /* s waits for e's record */
int oscd_event_query(void *e); /* 1 = done, 0 = pending, <0 error */
int oscd_event_elapsed_ms(void *a, void *b, float *ms);
int oscd_device_sync(void);
int oscd_mem_info(size_t *free_b, size_t *total_b);
int oscd_host_alloc(void **p, size_t bytes); /* pinned */
int oscd_host_free(void *p);
const char *oscd_last_hip_error(void);

/* ---- kernels ------------------------------------------------------------ */
/* Team reduction of n elements of `dtype` with `op`; every member's
 * dest_off gets the result.  fpmode: 0 accurate, 1 reference order.
 * shape: kernel shape index 0..3 (units per thread x slices in flight:
 * 8x2, 8x1, 2x4, 1x8), or -1 for the per-policy default. */
int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode, int shape,
                uint64_t dest_off, uint64_t src_off, uint64_t n, void *stream);

/* replicate: member j's [off, off + nbytes) = member 0's, for every j
 * (nbytes a multiple of 16), each loaded piece stored to `group` (1, 2, 4,
 * 8 or 16; -1 picks by size) consecutive members, the pieces interleaved
 * over `width` (16, 32 or 64) workgroups; the second kernel of a large
 * reduction.  skip (device bytes, one per 4096-byte piece, or NULL):
 * pieces with a nonzero byte are left alone. */
int oscd_replicate(const oscd_geom *g, uint64_t off, uint64_t nbytes,
                   int group, int width, const void *skip, void *stream);
/* fcollect/broadcast/alltoall blocks of 16-byte units from 64 KiB in the
 * interleaved order (on, the default) or the chunked orders (off) */
void oscd_set_move_interleave(int on);
/* read-only / write-only streams (16-byte aligned, bytes a multiple of
 * 16): XOR of every 16-byte unit into sink16 (cleared first); fill with a
 * 64-bit value in the replicate's write order */
int oscd_xor_read(const void *p, uint64_t bytes, void *sink16, void *stream);
int oscd_fill64(void *p, uint64_t bytes, uint64_t value, void *stream);
/* the kernels the calling thread's last oscd_reduce launched, as rocprof
 * names them, ';'-separated; returns how many */
int oscd_last_reduce_kernels(char *buf, size_t len);
/* replicate group the reduce path uses for a team of npes members and
 * nbytes per member: the fixed group, 8 below 8 MiB, the measured one, or
 * 0 while no call of that size class has measured it yet; *width (may be
 * NULL) receives its interleave width */
int oscd_replicate_group(int npes, uint64_t nbytes, int *width);
/* span flags of split float MIN/MAX reductions: per-stream device buffers
 * of one byte per 4096 bytes of the largest per-member call
 * (bytes_per_member, the heap slice); release frees them all */
void oscd_set_span_flag_capacity(uint64_t bytes_per_member);
void oscd_release_span_flags(void);
/* split calls whose bytes per member are not a multiple of 16 (1, the
 * default) or run them as one kernel (0) */
void oscd_set_split_partial(int on);
/* walk shape of split calls the library picks: 1 (the default) = by bytes
 * per member, 0 = the policy's split shape from 8 MiB, 1x8 below */
void oscd_set_split_walk_rule(int on);
/* which part of a reduction the calling thread's next oscd_reduce calls
 * issue: -1 (default) all; 0 the walk of a split call, or the whole call
 * when it does not split; 1 only the replicate of a split call.  Under 0
 * and 1 float MIN/MAX split only with the span flags set by
 * oscd_set_split_flags (one byte per 256 16-byte units of the call, the
 * same buffer for both phases; NULL: no split). */
void oscd_set_split_phase(int phase);
void oscd_set_split_flags(void *flags);
/* large-reduction split: 0 = one kernel (every workgroup stores to all
 * members), else the walk to member 0 + oscd_replicate with that group
 * size (-1, the default: picked by size) */
void oscd_set_reduce_split(int enabled, int group);

/* fcollect: dest_j[i*nbytes ..) = src_i[0 .. nbytes) for all members i, j */
int oscd_fcollect(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                  uint64_t nbytes, void *stream);

/* alltoall: dest_j[i*nbytes ..) = src_i[j*nbytes ..) for all i, j */
int oscd_alltoall(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                  uint64_t nbytes, void *stream);

/* broadcast: dest_j[0 .. nbytes) = src_root[0 .. nbytes) for all j
 * (j != root when skip_root, as the active-set broadcast does) */
int oscd_broadcast(const oscd_geom *g, int root, uint64_t dest_off,
                   uint64_t src_off, uint64_t nbytes, int skip_root,
                   void *stream);

/* collect, variable sizes: dest_j[off_i .. off_i + n_i) = src_i[0 .. n_i)
 * for all members i, j.  d_tab: device array of 2*npes uint64 (off_i, n_i)
 * in bytes; max_n = max n_i; unit_width W divides every off_i, n_i,
 * dest_off, src_off and the slice. */
int oscd_collect(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                 const uint64_t *d_tab, uint64_t max_n, int unit_width,
                 void *stream);

/* Allocation-free collect: the (offset, size) table lives in a static
 * device array of OSCD_STATIC_TAB_MEMBERS entries per GPU, whose address
 * oscd_scratch_init(dev) records (once per device, with dev current).
 * nbytes: host array of npes per-member byte counts; the table is written
 * by small kernels on `stream` whose arguments carry it, then the collect
 * runs on the same stream.  Calls that overlap on one GPU race on the
 * table. */
#define OSCD_STATIC_TAB_MEMBERS 4096
int oscd_scratch_init(int dev);
int oscd_collect_static(const oscd_geom *g, uint64_t dest_off,
                        uint64_t src_off, const uint64_t *nbytes,
                        uint64_t max_n, int unit_width, int dev, void *stream);

/* strided alltoall of nelems elements of es bytes per member pair:
 * dest_j[((i*nelems + t)*dst)*es ..+es) = src_i[((j*nelems + t)*sst)*es ..)
 * for t < nelems (dst, sst in elements, >= 1). */
int oscd_alltoalls(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                   uint64_t es, uint64_t dst, uint64_t sst, uint64_t nelems,
                   void *stream);

/* strided copy: dst[t*dst_step ..+es) = src[t*src_step ..+es), t < n
 * (steps in bytes) */
int oscd_copy_strided(void *dst, const void *src, uint64_t es,
                      uint64_t dst_step, uint64_t src_step, uint64_t n,
                      void *stream);

/* signal update after the preceding work on `stream`: *sig = value
 * (op 0, set) or *sig += value (op 1, add), as one 64-bit atomic. */
int oscd_signal(uint64_t *sig, uint64_t value, int op, void *stream);
/* Atomic memory operation on a 4- or 8-byte device word (dev_amo.hip).
 * With old != NULL the call waits (bounded spin on a pinned-memory slot,
 * then hipStreamSynchronize) and returns the value found; with old == NULL
 * it is only enqueued.  CSWAP writes value where the word equals cond. */
enum {
  OSCD_AMO_FETCH = 0,
  OSCD_AMO_SET,
  OSCD_AMO_SWAP,
  OSCD_AMO_CSWAP,
  OSCD_AMO_ADD,
  OSCD_AMO_AND,
  OSCD_AMO_OR,
  OSCD_AMO_XOR
};
int oscd_amo_init(void);
void oscd_amo_fini(void);
int oscd_amo(void *addr, int width, int op, uint64_t value, uint64_t cond,
             uint64_t *old, void *stream, unsigned spin_us);

/* ---- batched execution ------------------------------------------------- */
/* One collective call of a batch as the device sees it (64 bytes). */
typedef struct oscd_bcall {
  uint64_t src_off, dest_off;
  uint64_t count; /* reduce: elements; fcollect/alltoall/broadcast: bytes */
  int32_t start, stride, npes, lg;
  int32_t kind;   /* OSC_CALL_* */
  int32_t op, dtype;
  int32_t aux;    /* reduce: float mode; broadcast: root (team rank) */
  uint64_t pad_;
} oscd_bcall;

/* One tile: part of one call's work for one workgroup iteration.
 *  reduce    : units [begin, end) of VEC elements (16 bytes)
 *  fcollect  : source member `aux`, units [begin, end) of its block
 *  broadcast : units [begin, end) of the root's block
 *  alltoall  : if the block has >= 256 units: pair `aux`, units
 *              [begin, end); else pairs [begin, end), all units. */
typedef struct oscd_tile {
  uint32_t call;
  uint32_t aux;
  uint64_t begin, end;
} oscd_tile;

/* Persistent tile loop for one reduce family (dtype, op, mode, shape);
 * grid workgroups stride over ntiles tiles; lds_bytes covers the largest
 * team's special-element snapshot. */
int oscd_batch_reduce(int dtype, int op, int fpmode, int shape, int maxlg,
                      const oscd_bcall *calls, const oscd_tile *tiles,
                      uint32_t ntiles, uint32_t grid, size_t lds_bytes,
                      char *base, uint64_t slice, void *stream);
/* Persistent tile loop for data movement of unit width W (1..16). */
int oscd_batch_move(int unit_width, const oscd_bcall *calls,
                    const oscd_tile *tiles, uint32_t ntiles, uint32_t grid,
                    char *base, uint64_t slice, void *stream);
/* Per-call digest of every member's destination bytes (XOR of hashed
 * words, order-independent); digests[] must be zeroed first. */
int oscd_batch_digest(const oscd_bcall *calls, uint32_t ncalls,
                      uint64_t *digests, char *base, uint64_t slice,
                      void *stream);
int oscd_fill_u64(uint64_t *p, uint64_t value, uint64_t count, void *stream);
/* reduce kernel shape/unit geometry, for the host tile planner */
int oscd_reduce_span_units(int dtype, int op, int fpmode, int shape);
int oscd_reduce_vec(int dtype);
int oscd_cu_count(int *n);

/* graph capture of everything enqueued on `stream` between begin and end */
int oscd_graph_begin(void *stream);
int oscd_graph_end(void *stream, void **exec);
int oscd_graph_launch(void *exec, void *stream);
int oscd_graph_destroy(void *exec);

/* plain device copy (put/get between slices, roofline calibration) */
int oscd_copy(void *dst, const void *src, uint64_t bytes, void *stream);

/* seeded synthetic inputs for every member of g at sym offset `off` */
int oscd_fill(const oscd_geom *g, uint64_t off, uint64_t nelems, int dtype,
              int dist, double a, double b, uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OSC_DEV_H */
