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
int oscd_memset_async(void *dst, int value, size_t bytes, void *stream);
int oscd_stream_create(void **s);
int oscd_stream_destroy(void *s);
int oscd_stream_sync(void *s);
int oscd_stream_query(void *s); /* 1 = idle, 0 = busy, <0 error */
int oscd_event_create(void **e);
int oscd_event_destroy(void *e);
int oscd_event_record(void *e, void *s);
int oscd_event_sync(void *e);
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
 * 8x2, 4x4, 2x4, 1x8), or -1 for the per-policy default. */
int oscd_reduce(const oscd_geom *g, int op, int dtype, int fpmode, int shape,
                uint64_t dest_off, uint64_t src_off, uint64_t n, void *stream);

/* fcollect: dest_j[i*nbytes ..) = src_i[0 .. nbytes) for all members i, j */
int oscd_fcollect(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                  uint64_t nbytes, void *stream);

/* alltoall: dest_j[i*nbytes ..) = src_i[j*nbytes ..) for all i, j */
int oscd_alltoall(const oscd_geom *g, uint64_t dest_off, uint64_t src_off,
                  uint64_t nbytes, void *stream);

/* broadcast: dest_j[0 .. nbytes) = src_root[0 .. nbytes) for all j */
int oscd_broadcast(const oscd_geom *g, int root, uint64_t dest_off,
                   uint64_t src_off, uint64_t nbytes, void *stream);

/* plain device copy (put/get between slices, roofline calibration) */
int oscd_copy(void *dst, const void *src, uint64_t bytes, void *stream);

/* seeded synthetic inputs for every member of g at sym offset `off` */
int oscd_fill(const oscd_geom *g, uint64_t off, uint64_t nelems, int dtype,
              int dist, double a, double b, uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OSC_DEV_H */
