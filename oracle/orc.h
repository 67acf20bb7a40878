/*
 * orc.h - CPU oracle for the collective hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's default SHCOLL algorithms, run
 * for all PEs of a team inside one process (each PE = one host buffer),
 * step by step as the reference executes them:
 *   orc_reduce_rec_dbl   src/shcoll/src/reduce.c:182-293 (+ :35-43, :730-736)
 *   orc_reduce_linear    src/shcoll/src/reduce.c:55-91
 *   orc_reduce_binomial  src/shcoll/src/reduce.c:102-171
 *   orc_reduce_rabenseifner  src/shcoll/src/reduce.c:306-511, :521-724
 *   orc_fcollect_bruck   src/shcoll/src/fcollect.c:370-403 + util/rotate.c:22-54
 *   orc_alltoall_shift   src/shcoll/src/alltoall.c:75-110, SHIFT_PEER :206
 *   orc_broadcast_linear src/shcoll/src/broadcast.c:58-71, :527-531
 *   orc_broadcast_active_set  broadcast.c:451-470 (+ binomial :150-190)
 *   orc_collect_bruck    src/shcoll/src/collect.c:406-492 + util/scan.c:10-60
 *                        + util/rotate.c:56-64
 *   orc_alltoalls_shift  src/shcoll/src/alltoalls.c:54-90
 *   orc_strided_copy     src/api/rma/putget.h:130-186
 * plus high-precision references for float SUM/PROD:
 *   orc_reduce_accurate  exact (Shewchuk expansion) sums, double-double
 *                        products; result as an unevaluated hi + lo pair
 * and a CPU timing harness of the same reference algorithms with one forked
 * process per PE (orc_cpu_bench_procs), used as the CPU baseline.
 *
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg use
 * this library; the product (libosc) never links or calls it.
 *
 * Element types and ops use the numbering of include/osc.h.
 */
#ifndef ORC_H
#define ORC_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference reduce: src[i], dest[i] are PE i's buffers of n elements. */
int orc_reduce_rec_dbl(int op, int dtype, int npes, size_t n,
                       const void *const *src, void *const *dest);
/* the reference's other reductions (reduce.c:55-91, :102-171, :306-724);
 * binomial in the schedule where children report before their parent
 * polls; rabenseifner with the recursive-doubling (ring = 0) or the ring
 * (ring = 1, rabenseifner2) allgather */
int orc_reduce_linear(int op, int dtype, int npes, size_t n,
                      const void *const *src, void *const *dest);
int orc_reduce_binomial(int op, int dtype, int npes, size_t n,
                        const void *const *src, void *const *dest);
int orc_reduce_rabenseifner(int op, int dtype, int npes, size_t n, int ring,
                            const void *const *src, void *const *dest);

/* High-precision SUM/PROD of the npes inputs per element.  For complex
 * types the outputs hold 2n values (re, im interleaved). */
int orc_reduce_accurate(int op, int dtype, int npes, size_t n,
                        const void *const *src, double *hi, double *lo);

/* Exactly rounded double sum of x[0..m) (Shewchuk 1997 / fsum). */
double orc_fsum(const double *x, size_t m);

int orc_fcollect_bruck(int npes, size_t nbytes, const void *const *src,
                       void *const *dest);
int orc_alltoall_shift(int npes, size_t nbytes, const void *const *src,
                       void *const *dest);
int orc_broadcast_linear(int npes, int root, size_t nbytes,
                         const void *const *src, void *const *dest);
/* util/rotate.c:22-54 byte juggling rotation, exposed for its own test */
void orc_rotate_inplace(char *arr, size_t size, size_t dist);
/* util/rotate.c:56-64 rotation through a copy (used by collect) */
void orc_rotate(char *arr, size_t size, size_t dist);
/* collect with per-PE sizes nbytes[i] (Bruck + prefix sum + rotate) */
int orc_collect_bruck(int npes, const size_t *nbytes, const void *const *src,
                      void *const *dest);
/* strided alltoall, shift exchange; dst/sst in elements of es bytes */
int orc_alltoalls_shift(int npes, size_t es, size_t dst, size_t sst,
                        size_t nelems, const void *const *src,
                        void *const *dest);
/* active-set broadcast: every PE but the root receives the root's source */
int orc_broadcast_active_set(int npes, int root, size_t nbytes,
                             const void *const *src, void *const *dest);
/* strided copy of n elements of es bytes, strides in elements */
void orc_strided_copy(char *dst, const char *src, size_t es, ptrdiff_t tst,
                      ptrdiff_t sst, size_t n);

/* SplitMix64-based input generator shared with the device fill kernel:
 * integer distributions are bit-identical to it. */
uint64_t orc_hash3(uint64_t seed, uint64_t pe, uint64_t idx);

/* CPU baseline: run the reference algorithm with npes PEs spread over
 * nthreads threads (barrier between rounds), `iters` timed repetitions on
 * seeded inputs; *sec_median receives the median seconds per call. */

#ifdef __cplusplus
}
#endif
/* CPU baseline with one forked process per PE over one MAP_SHARED
 * mapping (kind 0: rec_dbl reduce of n elements; 1: Bruck fcollect of n
 * bytes; 2: shift-exchange alltoall of n-byte blocks); median over iters of
 * the slowest PE's time.  src / dest_out optional (P rows each). */
int orc_cpu_bench_procs(int kind, int op, int dtype, int npes, size_t n,
                        int iters, const void *const *src,
                        void *const *dest_out, double *sec_median);

#endif
