/*
 * osc_core.c - library state, initialisation, errors, symmetric addresses,
 * host<->PE copies, synthetic inputs and timing.
 *
 * Initialisation replaces the reference's shmemc_init / heap registration
 * sequence (src/shmemc/init.c:38-72, src/shmemc/heaps.c:33-52,
 * src/shmemc/ucx/init.c:174-218): there is no launcher, no PMIx and no UCX,
 * only one device arena holding every emulated PE's symmetric slice.
 */
#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_internal.h"

osc_state osc_g = {0};
static __thread char g_errmsg[512];
static __thread int g_errcode = 0;
static __thread int g_my_pe = 0;

int osc_set_error(int code, const char *fmt, ...) {
  va_list ap;
  g_errcode = code;
  va_start(ap, fmt);
  vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
  va_end(ap);
  if (osc_g.log_calls) fprintf(stderr, "[osc] error %d: %s\n", code, g_errmsg);
  return code;
}

int osc_hip_error(int code, const char *what) {
  return osc_set_error(code, "%s: %s", what, oscd_last_hip_error());
}

const char *osc_last_error(void) { return g_errmsg; }
int osc_last_error_code(void) { return g_errcode; }
void osc_clear_error(void) {
  g_errcode = OSC_OK;
  g_errmsg[0] = '\0';
}

const char *osc_strerror(int code) {
  switch (code) {
    case OSC_OK: return "success";
    case OSC_ERR_ARG: return "invalid argument";
    case OSC_ERR_NOTINIT: return "library not initialised";
    case OSC_ERR_TEAM: return "invalid team";
    case OSC_ERR_PTR: return "address not in the symmetric heap";
    case OSC_ERR_NOMEM: return "out of memory";
    case OSC_ERR_HIP: return "HIP runtime error";
    case OSC_ERR_UNSUPPORTED: return "type/op not supported on gfx950";
    case OSC_ERR_ALGO: return "unknown collective algorithm";
    case OSC_ERR_BUDGET: return "request exceeds the HBM admission budget";
    case OSC_ERR_STATE: return "invalid in the current state";
  }
  return "unknown error";
}

size_t osc_dtype_size(int dtype) {
  static const size_t sz[OSC_T_COUNT] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8, 8, 16};
  if (dtype < 0 || dtype >= OSC_T_COUNT) return 0;
  return sz[dtype];
}

/* Slice size: the usable heap rounded up to 2 MiB (so slices start on a
 * large-page boundary), plus OSC_SLICE_PAD bytes (default 0) to shift
 * equal offsets of different PEs apart if channel conflicts ever show. */
static size_t slice_for(size_t heap_bytes) {
  const size_t two_mib = (size_t)2 << 20;
  size_t s = (heap_bytes + two_mib - 1) / two_mib * two_mib;
  const char *pad = getenv("OSC_SLICE_PAD");
  if (pad != NULL) s += (size_t)strtoull(pad, NULL, 0) / 256 * 256;
  return s;
}

int osc_init(int npes, size_t heap_bytes_per_pe, int device) {
  int rc, ndev = 0;
  void *base = NULL;
  const char *lg = getenv("OSC_LOG");
  if (osc_g.initialized) return osc_set_error(OSC_ERR_STATE, "already initialised");
  const char *shp = getenv("OSC_REDUCE_SHAPE");
  const char *spin = getenv("OSC_WAIT_SPIN_US");
  osc_g.log_calls = (lg != NULL && lg[0] != '\0' && lg[0] != '0');
  osc_g.reduce_shape = (shp != NULL && shp[0] != '\0') ? atoi(shp) : -1;
  {
    const char *sp = getenv("OSC_REDUCE_SPLIT");
    if (sp == NULL || sp[0] == '\0' || osc_set_reduce_split(atoi(sp)) != OSC_OK)
      osc_set_reduce_split(-1);
  }
  {
    const char *gi = getenv("OSC_MOVE_INTERLEAVE");
    oscd_set_move_interleave(gi == NULL || gi[0] == '\0' || atoi(gi) != 0);
  }
  {
    const char *br = getenv("OSC_BATCH_BRANCHES");
    osc_g.batch_branches = (br != NULL && br[0] != '\0') ? atoi(br) : 4;
    if (osc_g.batch_branches < 1) osc_g.batch_branches = 1;
    if (osc_g.batch_branches > 8) osc_g.batch_branches = 8;
    {
      const char *ph = getenv("OSC_BATCH_PHASES");
      osc_g.batch_phases = (ph != NULL && ph[0] != '\0') ? atoi(ph) != 0 : 1;
    }
  }
  osc_g.wait_spin_us =
      (spin != NULL && spin[0] != '\0') ? (unsigned)strtoul(spin, NULL, 10) : 500u;
  if (npes < 1 || npes > OSC_MAX_NPES)
    return osc_set_error(OSC_ERR_ARG, "npes %d outside [1, %d]", npes,
                         OSC_MAX_NPES);
  if (heap_bytes_per_pe == 0)
    return osc_set_error(OSC_ERR_ARG, "zero heap size");
  heap_bytes_per_pe = (heap_bytes_per_pe + 255) / 256 * 256;
  rc = osc_registry_init_from_env();
  if (rc != OSC_OK) return rc;
  osc_tune_init_from_env();
  rc = oscd_get_device_count(&ndev);
  if (rc != OSC_OK) return osc_hip_error(rc, "hipGetDeviceCount");
  if (device < 0 || device >= ndev)
    return osc_set_error(OSC_ERR_ARG, "device %d not present (%d devices)",
                         device, ndev);
  rc = oscd_set_device(device);
  if (rc != OSC_OK) return osc_hip_error(rc, "hipSetDevice");
  /* the allocation-free collect's table: this device's address of it */
  rc = oscd_scratch_init(device);
  if (rc != OSC_OK) return osc_hip_error(rc, "hipGetSymbolAddress(scratch)");
  osc_g.slice = slice_for(heap_bytes_per_pe);
  oscd_set_span_flag_capacity(osc_g.slice);
  rc = oscd_malloc(&base, (size_t)npes * osc_g.slice);
  if (rc != OSC_OK)
    return osc_set_error(rc, "symmetric arena of %d x %zu bytes: %s", npes,
                         osc_g.slice, oscd_last_hip_error());
  rc = oscd_stream_create(&osc_g.stream);
  if (rc != OSC_OK) {
    oscd_free(base);
    return osc_hip_error(rc, "hipStreamCreate");
  }
  rc = oscd_wait_init();
  if (rc == OSC_OK) rc = oscd_amo_init();
  if (rc != OSC_OK) {
    oscd_wait_fini();
    oscd_stream_destroy(osc_g.stream);
    oscd_free(base);
    return osc_hip_error(rc, "completion slots");
  }
  rc = osc_heap_init(heap_bytes_per_pe);
  if (rc != OSC_OK) {
    oscd_amo_fini();
    oscd_wait_fini();
    oscd_stream_destroy(osc_g.stream);
    oscd_free(base);
    return osc_set_error(rc, "heap bookkeeping");
  }
  pthread_mutex_init(&osc_g.lock, NULL);
  osc_g.device = device;
  osc_g.npes = npes;
  osc_g.heap_bytes = heap_bytes_per_pe;
  osc_g.base = (char *)base;
  osc_team_world.magic = OSC_TEAM_MAGIC;
  osc_team_world.start = 0;
  osc_team_world.stride = 1;
  osc_team_world.npes = npes;
  osc_team_world.predefined = 1;
  osc_team_shared = osc_team_world;
  osc_g.initialized = 1;
  if (osc_g.log_calls)
    fprintf(stderr, "[osc] init: %d PEs x %zu B heap (slice %zu B) on device %d\n",
            npes, heap_bytes_per_pe, osc_g.slice, device);
  return OSC_OK;
}

int osc_finalize(void) {
  if (!osc_g.initialized) return OSC_OK;
  osc_ctx_release_all();
  oscd_stream_sync(osc_g.stream);
  oscd_release_span_flags();
  oscd_stream_destroy(osc_g.stream);
  oscd_wait_fini();
  oscd_amo_fini();
  if (osc_g.aux_dev) oscd_free(osc_g.aux_dev);
  if (osc_g.tab_event) oscd_event_destroy(osc_g.tab_event);
  if (osc_g.log_ev0) oscd_event_destroy(osc_g.log_ev0);
  if (osc_g.log_ev1) oscd_event_destroy(osc_g.log_ev1);
  if (osc_g.tab_dev) oscd_free(osc_g.tab_dev);
  if (osc_g.tab_host) oscd_host_free(osc_g.tab_host);
  oscd_free(osc_g.base);
  osc_heap_fini();
  pthread_mutex_destroy(&osc_g.lock);
  osc_team_world.magic = 0;
  osc_team_shared.magic = 0;
  memset(&osc_g, 0, sizeof(osc_g));
  return OSC_OK;
}

int osc_initialized(void) { return osc_g.initialized; }

/* -1: per-call default (by policy and size); 0..3: force one shape */
int osc_set_reduce_shape(int shape) {
  if (shape < -1 || shape > 3) return osc_set_error(OSC_ERR_ARG, "shape %d", shape);
  osc_g.reduce_shape = shape;
  return OSC_OK;
}
int osc_get_reduce_shape(void) { return osc_g.reduce_shape; }

int osc_set_batch_branches(int n) {
  if (n < 1 || n > 8) return osc_set_error(OSC_ERR_ARG, "batch branches %d (1..8)", n);
  osc_g.batch_branches = n;
  return OSC_OK;
}
int osc_set_batch_phases(int on) {
  if (on != 0 && on != 1)
    return osc_set_error(OSC_ERR_ARG, "batch phases %d (0 or 1)", on);
  osc_g.batch_phases = on;
  return OSC_OK;
}
int osc_set_reduce_split(int group) {
  if (group != -1 && group != 0 && group != 1 && group != 2 && group != 4 &&
      group != 8 && group != 16)
    return osc_set_error(OSC_ERR_ARG,
                         "replicate group %d (-1, 0, 1, 2, 4, 8, 16)", group);
  oscd_set_reduce_split(group != 0, group);
  return OSC_OK;
}
int osc_set_move_interleave(int on) {
  if (on != 0 && on != 1)
    return osc_set_error(OSC_ERR_ARG, "move interleave %d (0 or 1)", on);
  oscd_set_move_interleave(on);
  return OSC_OK;
}
int osc_reduce_split_group(int npes, size_t nbytes, int *width) {
  return oscd_replicate_group(npes, (uint64_t)nbytes, width);
}
int osc_npes(void) { return osc_g.npes; }
size_t osc_slice_bytes(void) { return osc_g.slice; }
size_t osc_heap_bytes(void) { return osc_g.heap_bytes; }
void *osc_heap_base(void) { return osc_g.base; }
void *osc_default_stream(void) { return osc_g.stream; }

int osc_set_my_pe(int pe) {
  if (osc_g.initialized && (pe < 0 || pe >= osc_g.npes))
    return osc_set_error(OSC_ERR_ARG, "PE %d out of range", pe);
  g_my_pe = pe;
  return OSC_OK;
}

int osc_my_pe_get(void) { return g_my_pe; }

int osc_sym_offset(const void *sym, int *pe, size_t *offset) {
  const char *p = (const char *)sym;
  size_t d, off;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (p < osc_g.base || p >= osc_g.base + (size_t)osc_g.npes * osc_g.slice)
    return osc_set_error(OSC_ERR_PTR, "%p is not a symmetric address", sym);
  d = (size_t)(p - osc_g.base);
  off = d % osc_g.slice;
  if (off >= osc_g.heap_bytes)
    return osc_set_error(OSC_ERR_PTR, "%p is in slice padding", sym);
  if (pe != NULL) *pe = (int)(d / osc_g.slice);
  if (offset != NULL) *offset = off;
  return OSC_OK;
}

/* [p, p+bytes) must lie inside one PE's usable heap; returns its offset */
int osc_sym_range(const void *p, size_t bytes, size_t *off) {
  size_t o;
  int rc = osc_sym_offset(p, NULL, &o);
  if (rc != OSC_OK) return rc;
  if (bytes > osc_g.heap_bytes - o)
    return osc_set_error(OSC_ERR_PTR,
                         "range of %zu bytes at offset %zu runs past the %zu-byte heap",
                         bytes, o, osc_g.heap_bytes);
  *off = o;
  return OSC_OK;
}

void *osc_pe_ptr(const void *sym, int pe) {
  size_t off;
  if (osc_sym_offset(sym, NULL, &off) != OSC_OK) return NULL;
  if (pe < 0 || pe >= osc_g.npes) return NULL;
  return osc_g.base + (size_t)pe * osc_g.slice + off;
}

int osc_put_host(void *sym_dest, const void *host_src, size_t bytes, int pe) {
  size_t off;
  int rc;
  if (pe < 0 || pe >= osc_g.npes) return osc_set_error(OSC_ERR_ARG, "bad PE %d", pe);
  rc = osc_sym_range(sym_dest, bytes, &off);
  if (rc != OSC_OK) return rc;
  /* the library stream is non-blocking: the copy below (null stream) must
   * not overtake work enqueued on it before this call */
  rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  if (rc != OSC_OK) return osc_hip_error(rc, "stream wait");
  rc = oscd_memcpy_h2d(osc_g.base + (size_t)pe * osc_g.slice + off, host_src, bytes);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "hipMemcpy H2D");
}

int osc_get_host(void *host_dest, const void *sym_src, size_t bytes, int pe) {
  size_t off;
  int rc;
  if (pe < 0 || pe >= osc_g.npes) return osc_set_error(OSC_ERR_ARG, "bad PE %d", pe);
  rc = osc_sym_range(sym_src, bytes, &off);
  if (rc != OSC_OK) return rc;
  rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  if (rc != OSC_OK) return osc_hip_error(rc, "stream wait");
  rc = oscd_memcpy_d2h(host_dest, osc_g.base + (size_t)pe * osc_g.slice + off, bytes);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "hipMemcpy D2H");
}

int osc_fill(void *team, void *sym, size_t nelems, int dtype, int dist,
             double a, double b, uint64_t seed, void *stream) {
  const osc_team *t = osc_team_get(team);
  oscd_geom g;
  size_t off;
  int rc;
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (t == NULL) return osc_set_error(OSC_ERR_TEAM, "invalid team");
  if (osc_dtype_size(dtype) == 0) return osc_set_error(OSC_ERR_ARG, "bad dtype %d", dtype);
  rc = osc_sym_range(sym, nelems * osc_dtype_size(dtype), &off);
  if (rc != OSC_OK) return rc;
  osc_team_geom(t, &g);
  rc = oscd_fill(&g, off, nelems, dtype, dist, a, b, seed,
                 stream ? stream : osc_g.stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "fill launch");
}

int osc_wait(void) {
  const int rc = oscd_stream_wait(osc_g.stream, osc_g.wait_spin_us);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "stream wait");
}

int osc_stream_sync(void *stream) {
  int rc = oscd_stream_sync(stream ? stream : osc_g.stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "hipStreamSynchronize");
}

int osc_event_create(void **ev) { return oscd_event_create(ev); }
int osc_event_destroy(void *ev) { return oscd_event_destroy(ev); }
int osc_event_record(void *ev, void *stream) {
  return oscd_event_record(ev, stream ? stream : osc_g.stream);
}
int osc_event_elapsed_ms(void *a, void *b, float *ms) {
  return oscd_event_elapsed_ms(a, b, ms);
}

int osc_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
  int rc = oscd_copy(dst, src, bytes, stream ? stream : osc_g.stream);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "copy launch");
}

int osc_device_xor_read(const void *p, size_t bytes, void *sink16,
                        void *stream) {
  int rc = oscd_xor_read(p, bytes, sink16, stream ? stream : osc_g.stream);
  if (rc == OSC_ERR_ARG)
    return osc_set_error(rc, "xor read: %p (%zu bytes) needs 16-byte alignment",
                         p, bytes);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "xor read launch");
}

int osc_device_fill64(void *p, size_t bytes, uint64_t value, void *stream) {
  int rc = oscd_fill64(p, bytes, value, stream ? stream : osc_g.stream);
  if (rc == OSC_ERR_ARG)
    return osc_set_error(rc, "fill: %p (%zu bytes) needs 16-byte alignment",
                         p, bytes);
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "fill launch");
}

int osc_last_reduce_kernels(char *buf, size_t len) {
  return oscd_last_reduce_kernels(buf, len);
}

/* "32M", "1G", "4096", "512k": binary multiples, as the reference's
 * SHMEM_SYMMETRIC_SIZE parser (src/shmemu/unitparse.c:72). */
int osc_parse_size(const char *s, size_t *out) {
  char *end = NULL;
  double v;
  size_t mul = 1;
  if (s == NULL || *s == '\0') return OSC_ERR_ARG;
  v = strtod(s, &end);
  if (end == s || v < 0) return OSC_ERR_ARG;
  while (*end == ' ') ++end;
  switch (toupper((unsigned char)*end)) {
    case '\0': break;
    case 'K': mul = (size_t)1 << 10; break;
    case 'M': mul = (size_t)1 << 20; break;
    case 'G': mul = (size_t)1 << 30; break;
    case 'T': mul = (size_t)1 << 40; break;
    default: return OSC_ERR_ARG;
  }
  if (*end != '\0') {
    ++end;
    if (toupper((unsigned char)*end) == 'B') ++end;
    if (*end != '\0') return OSC_ERR_ARG;
  }
  *out = (size_t)(v * (double)mul);
  return OSC_OK;
}
