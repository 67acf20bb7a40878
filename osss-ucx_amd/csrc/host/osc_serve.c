/*
 * osc_serve.c - the served request loop (continuous batching).
 *
 * One host thread runs the loop; the device does the work on several
 * streams.  Each iteration:
 *   1. retire: launch groups whose event has completed release their
 *      destination pages and budget (after digesting results if asked);
 *   2. arrive: requests whose arrival time has passed are validated
 *      (those that can never run are rejected) and enter the ready queue,
 *      a binary heap ordered by arrival time or, in the size-aware order,
 *      by algorithmic bytes + age_bytes_per_s * arrival time (shortest
 *      first with aging, after size-based scheduling of web servers,
 *      Harchol-Balter et al., ACM TOCS 21(2), 2003);
 *   3. admit: the head of the ready queue is admitted while a slot is
 *      free, the HBM budget allows its destination bytes and the team-span
 *      page allocator finds room for it;
 *   4. launch: in batched mode every admitted, not yet running request
 *      joins the next launch group on a free stream (tiled persistent
 *      launches per kernel family, see osc_sched.c); in per-call mode each
 *      request is one per-call kernel on a single stream, in order.
 * Requests therefore join while others are on the GPU, their work overlaps
 * on the device, and memory is recycled as soon as a group finishes.
 *
 * Destinations come from a per-PE arena cut into pages; a request needs
 * the same page range on every PE of its team span (so its destination is
 * one symmetric object for the team), found by OR-ing the span's page
 * bitmaps.
 */
#define _GNU_SOURCE
#include <sched.h>
#include <stdio.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "osc_sched.h"

#define SLOT_TILES ((uint64_t)4 << 20)

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- team-span page allocator ------------------------------------------- */
typedef struct {
  int npes, npages, words;
  uint64_t *bits; /* npes x words */
  uint64_t *tmp;
} pagemap;

static int pm_init(pagemap *m, int npes, int npages) {
  m->npes = npes;
  m->npages = npages;
  m->words = (npages + 63) / 64;
  m->bits = (uint64_t *)calloc((size_t)npes * (size_t)m->words, sizeof(uint64_t));
  m->tmp = (uint64_t *)calloc((size_t)m->words, sizeof(uint64_t));
  return (m->bits && m->tmp) ? OSC_OK : OSC_ERR_NOMEM;
}

static void pm_fini(pagemap *m) {
  free(m->bits);
  free(m->tmp);
}

static void pm_mark(pagemap *m, int lo, int hi, int first, int k, int set) {
  int p, i;
  for (p = lo; p <= hi; ++p) {
    uint64_t *row = m->bits + (size_t)p * (size_t)m->words;
    for (i = first; i < first + k; ++i) {
      if (set)
        row[i >> 6] |= 1ull << (i & 63);
      else
        row[i >> 6] &= ~(1ull << (i & 63));
    }
  }
}

/* first run of k pages free on every PE of [lo, hi]; -1 if none */
static int pm_find(pagemap *m, int lo, int hi, int k) {
  int p, w, i, run = 0;
  memset(m->tmp, 0, sizeof(uint64_t) * (size_t)m->words);
  for (p = lo; p <= hi; ++p) {
    const uint64_t *row = m->bits + (size_t)p * (size_t)m->words;
    for (w = 0; w < m->words; ++w) m->tmp[w] |= row[w];
  }
  for (i = 0; i < m->npages; ++i) {
    if ((m->tmp[i >> 6] >> (i & 63)) & 1ull) {
      run = 0;
      if (m->tmp[i >> 6] == ~0ull) i |= 63;
      continue;
    }
    if (++run == k) return i - k + 1;
  }
  return -1;
}

/* ---- launch slots ------------------------------------------------------------ */
typedef struct {
  void *stream;
  void *event;
  int busy, group;
  int *reqs;
  int nreq;
  osc_call *calls;
  oscd_bcall *h_calls, *d_calls;
  oscd_tile *h_tiles, *d_tiles;
  uint64_t *d_digest;
  uint64_t *h_digest;
  osc_plan plan;
} slot_t;

typedef struct {
  int first, npages, lo, hi;
  double area, key;
  uint64_t tiles;
} req_state;

/* ---- ready queue: a binary min-heap of request indices by (key, index) ----- */
static int rq_before(const req_state *rs, int a, int b) {
  return rs[a].key < rs[b].key || (rs[a].key == rs[b].key && a < b);
}

static void rq_push(int *h, int *n, const req_state *rs, int r) {
  int i = (*n)++;
  while (i > 0) {
    const int up = (i - 1) / 2;
    if (!rq_before(rs, r, h[up])) break;
    h[i] = h[up];
    i = up;
  }
  h[i] = r;
}

static void rq_pop(int *h, int *n, const req_state *rs) {
  const int last = h[--(*n)];
  int i = 0;
  for (;;) {
    int c = 2 * i + 1;
    if (c >= *n) break;
    if (c + 1 < *n && rq_before(rs, h[c + 1], h[c])) ++c;
    if (!rq_before(rs, h[c], last)) break;
    h[i] = h[c];
    i = c;
  }
  if (*n > 0) h[i] = last;
}

static int slot_init(slot_t *s, int max_group, void *stream) {
  int rc;
  memset(s, 0, sizeof(*s));
  osc_plan_init(&s->plan);
  s->stream = stream;
  s->reqs = (int *)calloc((size_t)max_group, sizeof(int));
  s->calls = (osc_call *)calloc((size_t)max_group, sizeof(osc_call));
  s->h_digest = (uint64_t *)calloc((size_t)max_group, sizeof(uint64_t));
  if (!s->reqs || !s->calls || !s->h_digest) return OSC_ERR_NOMEM;
  rc = oscd_event_create(&s->event);
  if (rc == OSC_OK) rc = oscd_host_alloc((void **)&s->h_calls, sizeof(oscd_bcall) * (size_t)max_group);
  if (rc == OSC_OK) rc = oscd_malloc((void **)&s->d_calls, sizeof(oscd_bcall) * (size_t)max_group);
  if (rc == OSC_OK) rc = oscd_host_alloc((void **)&s->h_tiles, sizeof(oscd_tile) * SLOT_TILES);
  if (rc == OSC_OK) rc = oscd_malloc((void **)&s->d_tiles, sizeof(oscd_tile) * SLOT_TILES);
  if (rc == OSC_OK) rc = oscd_malloc((void **)&s->d_digest, sizeof(uint64_t) * (size_t)max_group);
  return rc;
}

static void slot_fini(slot_t *s) {
  if (s->event) oscd_event_destroy(s->event);
  if (s->h_calls) oscd_host_free(s->h_calls);
  if (s->d_calls) oscd_free(s->d_calls);
  if (s->h_tiles) oscd_host_free(s->h_tiles);
  if (s->d_tiles) oscd_free(s->d_tiles);
  if (s->d_digest) oscd_free(s->d_digest);
  osc_plan_free(&s->plan);
  free(s->reqs);
  free(s->calls);
  free(s->h_digest);
}

/* digests of the slot's calls (already in d_calls), synchronously */
static int slot_digest(slot_t *s) {
  int rc = oscd_fill_u64(s->d_digest, 0, (uint64_t)s->nreq, s->stream);
  if (rc == OSC_OK)
    rc = oscd_batch_digest(s->d_calls, (uint32_t)s->nreq, s->d_digest, osc_g.base,
                           osc_g.slice, s->stream);
  if (rc == OSC_OK) rc = oscd_stream_sync(s->stream);
  if (rc == OSC_OK)
    rc = oscd_memcpy_d2h(s->h_digest, s->d_digest, sizeof(uint64_t) * (size_t)s->nreq);
  return rc;
}

static int cmp_double(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double pct(const double *v, int n, double q) {
  int i;
  if (n == 0) return 0.0;
  i = (int)(q * (double)n + 0.999999) - 1; /* nearest rank */
  if (i < 0) i = 0;
  if (i >= n) i = n - 1;
  return v[i];
}

int osc_serve_run(osc_request *reqs, int nreq, const osc_serve_cfg *cfg,
                  osc_serve_stats *st) {
  const int batched = cfg->mode == OSC_SERVE_BATCHED;
  const int nslots = batched ? (cfg->nstreams > 0 ? cfg->nstreams : 1) : 1;
  const int max_group = batched ? (cfg->max_group > 0 ? cfg->max_group : 64) : 1;
  const uint64_t page = cfg->page_bytes ? cfg->page_bytes : ((uint64_t)256 << 10);
  slot_t *slots = NULL;
  void **streams = NULL;
  void **events = NULL; /* per-call mode: one event per request */
  req_state *rs = NULL;
  int *ready = NULL;   /* arrived, not yet admitted (heap) */
  int nready = 0;
  int *waiting = NULL; /* admitted, not yet launched (FIFO) */
  int nwait = 0, wait_head = 0;
  int *running = NULL; /* per-call mode: launched, not retired (FIFO) */
  int run_head = 0, run_tail = 0;
  pagemap pm;
  int i, rc = OSC_OK, next = 0, done = 0, inflight = 0, groups = 0;
  double inflight_area = 0.0, t0;
  memset(&pm, 0, sizeof(pm));
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  if (reqs == NULL || nreq < 0 || cfg == NULL || st == NULL || cfg->max_inflight < 1)
    return osc_set_error(OSC_ERR_ARG, "serve arguments");
  if ((cfg->order != OSC_SERVE_ORDER_ARRIVAL && cfg->order != OSC_SERVE_ORDER_SIZE) ||
      !(cfg->age_bytes_per_s >= 0.0 && cfg->age_bytes_per_s <= 1e30))
    return osc_set_error(OSC_ERR_ARG, "serve order %d / age weight %g", cfg->order,
                         cfg->age_bytes_per_s);
  /* written so that no sum wraps: base, then size within what is left */
  if (cfg->dest_base > osc_g.heap_bytes ||
      cfg->dest_bytes > osc_g.heap_bytes - cfg->dest_base || page % 256 != 0 ||
      cfg->dest_base % 256 != 0)
    return osc_set_error(OSC_ERR_ARG, "destination arena outside the heap or misaligned");
  if (cfg->dest_bytes / page < 1 || cfg->dest_bytes / page > (uint64_t)INT_MAX)
    return osc_set_error(OSC_ERR_ARG, "destination arena of %llu pages",
                         (unsigned long long)(cfg->dest_bytes / page));
  for (i = 1; i < nreq; ++i)
    if (reqs[i].arrival_s < reqs[i - 1].arrival_s)
      return osc_set_error(OSC_ERR_ARG, "arrivals must be non-decreasing (request %d)", i);
  memset(st, 0, sizeof(*st));
  rs = (req_state *)calloc((size_t)(nreq ? nreq : 1), sizeof(req_state));
  ready = (int *)calloc((size_t)(nreq ? nreq : 1), sizeof(int));
  waiting = (int *)calloc((size_t)(nreq ? nreq : 1), sizeof(int));
  running = (int *)calloc((size_t)(nreq ? nreq : 1), sizeof(int));
  slots = (slot_t *)calloc((size_t)nslots, sizeof(slot_t));
  streams = (void **)calloc((size_t)nslots, sizeof(void *));
  if (!rs || !ready || !waiting || !running || !slots || !streams) {
    rc = osc_set_error(OSC_ERR_NOMEM, "serve state");
    goto out;
  }
  rc = pm_init(&pm, osc_g.npes, (int)(cfg->dest_bytes / page));
  if (rc != OSC_OK) goto out;
  for (i = 0; i < nslots && rc == OSC_OK; ++i) {
    rc = oscd_stream_create(&streams[i]);
    if (rc == OSC_OK) rc = slot_init(&slots[i], max_group, streams[i]);
  }
  if (rc == OSC_OK && !batched) {
    events = (void **)calloc((size_t)(nreq ? nreq : 1), sizeof(void *));
    if (events == NULL) rc = OSC_ERR_NOMEM;
  }
  if (rc != OSC_OK) {
    osc_set_error(rc, "serve resources: %s", oscd_last_hip_error());
    goto out;
  }
  for (i = 0; i < nreq; ++i) {
    reqs[i].status = 0;
    reqs[i].group = -1;
    reqs[i].admit_s = reqs[i].launch_s = reqs[i].done_s = -1.0;
    reqs[i].digest = 0;
  }

  t0 = now_s();
  while (done < nreq && rc == OSC_OK) {
    int progress = 0;
    double now = now_s() - t0;
    /* 1. retire */
    if (batched) {
      for (i = 0; i < nslots; ++i) {
        slot_t *s = &slots[i];
        int k, q = 0;
        if (!s->busy) continue;
        q = oscd_event_query(s->event);
        if (q < 0) {
          rc = osc_hip_error(q, "event query");
          break;
        }
        if (q == 0) continue;
        if (cfg->digests && (rc = slot_digest(s)) != OSC_OK) {
          osc_hip_error(rc, "digest");
          break;
        }
        now = now_s() - t0;
        for (k = 0; k < s->nreq; ++k) {
          const int r = s->reqs[k];
          reqs[r].done_s = now;
          if (cfg->digests) reqs[r].digest = s->h_digest[k];
          pm_mark(&pm, rs[r].lo, rs[r].hi, rs[r].first, rs[r].npages, 0);
          inflight_area -= rs[r].area;
          --inflight;
          ++done;
        }
        s->busy = 0;
        progress = 1;
      }
    } else {
      while (run_head < run_tail) {
        const int r = running[run_head];
        const int q = oscd_event_query(events[r]);
        if (q < 0) {
          rc = osc_hip_error(q, "event query");
          break;
        }
        if (q == 0) break;
        now = now_s() - t0;
        if (cfg->digests) {
          uint64_t d = 0;
          {
            slot_t *s = &slots[0];
            s->nreq = 1;
            osc_call_to_device(&reqs[r].call, &s->h_calls[0]);
            rc = oscd_memcpy_h2d(s->d_calls, s->h_calls, sizeof(oscd_bcall));
            if (rc == OSC_OK) rc = slot_digest(s);
            d = s->h_digest[0];
          }
          if (rc != OSC_OK) {
            osc_hip_error(rc, "digest");
            break;
          }
          reqs[r].digest = d;
          now = now_s() - t0;
        }
        reqs[r].done_s = now;
        oscd_event_destroy(events[r]);
        events[r] = NULL;
        pm_mark(&pm, rs[r].lo, rs[r].hi, rs[r].first, rs[r].npages, 0);
        inflight_area -= rs[r].area;
        --inflight;
        ++done;
        ++run_head;
        progress = 1;
      }
    }
    if (rc != OSC_OK) break;
    /* 2. arrivals: reject what can never run, queue the rest */
    now = now_s() - t0;
    while (next < nreq && reqs[next].arrival_s <= now) {
      osc_request *q = &reqs[next];
      const osc_call *c = &q->call;
      const uint64_t db = osc_call_dest_bytes(c);
      const double area = (double)db * (double)c->team_npes;
      int vrc;
      { /* validate with the destination that will be assigned */
        osc_call tmp = *c;
        tmp.dest_off = cfg->dest_base;
        vrc = osc_call_validate(&tmp, next);
        /* a source inside the destination arena could be overwritten by
         * another request in flight */
        if (vrc == OSC_OK && osc_call_src_bytes(c) > 0 &&
            c->src_off < cfg->dest_base + cfg->dest_bytes &&
            cfg->dest_base < c->src_off + osc_call_src_bytes(c))
          vrc = osc_set_error(OSC_ERR_ARG, "request %d: source inside the destination arena",
                              next);
        /* the arena holds dest_bytes / page whole pages */
        if (vrc == OSC_OK && (db + page - 1) / page > cfg->dest_bytes / page)
          vrc = OSC_ERR_BUDGET;
        if (vrc == OSC_OK && area > cfg->budget_bytes) vrc = OSC_ERR_BUDGET;
        if (vrc == OSC_OK && batched) {
          rs[next].tiles = osc_call_tile_count(&tmp);
          if (rs[next].tiles > SLOT_TILES) vrc = OSC_ERR_BUDGET;
        }
      }
      if (vrc != OSC_OK) { /* can never run: reject */
        q->status = vrc;
        q->admit_s = q->done_s = now;
        ++st->rejected;
        ++done;
      } else {
        rs[next].npages = (int)((db + page - 1) / page);
        rs[next].lo = c->team_start;
        rs[next].hi = c->team_start + (c->team_npes - 1) * c->team_stride;
        rs[next].area = area;
        rs[next].key = cfg->order == OSC_SERVE_ORDER_SIZE
                           ? osc_call_alg_bytes(c) + cfg->age_bytes_per_s * q->arrival_s
                           : q->arrival_s;
        rq_push(ready, &nready, rs, next);
      }
      ++next;
      progress = 1;
    }
    /* 3. admit the head of the ready queue while slots and memory allow */
    while (nready > 0 && inflight < cfg->max_inflight) {
      const int r = ready[0];
      osc_request *q = &reqs[r];
      int first = 0;
      if (inflight_area + rs[r].area > cfg->budget_bytes) break;
      if (rs[r].npages > 0) {
        first = pm_find(&pm, rs[r].lo, rs[r].hi, rs[r].npages);
        if (first < 0) break;
        pm_mark(&pm, rs[r].lo, rs[r].hi, first, rs[r].npages, 1);
      }
      rq_pop(ready, &nready, rs);
      rs[r].first = first;
      q->call.dest_off = cfg->dest_base + (uint64_t)first * page;
      q->admit_s = now;
      inflight_area += rs[r].area;
      if (inflight_area > st->peak_inflight_bytes) st->peak_inflight_bytes = inflight_area;
      ++inflight;
      if (inflight > st->max_inflight_seen) st->max_inflight_seen = inflight;
      waiting[nwait++] = r;
      progress = 1;
    }
    /* 4. launch */
    if (batched) {
      for (i = 0; i < nslots && wait_head < nwait; ++i) {
        slot_t *s = &slots[i];
        uint64_t tiles = 0;
        int k;
        if (s->busy) continue;
        s->nreq = 0;
        while (wait_head < nwait && s->nreq < max_group) {
          const int r = waiting[wait_head];
          if (s->nreq > 0 && tiles + rs[r].tiles > SLOT_TILES) break;
          tiles += rs[r].tiles;
          s->reqs[s->nreq] = r;
          s->calls[s->nreq] = reqs[r].call;
          s->nreq++;
          wait_head++;
        }
        rc = osc_plan_build_into(&s->plan, s->calls, s->nreq, 0, s->h_tiles, SLOT_TILES);
        if (rc != OSC_OK) break;
        for (k = 0; k < s->nreq; ++k) osc_call_to_device(&s->calls[k], &s->h_calls[k]);
        rc = oscd_memcpy_h2d_async(s->d_calls, s->h_calls, sizeof(oscd_bcall) * (size_t)s->nreq,
                                   s->stream);
        if (rc == OSC_OK && s->plan.ntiles)
          rc = oscd_memcpy_h2d_async(s->d_tiles, s->h_tiles, sizeof(oscd_tile) * s->plan.ntiles,
                                     s->stream);
        if (rc == OSC_OK) rc = osc_plan_enqueue(&s->plan, s->calls, s->d_calls, s->d_tiles, s->stream);
        if (rc == OSC_OK) rc = oscd_event_record(s->event, s->stream);
        if (rc != OSC_OK) {
          osc_hip_error(rc, "group launch");
          break;
        }
        now = now_s() - t0;
        for (k = 0; k < s->nreq; ++k) {
          reqs[s->reqs[k]].launch_s = now;
          reqs[s->reqs[k]].group = groups;
        }
        s->group = groups++;
        s->busy = 1;
        progress = 1;
      }
    } else {
      while (wait_head < nwait) {
        const int r = waiting[wait_head++];
        rc = osc_call_launch_single(&reqs[r].call, streams[0]);
        if (rc == OSC_OK) rc = oscd_event_create(&events[r]);
        if (rc == OSC_OK) rc = oscd_event_record(events[r], streams[0]);
        if (rc != OSC_OK) {
          osc_hip_error(rc, "per-call launch");
          break;
        }
        reqs[r].launch_s = now_s() - t0;
        reqs[r].group = groups++;
        running[run_tail++] = r;
        progress = 1;
      }
    }
    if (!progress) sched_yield();
  }
  st->wall_s = now_s() - t0;

  if (rc == OSC_OK) {
    double *lat = (double *)malloc(sizeof(double) * (size_t)(nreq ? nreq : 1));
    int n = 0;
    double sum = 0.0;
    if (lat == NULL) {
      rc = OSC_ERR_NOMEM;
      goto out;
    }
    for (i = 0; i < nreq; ++i) {
      if (reqs[i].status != 0) continue;
      lat[n] = (reqs[i].done_s - reqs[i].arrival_s) * 1e3;
      sum += lat[n];
      st->algorithmic_bytes += osc_call_alg_bytes(&reqs[i].call);
      ++n;
    }
    qsort(lat, (size_t)n, sizeof(double), cmp_double);
    st->completed = n;
    st->groups = groups;
    st->mean_ms = n ? sum / n : 0.0;
    st->p50_ms = pct(lat, n, 0.50);
    st->p90_ms = pct(lat, n, 0.90);
    st->p99_ms = pct(lat, n, 0.99);
    st->max_ms = n ? lat[n - 1] : 0.0;
    st->gbps = st->wall_s > 0 ? st->algorithmic_bytes / st->wall_s / 1e9 : 0.0;
    free(lat);
  }
out:
  for (i = 0; slots && i < nslots; ++i) {
    if (streams && streams[i]) oscd_stream_sync(streams[i]);
    slot_fini(&slots[i]);
    if (streams && streams[i]) oscd_stream_destroy(streams[i]);
  }
  if (events) {
    for (i = 0; i < nreq; ++i)
      if (events[i]) oscd_event_destroy(events[i]);
    free(events);
  }
  pm_fini(&pm);
  free(slots);
  free(streams);
  free(rs);
  free(ready);
  free(waiting);
  free(running);
  return rc;
}

int osc_request_alone(const osc_call *call, uint64_t dest_off,
                      uint64_t *digest) {
  osc_call c;
  oscd_bcall h;
  oscd_bcall *d_call;
  uint64_t *d_dig;
  int rc;
  if (call == NULL) return osc_set_error(OSC_ERR_ARG, "null call");
  if (!osc_g.initialized) return osc_set_error(OSC_ERR_NOTINIT, "not initialised");
  c = *call;
  c.dest_off = dest_off;
  rc = osc_call_validate(&c, 0);
  if (rc != OSC_OK) return rc;
  if (osc_g.aux_dev == NULL) {
    rc = oscd_malloc(&osc_g.aux_dev, 256);
    if (rc != OSC_OK) return osc_hip_error(rc, "aux buffer");
  }
  d_call = (oscd_bcall *)osc_g.aux_dev;
  d_dig = (uint64_t *)((char *)osc_g.aux_dev + 128);
  osc_call_to_device(&c, &h);
  rc = osc_call_launch_single(&c, osc_g.stream);
  if (rc == OSC_OK) rc = oscd_stream_sync(osc_g.stream);
  if (rc == OSC_OK) rc = oscd_memcpy_h2d(d_call, &h, sizeof(h));
  if (rc == OSC_OK) rc = oscd_fill_u64(d_dig, 0, 1, osc_g.stream);
  if (rc == OSC_OK) rc = oscd_batch_digest(d_call, 1, d_dig, osc_g.base, osc_g.slice, osc_g.stream);
  if (rc == OSC_OK) rc = oscd_stream_sync(osc_g.stream);
  if (rc == OSC_OK && digest) rc = oscd_memcpy_d2h(digest, d_dig, sizeof(uint64_t));
  return rc == OSC_OK ? OSC_OK : osc_hip_error(rc, "request alone");
}
