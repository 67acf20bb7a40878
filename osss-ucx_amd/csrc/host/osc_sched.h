/*
 * osc_sched.h - tile plans shared by the batched and served modes.
 *
 * A plan turns a list of calls into tiles (oscd_tile) grouped by kernel
 * family into launches; enqueueing a plan issues one persistent tile-loop
 * kernel per launch on a stream.
 */
#ifndef OSC_SCHED_H
#define OSC_SCHED_H 1

#include "osc_internal.h"

/* launch kinds besides OSC_CALL_REDUCE */
#define OSC_LAUNCH_MOVE (-1)   /* tile loop over data-movement tiles */
#define OSC_LAUNCH_DIRECT (-2) /* one call through the per-call kernel */

typedef struct {
  int wave;
  int kind; /* OSC_CALL_REDUCE, OSC_LAUNCH_MOVE or OSC_LAUNCH_DIRECT */
  int call; /* OSC_LAUNCH_DIRECT: index of the call */
  int dtype, op, mode, shape, width;
  int maxlg, maxpes;
  uint32_t grid;
  uint64_t tile0, ntiles;
  double bytes; /* algorithmic bytes of the launch's calls */
} osc_launch;

typedef struct {
  osc_launch *launches;
  int nlaunches, cap_launches;
  oscd_tile *tiles; /* host copy, grouped by launch */
  uint64_t ntiles, cap_tiles;
} osc_plan;

uint64_t osc_call_dest_bytes(const osc_call *c);
uint64_t osc_call_src_bytes(const osc_call *c);
double osc_call_alg_bytes(const osc_call *c);
/* 1 if the call's team (start, stride, size) lies inside a world of
 * world_npes PEs */
int osc_call_team_fits(const osc_call *c, int world_npes);
int osc_call_validate(const osc_call *c, int i);
void osc_call_to_device(const osc_call *c, oscd_bcall *d);
/* upper bound of the tiles osc_plan_build makes for one call */
uint64_t osc_call_tile_count(const osc_call *c);

void osc_plan_init(osc_plan *p);
void osc_plan_free(osc_plan *p);
/* Tile calls[0..ncalls) (tile.call = index into calls) and group the tiles
 * into launches ordered by (wave, family).  use_waves = 0 puts every call
 * in wave 0 (the caller guarantees disjoint destinations). */
int osc_plan_build(osc_plan *p, const osc_call *calls, int ncalls,
                   int use_waves);
/* The same, writing the tiles to `out` (capacity `cap` tiles) instead of
 * p->tiles; OSC_ERR_BUDGET if they do not fit. */
int osc_plan_build_into(osc_plan *p, const osc_call *calls, int ncalls,
                        int use_waves, oscd_tile *out, uint64_t cap);
/* Issue the plan's launches; d_tiles holds p->tiles (same order), calls
 * and d_calls the host and device call arrays the plan was built from. */
int osc_plan_enqueue(const osc_plan *p, const osc_call *calls,
                     const oscd_bcall *d_calls, const oscd_tile *d_tiles,
                     void *stream);
/* The same, with the launches of each wave spread over nbranch streams
 * (float MIN/MAX reductions with flag_off[call] != UINT64_MAX use
 * d_flags + flag_off[call] as their span flags in two-phase waves; either
 * may be NULL):
 * `stream` and aux[0 .. nbranch-2], forked from `stream` at the start of
 * the wave through ev_fork and joined back at its end through
 * ev_join[0 .. nbranch-2] (launches of one wave are independent: disjoint
 * destinations, read-only sources).  Branches are balanced by algorithmic
 * bytes.  Works under stream capture, where it records a graph whose
 * waves are sets of parallel branches. */
int osc_plan_enqueue_branches(const osc_plan *p, const osc_call *calls,
                              const oscd_bcall *d_calls,
                              const oscd_tile *d_tiles, void *stream,
                              int nbranch, void *const *aux, void *ev_fork,
                              void *const *ev_join, uint8_t *d_flags,
                              const uint64_t *flag_off);
/* Bytes of span flags a float MIN/MAX reduction may need in a two-phase
 * wave (0 for every other call). */
uint64_t osc_call_flag_bytes(const osc_call *c);
/* One call through the per-call kernels (the "one by one" path). */
int osc_call_launch_single(const osc_call *c, void *stream);

#endif
