/*
 * osc_tune.c - the tile/schedule table the batched planner reads.
 *
 * One row per (call class, size bucket): the reduce kernel shape, the tile
 * length and the grid cap of the family launch.  Classes: 0 reduce of
 * 1..4-byte elements, 1 reduce of 8/16-byte elements, 2 fcollect,
 * 3 alltoall, 4 broadcast.  Buckets by destination bytes per PE of the
 * call: <64 KiB, <1 MiB, <16 MiB, >=16 MiB.  The tile length is in block
 * iterations for reductions and in KiB for data movement; tile 0 runs each
 * call of the row on its own through the per-call kernel (a call that
 * fills the chip alone gains nothing from sharing a tile-loop launch).
 *
 * The table starts from built-in defaults.  The autotuner (the library's
 * osc_tune_* calls driven by osss-ucx_amd/autotune.py) sweeps candidates on
 * the GPU per row, keeps the fastest and saves the table; a saved table is
 * loaded at init from OSC_TUNE_FILE, else from the library's own
 * ../tuning/mi355x.tune when present.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "osc_internal.h"

#define NCLS 5
#define NBKT 4

typedef struct {
  int shape, tile, grid;
} tune_row;

static tune_row g_tab[NCLS][NBKT] = {
    /* reduce, small elements */
    {{3, 1, 2048}, {3, 2, 2048}, {0, 1, 2048}, {0, 2, 2048}},
    /* reduce, 8/16-byte elements */
    {{3, 1, 2048}, {3, 2, 2048}, {0, 1, 2048}, {0, 2, 2048}},
    /* fcollect */
    {{0, 16, 2048}, {0, 64, 2048}, {0, 64, 2048}, {0, 64, 2048}},
    /* alltoall */
    {{0, 16, 2048}, {0, 64, 2048}, {0, 128, 2048}, {0, 128, 2048}},
    /* broadcast */
    {{0, 16, 2048}, {0, 64, 2048}, {0, 64, 2048}, {0, 64, 2048}},
};

int osc_tune_bucket(uint64_t bytes_per_pe) {
  if (bytes_per_pe < ((uint64_t)64 << 10)) return 0;
  if (bytes_per_pe < ((uint64_t)1 << 20)) return 1;
  if (bytes_per_pe < ((uint64_t)16 << 20)) return 2;
  return 3;
}

int osc_tune_set(int cls, int bucket, int shape, int tile, int grid) {
  if (cls < 0 || cls >= NCLS || bucket < 0 || bucket >= NBKT || tile < 0 ||
      grid < 1 || shape < 0 || shape > 3)
    return OSC_ERR_ARG;
  g_tab[cls][bucket].shape = shape;
  g_tab[cls][bucket].tile = tile;
  g_tab[cls][bucket].grid = grid;
  return OSC_OK;
}

int osc_tune_get(int cls, int bucket, int *shape, int *tile, int *grid) {
  if (cls < 0 || cls >= NCLS || bucket < 0 || bucket >= NBKT) return OSC_ERR_ARG;
  if (shape) *shape = g_tab[cls][bucket].shape;
  if (tile) *tile = g_tab[cls][bucket].tile;
  if (grid) *grid = g_tab[cls][bucket].grid;
  return OSC_OK;
}

/* text: "class bucket shape tile grid" per line, '#' comments */
int osc_tune_load(const char *path) {
  FILE *f = fopen(path, "r");
  char line[256];
  int n = 0;
  if (f == NULL) return osc_set_error(OSC_ERR_ARG, "cannot open %s", path);
  while (fgets(line, sizeof(line), f) != NULL) {
    int c, b, s, t, g;
    if (line[0] == '#' || line[0] == '\n') continue;
    if (sscanf(line, "%d %d %d %d %d", &c, &b, &s, &t, &g) != 5 ||
        osc_tune_set(c, b, s, t, g) != OSC_OK) {
      fclose(f);
      return osc_set_error(OSC_ERR_ARG, "bad tuning line in %s: %s", path, line);
    }
    ++n;
  }
  fclose(f);
  return OSC_OK;
}

int osc_tune_save(const char *path) {
  FILE *f = fopen(path, "w");
  int c, b;
  if (f == NULL) return osc_set_error(OSC_ERR_ARG, "cannot write %s", path);
  fprintf(f, "# class bucket shape tile grid\n"
             "# class: 0 reduce<=4B, 1 reduce 8/16B, 2 fcollect, 3 alltoall, 4 broadcast\n"
             "# bucket: dest bytes per PE <64KiB, <1MiB, <16MiB, >=16MiB\n"
             "# tile: block iterations (reduce) or KiB (data movement);\n"
             "#   0 = each call alone through the per-call kernel\n");
  for (c = 0; c < NCLS; ++c)
    for (b = 0; b < NBKT; ++b)
      fprintf(f, "%d %d %d %d %d\n", c, b, g_tab[c][b].shape, g_tab[c][b].tile,
              g_tab[c][b].grid);
  fclose(f);
  return OSC_OK;
}

/* Load OSC_TUNE_FILE if set, else <libosc.so dir>/../tuning/mi355x.tune
 * if it exists (errors are reported, defaults kept). */
void osc_tune_init_from_env(void) {
  const char *p = getenv("OSC_TUNE_FILE");
  Dl_info info;
  if (p != NULL && p[0] != '\0') {
    if (osc_tune_load(p) != OSC_OK) fprintf(stderr, "[osc] %s\n", osc_last_error());
    return;
  }
  if (dladdr((void *)&osc_tune_init_from_env, &info) != 0 && info.dli_fname) {
    char path[4096];
    const char *slash = strrchr(info.dli_fname, '/');
    const size_t dir = slash ? (size_t)(slash - info.dli_fname) : 0;
    FILE *f;
    if (dir + 32 >= sizeof(path)) return;
    memcpy(path, info.dli_fname, dir);
    strcpy(path + dir, dir ? "/../tuning/mi355x.tune" : "../tuning/mi355x.tune");
    f = fopen(path, "r");
    if (f != NULL) {
      fclose(f);
      if (osc_tune_load(path) != OSC_OK) fprintf(stderr, "[osc] %s\n", osc_last_error());
    }
  }
}
