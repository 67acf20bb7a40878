/*
 * osc_spmd_main.c - main() for SPMD OpenSHMEM programs (lib/libosc_spmd_main.a).
 *
 * Build the program with -Dmain=osc_pe_main and link this archive and
 * libosc.so.  This main() starts OSC_NPES host threads (default 8), one per
 * emulated PE, each running the program's own main, as the reference's
 * oshrun launcher starts one process per PE.  SHMEM_SYMMETRIC_SIZE sets the
 * heap per PE (default 32M, the reference's default) and OSC_DEVICE the
 * HIP device.  The exit status is 0 when every PE's main returned 0.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../../include/osc.h"

int osc_pe_main(int argc, char **argv);

typedef struct {
  int argc;
  char **argv;
} main_args;

static int run_pe(int pe, void *a) {
  main_args *m = (main_args *)a;
  (void)pe;
  return osc_pe_main(m->argc, m->argv);
}

int main(int argc, char **argv) {
  const char *np = getenv("OSC_NPES");
  const char *hs = getenv("SHMEM_SYMMETRIC_SIZE");
  const char *dv = getenv("OSC_DEVICE");
  size_t heap = (size_t)32 << 20;
  main_args m;
  int rc;
  if (hs != NULL && osc_parse_size(hs, &heap) != 0) {
    fprintf(stderr, "bad SHMEM_SYMMETRIC_SIZE=%s\n", hs);
    return 2;
  }
  m.argc = argc;
  m.argv = argv;
  rc = osc_spmd_run(np ? atoi(np) : 8, heap, dv ? atoi(dv) : 0, run_pe, &m);
  if (rc < 0)
    fprintf(stderr, "osc_spmd_run: %s (%s)\n", osc_strerror(rc), osc_last_error());
  return rc == 0 ? 0 : 1;
}
