/*
 * shmemx.h - extensions beside the OpenSHMEM API (the reference's
 * include/shmemx.h); only the wall-clock timer has a counterpart here.
 */
#ifndef OSC_SHMEMX_H
#define OSC_SHMEMX_H

#include "shmem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* seconds from an arbitrary fixed point, monotonic (reference
 * include/shmemx.h:43) */
double shmemx_wtime(void);

#ifdef __cplusplus
}
#endif
#endif /* OSC_SHMEMX_H */
