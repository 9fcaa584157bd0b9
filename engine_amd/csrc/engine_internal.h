/*
 * engine_internal.h -- internal C interface between the C host layer
 * (graph.c, legacy.c, packing_ops.c) and the HIP engine (engine.hip).
 * Not installed; callers use paradmm.h / paradmm_gpu.h.
 */
#ifndef PARADMM_ENGINE_INTERNAL_H
#define PARADMM_ENGINE_INTERNAL_H

#include <stdint.h>

#include "paradmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pass bits of a legacy update call, run in this order: X, M, Z, U, N */
enum {
    PDM_PASS_X = 1,
    PDM_PASS_M = 2,
    PDM_PASS_Z = 4,
    PDM_PASS_U = 8,
    PDM_PASS_N = 16
};

/* Runs the given passes once on the GPU with host coherence: uploads what
 * they read, downloads what they write.  Returns 0 or an error code. */
int pdm_legacy_passes(graph *G, int passes);
/* maxiter whole iterations on the GPU, host synchronised before and after. */
int pdm_legacy_iterations(graph *G, int maxiter);
/* the warm-start sweeps of initialize_X_N_Z_M_U_optimal on the GPU */
int pdm_legacy_warm_start(graph *G, int maxiter);

/* The host arrays were (re)allocated or rewritten by an initialiser. */
void pdm_host_written(graph *G);
/* Brings the host arrays up to date when the device holds newer state
 * (no-op without device state). */
int pdm_sync_if_device_newer(graph *G);
/* Frees the device state of G (closeG). */
void pdm_release_graph(graph *G);

/* Legacy entry points return void: on error they print the message and exit
 * with EXIT_FAILURE, as the reference does on allocation failure, unless the
 * environment sets PARADMM_NO_EXIT=1 (then the error is only recorded). */
void pdm_legacy_fail(int code);

void pdm_set_error(int code, const char *fmt, ...);

/* uniform double in [0, 1) from a splitmix64 stream (packing_ops.c) */
double pdm_uniform01(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
