/*
 * paradmm_gpu.h -- MI355X extension of the parADMM C API.
 *
 * Everything here is additive to paradmm.h: plain C, HIP-free, pointers and
 * sizes only.  Every function returns PARADMM_OK (0) or a positive error code
 * whose text paradmm_last_error() returns.
 *
 * Reference counterparts are cited per entry point.  The reference has no
 * residuals, no device-resident iteration API, no batching and no rho policy;
 * those entry points say so.
 */
#ifndef PARADMM_GPU_H
#define PARADMM_GPU_H

#include <stdint.h>
#include "paradmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*paradmm_prox_fn)(double *x, double *n, double *rhos, int nvars, int ndims, void *params);

enum {
    PARADMM_OK = 0,
    PARADMM_ERR_ARG = 1,     /* bad argument */
    PARADMM_ERR_KIND = 2,    /* a factor's callback is not a known device prox kind */
    PARADMM_ERR_SHAPE = 3,   /* factor degree / num_dims do not fit its kind */
    PARADMM_ERR_HIP = 4,     /* HIP runtime error */
    PARADMM_ERR_NOMEM = 5,   /* host or device allocation failed */
    PARADMM_ERR_STATE = 6    /* graph not initialised / arrays missing */
};

/* Device prox kinds.  Each is a fixed operator of the packing factor library. */
enum {
    PARADMM_KIND_NONE = 0,
    PARADMM_KIND_PAIR_2D = 1,    /* packing/host.c:83-132  disk_collision_variable_radius, 4 edges, dim 2 */
    PARADMM_KIND_WALL_2D = 2,    /* packing/host.c:31-65   disk_plane_collision_variable_radius, 2 edges, dim 2,
                                    params: double[4] = unit normal, point */
    PARADMM_KIND_RADIUS_2D = 3,  /* packing/host.c:67-81   maximize_sum_of_squared_disk_radii_alpha, per edge,
                                    params: double[1] = exponent a */
    PARADMM_KIND_PAIR_3D = 4,    /* dim-3 (cx,cy,r) form of packing/host.c:83-132, 2 edges */
    PARADMM_KIND_BOX_3D = 5,     /* projection onto the unit square container, per edge (no reference function) */
    PARADMM_KIND_RADIUS_3D = 6,  /* dim-3 form of packing/host.c:67-81, per edge, params: double[1] = a */
    PARADMM_KIND_COUNT = 7
};

/* Last error message of this thread ("" if none). */
const char *paradmm_last_error(void);

/* ---- prox kinds ------------------------------------------------------------
 * The reference calls host function pointers from the X pass
 * (core/paradmm.c:590) and device function pointers on the GPU
 * (core/paradmm-gpu.cu:45).  Here a callback address is mapped to a device
 * kind: the operators below are pre-registered, the reference demo's operator
 * names are recognised through the dynamic symbol table (link the caller with
 * -rdynamic), and any other address can be registered explicitly. */
int paradmm_register_prox_kind(paradmm_prox_fn fn, int kind);
int paradmm_prox_kind_of(paradmm_prox_fn fn);

/* Built-in operators, usable with addNode.  They are tokens for the device
 * kinds above: the X pass runs them on the GPU.  Called on the host they
 * print an error and abort. */
void paradmm_prox_disk_collision_2d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void paradmm_prox_disk_plane_2d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void paradmm_prox_max_radius_2d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void paradmm_prox_disk_collision_3d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void paradmm_prox_box_3d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void paradmm_prox_max_radius_3d(double *x, double *n, double *rhos, int nvars, int ndims, void *params);

/* ---- packing problem helpers ---------------------------------------------
 * Unit-square packing with one (cx,cy,r) variable per circle (num_dims 3):
 * N radius factors (exponent *a_exp, stored by the graph), N box factors,
 * N(N-1)/2 pair factors in i-major order -- the factor order of
 * packing/host.c:298-335 with the triangle walls replaced by one box. */
int paradmm_packing_square(graph *G, int ncircles, double a_exp);
/* Allocates X, N, M, U, Z (as initialize_X_N_Z_M_U_zeros, core/paradmm.c:210)
 * and fills them from splitmix64(seed): cx, cy ~ U[0,1], r ~ U[0,0.05] per
 * circle; X = N = M = Z of the edge's variable, U = 0.  The reference seeds
 * from time(NULL) (core/paradmm.c:247); this is the reproducible form. */
int paradmm_packing_init_square(graph *G, uint64_t seed);

/* ---- device-resident iteration --------------------------------------------
 * Equivalent to the reference's GPU driver (packing/gpu.cu:250-291: copy the
 * graph, iterate, copy Z back) without per-iteration host traffic.  After
 * prepare/iterate the device holds the newest state: read it back with
 * paradmm_gpu_sync_to_host; after host writes, call paradmm_gpu_sync_from_host.
 * Legacy update calls made in between synchronise by themselves. */
int paradmm_gpu_prepare(graph *G);          /* finalise the layout and upload (implicit on first use) */
int paradmm_gpu_iterate(graph *G, int iters);
int paradmm_gpu_sync_to_host(graph *G);     /* download X, M, N, U, Z, RHOS */
int paradmm_gpu_sync_from_host(graph *G);   /* upload X, N, U, Z, RHOS, ALPHAS */
int paradmm_gpu_release(graph *G);          /* free device state (closeG does this too) */
/* 1 = generic CSR layout, 2 = dense pair-tile layout; 0 before prepare */
int paradmm_gpu_layout(graph *G);
/* NEW: FP64 residuals of the last device iteration (the reference has none):
 * primal2 = sum_e ||X_e - Z_v(e)||^2, dual2 = sum_e rho_e^2 ||Z_v - Zprev_v||^2. */
int paradmm_gpu_residuals(graph *G, double *primal2, double *dual2);
/* Runs iters iterations and returns the elapsed milliseconds measured with HIP
 * events on the engine's stream (inputs resident, no host traffic). */
int paradmm_gpu_time_iterations(graph *G, int iters, double *ms);
/* Runs iters iterations launched one kernel at a time with HIP events around
 * every kernel on the engine's stream; per distinct kernel name returns the
 * mean duration (avg_ms[j], ms) and launch count.  names receives up to
 * max_kernels NUL-terminated names of 64 bytes each. */
int paradmm_gpu_profile_kernels(graph *G, int iters, int max_kernels, char *names, double *avg_ms,
                                int *count, int *nkernels);

#ifdef __cplusplus
}
#endif
#endif
