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

/* ---- bulk construction -------------------------------------------------------
 * `count` factors of one operator and degree in one call: factor i has the
 * variables vars[i*nvars .. i*nvars+nvars) and the parameters
 * params + i*param_stride bytes (param_stride 0: one shared block).  Edge
 * numbering is that of `count` addNode calls (core/paradmm.c:381-466);
 * the reference has no bulk form. */
int paradmm_add_nodes(graph *G, paradmm_prox_fn op, void *params, size_t param_stride, int count, int nvars,
                      const int *vars);

/* ---- snapshots -----------------------------------------------------------------
 * The iteration state held in the graph's host arrays (X, M, N, U, Z, RHOS,
 * ALPHAS; core/paradmm.h:53-61) to and from a file, with the graph's
 * dimensions and a hash of its edge numbering; loading into a graph of
 * another structure fails with PARADMM_ERR_SHAPE.  Saving first downloads
 * newer device state; loading refreshes the device state before its next
 * use.  No reference counterpart (the reference has no file format). */
int paradmm_save_state(graph *G, const char *path);
int paradmm_load_state(graph *G, const char *path);

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
/* initialize_X_N_Z_M_U_rand (core/paradmm.c:229-262) with a seed: the same
 * bounds and order of draws, from splitmix64 instead of the time-seeded
 * rand(), so runs are reproducible */
int paradmm_init_rand_seeded(graph *G, uint64_t seed, double X_l, double X_u, double N_l, double N_u, double Z_l,
                             double Z_u, double M_l, double M_u, double U_l, double U_u);

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
/* One device array into out (double), leaving the graph's host arrays as
 * they are: PARADMM_READ_Z num_vars x num_dims; PARADMM_READ_X, _U
 * num_edges x num_dims in edge order; PARADMM_READ_RHO num_edges.  A read
 * of Z after a solve moves num_vars x num_dims values instead of every edge
 * array (paradmm_gpu_sync_to_host). */
enum { PARADMM_READ_Z = 0, PARADMM_READ_X = 1, PARADMM_READ_U = 2, PARADMM_READ_RHO = 3 };
int paradmm_gpu_read(graph *G, int which, double *out);
/* 1 = generic CSR layout, 2 = dense pair-tile layout; 0 before prepare */
int paradmm_gpu_layout(graph *G);
/* Whether the dense layout is in its contact state (one rho, alpha = 1:
 * only pairs whose prox was active keep per-edge values), its contact tile
 * side and the pairs its last iteration listed as active; 0, 0, 0 outside
 * it.  No reference counterpart (observability of this engine's state). */
int paradmm_gpu_contact_info(graph *G, int *in_contact, int *tile_side, long *listed_pairs);
/* NEW: FP64 residuals of the last device iteration (the reference has none):
 * primal2 = sum_e ||X_e - Z_v(e)||^2, dual2 = sum_e rho_e^2 ||Z_v - Zprev_v||^2. */
int paradmm_gpu_residuals(graph *G, double *primal2, double *dual2);
/* Runs iters iterations and returns the elapsed milliseconds measured with HIP
 * events on the engine's stream (inputs resident, no host traffic). */
int paradmm_gpu_time_iterations(graph *G, int iters, double *ms);
/* Runs iters iterations launched one kernel at a time with HIP events around
 * every kernel on the engine's stream; per distinct kernel name returns the
 * mean duration (avg_ms[j], ms) and launch count.  names receives up to
 * max_kernels NUL-terminated names of 64 bytes each.  With a rho policy
 * attached (paradmm_gpu_set_rho_policy) the sequence also holds one
 * application of it, as a solve makes it, and one more iteration. */
int paradmm_gpu_profile_kernels(graph *G, int iters, int max_kernels, char *names, double *avg_ms,
                                int *count, int *nkernels);
/* ---- solving to tolerance (NEW: the reference runs a fixed iteration count,
 * packing/host.c:237-244, and has no stopping test) ---------------------------
 * Iterates in chunks of check_every, after each chunk computes the FP64
 * residuals as RMS values, sqrt(primal2 / (num_dims * num_edges)) and
 * sqrt(dual2 / (num_dims * num_edges)), and stops when both are <= their
 * tolerance and, if tol_feas >= 0, the unit-square packing read from Z has
 * max pair overlap and max container violation <= tol_feas.  A rho policy
 * attached with paradmm_gpu_set_rho_policy runs every policy_every
 * iterations (0: fixed rho).  elapsed_ms is host wall time of the whole solve
 * including the checks. */
typedef struct paradmm_solve_opts {
    int max_iters;
    int check_every;
    double tol_primal;
    double tol_dual;
    double tol_feas;
    int policy_every;
} paradmm_solve_opts;

typedef struct paradmm_solve_result {
    int iterations;
    int converged;
    double primal_rms;
    double dual_rms;
    double max_overlap;   /* < 0 when not evaluated */
    double max_box;
    double objective;     /* sum of radii */
    double elapsed_ms;
} paradmm_solve_result;

int paradmm_gpu_solve(graph *G, const paradmm_solve_opts *opts, paradmm_solve_result *res);
/* Feasibility and objective of a unit-square packing (num_dims 3, Z rows are
 * circles): max pair overlap, max container violation, sum of radii. */
int paradmm_gpu_square_quality(graph *G, double *max_overlap, double *max_box, double *sum_r);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the computation of paradmm_gpu_square_quality on a
 * caller-owned device array Z (ncircles x 3 fp32, rows cx, cy, r), writing
 * out3 (device memory, 3 doubles: max pair overlap, max container
 * violation, sum of radii) asynchronously on `stream` (a hipStream_t; NULL
 * is the null stream).  Its per-circle partials live in a static device
 * array of PARADMM_QUALITY_NOALLOC_MAX circles (1.5 MB per device, held
 * until the process ends); more circles is PARADMM_ERR_ARG.  It allocates
 * nothing; calls that overlap on one device race on that array. */
#define PARADMM_QUALITY_NOALLOC_MAX 65536
int paradmm_square_quality_noalloc(const float *Z, int ncircles, double *out3, void *stream);

/* ---- learned rho (NEW: the reference's hook update_outgoing_rhos,
 * core/paradmm.c:564-566, is empty) ---------------------------------------
 * A 6 -> 32 -> 32 -> 1 ReLU MLP maps per-edge features to log10(rho_e);
 * forward and backward run on bf16 MFMA with fp32 accumulation, training is
 * Adam on fp32 master weights.  Features of edge e (variable v):
 *   log10 |X_e - Z_v| / 5, log10(rho_e |Z_v - Zprev_v|) / 5, log10(rho_e) / 2,
 *   [factor is a pair], [factor is a radius], iteration fraction.
 * Applying the policy sets rho_e = 10^clamp(y, -1, 4) and rescales the scaled
 * dual U_e by rho_old / rho_new.  Parameters: 1313 fp32 in the order
 * W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3. */
typedef struct paradmm_rho_mlp paradmm_rho_mlp;
/* params NULL: He-normal init from seed; else nparams must be 1313 */
int paradmm_rho_mlp_create(paradmm_rho_mlp **out, const float *params, int nparams, uint64_t seed);
void paradmm_rho_mlp_destroy(paradmm_rho_mlp *m);
int paradmm_rho_mlp_get_params(paradmm_rho_mlp *m, float *params, int nparams);
/* y[i] = MLP(feat[i][0..5]) for n rows */
int paradmm_rho_mlp_forward(paradmm_rho_mlp *m, const float *feat, int n, float *y);
/* gradient of 0.5 mean((y - target)^2) over the n rows (one batch) and the loss */
int paradmm_rho_mlp_gradient(paradmm_rho_mlp *m, const float *feat, const float *target, int n, float *grad,
                             double *loss);
/* epochs of shuffled minibatch Adam; final_loss = mean loss of the last epoch */
int paradmm_rho_mlp_train(paradmm_rho_mlp *m, const float *feat, const float *target, int n, int epochs, int batch,
                          float lr, uint64_t seed, double *final_loss);
/* one 16x16x16 bf16 MFMA on fixed operands, D written row-major (layout check) */
int paradmm_rho_mlp_layout_probe(float *out256);
/* per-edge features of the current device state, num_edges x 6 in api edge order */
int paradmm_gpu_rho_features(graph *G, float iter_frac, float *feat);
/* attach (m != NULL) or detach the policy used by paradmm_gpu_solve */
int paradmm_gpu_set_rho_policy(graph *G, paradmm_rho_mlp *m);
/* apply the policy once now */
int paradmm_gpu_apply_rho_policy(graph *G, paradmm_rho_mlp *m, float iter_frac);
/* How the network's per-edge outputs (log10 rho) are applied:
 *   mode 0 (default): each edge takes its own output;
 *   mode 1: each edge takes 10^(mean of the outputs over the edges of its
 *           factor type: pair, container, radius);
 *   mode 2: every edge takes 10^(mean over all edges) -- one rho per graph;
 *   mode 3: as mode 2, except that an edge whose own output exceeds that
 *           mean by more than log10(3.2) keeps its own output (the edges
 *           the network singles out for a higher rho).
 * average = 1 (modes 1-3): the applied mean is the average of this and
 * every earlier application's estimate since the mode was set.  Modes 1-3
 * reduce in a fixed order (deterministic). */
int paradmm_gpu_set_rho_policy_mode(graph *G, int mode, int average);
/* rho_e *= factor on every edge, U_e /= factor (uniform change of rho) */
int paradmm_gpu_scale_rho(graph *G, double factor);
/* save / restore the device iterate (X, U, rho, Z, Zprev): counterfactual rollouts */
int paradmm_gpu_state_save(graph *G);
int paradmm_gpu_state_restore(graph *G);

/* NEW: bandwidth of a plain device-to-device copy of `bytes` (read + write
 * counted), iters launches timed with HIP events: the measured HBM ceiling
 * reported next to the nominal peak. */
int paradmm_gpu_copy_bandwidth(size_t bytes, int iters, double *gbs);

/* ---- batched solve -----------------------------------------------------------
 * Many independent unit-square packings (each graph the all-pairs dim-3
 * packing built by paradmm_packing_square, or the same structure built with
 * addNode) solved together by the plain fixed-rho method in ONE persistent
 * cooperative launch.  Each instance is checked every check_every iterations
 * (primal and dual RMS, max overlap and container violation) and retired
 * when within tolerance; retired instances drop out of the remaining
 * iterations.  No reference counterpart: the reference runs one graph per
 * process (core/paradmm.c:736-741); its single-block GPU loop
 * (core/paradmm-gpu.cu:177-281) is the closest relative.  At most 1024
 * graphs; the graphs must outlive the batch. */
typedef struct paradmm_batch paradmm_batch;
typedef struct {
    int tile;            /* pair-tile side: 32 or 64 circles */
    int blocks_per_cu;   /* persistent workgroups per compute unit */
    int barrier;         /* grid barrier: 0 cooperative-groups sync, 1 two-level (bounded wait) */
    double ms_per_iter;  /* measured by the autotuner */
    int contact_tile;    /* 0: the per-edge kernel; 32 or 64: the contact-state kernel with that contact
                            tile side (blocks_per_cu and barrier are then its own) */
} paradmm_batch_config;

int paradmm_batch_create(paradmm_batch **out, graph *const *graphs, int count);
void paradmm_batch_destroy(paradmm_batch *b);
/* Fixes the tile side, grid and barrier (default on first solve: tile 64,
 * 1 block/CU, two-level barrier). */
int paradmm_batch_configure(paradmm_batch *b, int tile, int blocks_per_cu, int barrier);
/* The plain fixed-rho batch (alpha = 1 and one rho on every edge of each
 * instance, no policy) runs in the contact state of the dense layout: per
 * instance only the pairs whose prox is active keep per-edge values, and
 * each instance's iterations equal the single-instance contact iterations.
 * Either one cooperative launch iterates every instance (blocks_per_cu >=
 * 1 workgroups per CU, one grid barrier per iteration) or, blocks_per_cu =
 * 0, one launch per iteration does (the hardware keeps as many work items
 * resident as fit; barrier ignored).  This fixes the contact tile side (32
 * or 64; 0 keeps the per-edge kernel for every solve), the form and the
 * barrier. */
int paradmm_batch_configure_contact(paradmm_batch *b, int contact_tile, int blocks_per_cu, int barrier);
/* The contact kernel's configuration (blocks_per_cu 0: one launch per
 * iteration), its contact tiles over all instances, the pairs listed by
 * every instance's last iteration of the last contact solve (synchronises),
 * and phase_ms[4] of that solve: the after-upload pass, the entry, the
 * remaining iterations, the exit (with the pending U update). */
int paradmm_batch_contact_info(paradmm_batch *b, int *contact_tile, int *blocks_per_cu, int *grid,
                               long *contact_tiles, long *listed_pairs, double *phase_ms);
/* Times every candidate and configures the fastest; *chosen gets it,
 * tried[0..max_tried) the candidates measured.  Where the contact state
 * applies the candidates are its kernel's (contact tile 32/64 x blocks per
 * CU x barrier kind, `iters` iterations from a fresh upload each); else the
 * per-edge kernel's (tile 32/64 x blocks per CU up to the occupancy x
 * barrier kind, `iters` iterations). */
int paradmm_batch_autotune(paradmm_batch *b, int iters, paradmm_batch_config *chosen,
                           paradmm_batch_config *tried, int max_tried, int *ntried);
/* Uploads every graph's host state (X, U, N, Z, RHOS, ALPHAS) and solves;
 * res[i] per instance (iterations at retirement, residuals and quality of the
 * last check, elapsed_ms from the launch start to its retirement, device
 * clock); *kernel_ms the whole launch (with a rho policy: the launches and
 * the policy steps). */
int paradmm_batch_solve(paradmm_batch *b, const paradmm_solve_opts *opts, paradmm_solve_result *res,
                        double *kernel_ms);
/* Learned rho on the batch (m = NULL detaches): with opts->policy_every > 0
 * (and check_every == policy_every) the batch runs policy_every iterations
 * per cooperative launch; after each, every active instance is checked and
 * then takes one rho, 10^(mean of the network's log10 rho over its edges),
 * with U rescaled by rho_old / rho_new -- the per-graph mode of
 * paradmm_gpu_set_rho_policy_mode applied per instance. */
int paradmm_batch_set_rho_policy(paradmm_batch *b, paradmm_rho_mlp *m);
/* 2 (default): one rho per instance, the mean of the network's estimates;
 * 3: that rho plus, per edge, the network's own rho where it exceeds the
 * instance's by more than 10^0.5; average = 1: each instance's rho is the
 * running mean of its estimates (paradmm_gpu_set_rho_policy_mode, per
 * instance) */
int paradmm_batch_set_rho_policy_mode(paradmm_batch *b, int mode, int average);
/* Writes Z, X, U, RHOS (and M = X + U, N = Z - U) back into every graph. */
int paradmm_batch_sync_to_host(paradmm_batch *b);
int paradmm_batch_info(paradmm_batch *b, int *tile, int *blocks_per_cu, int *grid, long *slots, int *tiles);

#ifdef __cplusplus
}
#endif
#endif
