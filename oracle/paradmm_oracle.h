/*
 * paradmm_oracle.h -- FP64 CPU oracle for the parADMM message-passing iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP engine is
 * compared against; it is never linked into libparadmm and nothing in the
 * product calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * It restates, in double precision and in the reference's loop order:
 *   - the factor-graph builder            core/paradmm.c:381-466 (addNode)
 *   - the serial X/M/Z/U/N passes          core/paradmm.c:569-741
 *   - the OpenMP X/M/Z/U/N passes          core/paradmm.c:812-945
 *   - the packing prox operators (dim 2)   packing/host.c:10-132
 * and defines the pieces the reference does not have (seeded init, dim-3
 * unit-square prox operators, residuals, feasibility), each pinned by tests
 * against closed forms, brute force or the reference compiled from source.
 */
#ifndef PARADMM_ORACLE_H
#define PARADMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Callback signature of the reference (core/paradmm.h:41). */
typedef void (*or_prox_fn)(double *x, double *n, double *rhos, int nvars, int ndims, void *params);

typedef struct or_graph {
    int ndims, nfac, nvar, nedge;
    int cap_fac, cap_edge, cap_var;
    or_prox_fn *fac_fn;
    void **fac_params;
    int *fac_deg;
    int *fac_first;      /* first edge of each factor: edges of a factor are contiguous */
    int *var_deg;
    int *var_cap;
    int **var_edges;     /* per-variable edge list in creation order (pointer-of-pointers, as the reference) */
    int *edge_var;
    double *X, *M, *N, *U, *Z, *RHO, *ALPHA, *Zprev;
    double *owned_params; /* parameter storage owned by the packing builders */
    int iter;
} or_graph;

/* ---- builder ---- */
or_graph *or_new(int ndims);
void or_free(or_graph *g);
/* Adds one factor; vars are copied and sorted ascending (the reference sorts the caller's list). */
void or_add_factor(or_graph *g, or_prox_fn fn, void *params, int nvars, const int *vars);
/* Allocates X, M, N, U, Z, RHO, ALPHA, Zprev (zeroed) once all factors are added. */
void or_alloc_state(or_graph *g);
void or_set_rho_alpha(or_graph *g, double rho, double alpha);

/* ---- serial passes (core/paradmm.c:569-741) ---- */
void or_update_XM(or_graph *g);
void or_update_X(or_graph *g);
void or_update_M(or_graph *g);
void or_update_Z(or_graph *g);
void or_update_UN(or_graph *g);
void or_update_U(or_graph *g);
void or_update_N(or_graph *g);
/* iters x (X, M, Z, U, N); keeps Zprev = Z of the previous iteration. zhist (optional)
 * receives Z after every iteration, iters*nvar*ndims doubles. */
void or_iterate(or_graph *g, int iters, double *zhist);

/* ---- OpenMP passes (core/paradmm.c:812-945); used for the timed CPU baseline ---- */
void or_update_X_omp(or_graph *g);
void or_update_M_omp(or_graph *g);
void or_update_Z_omp(or_graph *g);
void or_update_U_omp(or_graph *g);
void or_update_N_omp(or_graph *g);
/* iters x the five OpenMP calls, as the reference's OpenMP driver loop (packing/host.c:237-244).
 * Returns wall seconds measured with gettimeofday around the loop only. */
double or_iterate_omp(or_graph *g, int iters, int nthreads);
/* or_iterate (Zprev kept, Z recorded) through the OpenMP passes on nthreads:
 * the same results, faster, for the checks at large sizes. */
void or_iterate_par(or_graph *g, int iters, int nthreads, double *zhist);

/* ---- packing prox operators ---- */
/* dim-2 operators of the reference demo (packing/host.c) */
double or_max_alpha_radius(double n, double rho, double a);
void or_prox_pair2(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void or_prox_wall2(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void or_prox_radius2(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
/* dim-3 (cx, cy, r) operators for the unit-square packing */
void or_prox_pair3(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void or_prox_box3(double *x, double *n, double *rhos, int nvars, int ndims, void *params);
void or_prox_radius3(double *x, double *n, double *rhos, int nvars, int ndims, void *params);

/* ---- seeded generator (splitmix64, Vigna 2015) ---- */
uint64_t or_splitmix64(uint64_t *state);
double or_uniform01(uint64_t *state);

/* ---- packing problem builders ---- */
/* Unit-square packing, one (cx,cy,r) variable per circle.  Factor order: N radius
 * factors (exponent a), N box factors, N(N-1)/2 pair factors in i-major order.
 * Init: cx,cy ~ U[0,1], r ~ U[0,0.05] from splitmix64(seed); X = Z, U = 0, N = Z, M = X. */
or_graph *or_build_square(int ncircles, uint64_t seed, double rho, double alpha, double a_exp);
/* Triangle packing of the reference demo (packing/host.c:256-343): variables 2i (centre)
 * and 2i+1 (radius), dim 2, N radius factors (a = 2), 3N wall factors, N(N-1)/2 pair
 * factors.  rho = 10 except 5 on the first N edges, alpha = 0.1.  The reference draws the
 * init from time-seeded rand(); here X, N, M, U, Z are drawn from splitmix64(seed) in
 * [lo, hi] in the same interleaving (X,N,M,U per entry, then Z). */
or_graph *or_build_triangle(int ncircles, uint64_t seed, double lo, double hi);

/* ---- warm start: initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375) ---- */
void or_warm_start(or_graph *g, const double *Zinit, const double *Uinit, int maxiter);

/* ---- residuals, feasibility, objective (not in the reference) ---- */
/* primal2 = sum_e ||X_e - Z_v(e)||^2, dual2 = sum_e rho_e^2 ||Z_v - Zprev_v||^2 */
void or_residuals(const or_graph *g, double *primal2, double *dual2);
/* For the dim-3 square packing: max pair overlap (r_i + r_j - |c_i - c_j|)+, max box
 * violation, and objective sum r_i, evaluated on Z. */
void or_square_quality(const or_graph *g, int ncircles, double *max_overlap, double *max_box, double *sum_r);

/* ---- accessors for ctypes ---- */
int or_nvar(const or_graph *g);
int or_nedge(const or_graph *g);
int or_nfac(const or_graph *g);
int or_ndims(const or_graph *g);
double *or_ptr(or_graph *g, int which); /* 0 X 1 M 2 N 3 U 4 Z 5 RHO 6 ALPHA 7 Zprev */
const int *or_edge_var(const or_graph *g);
const int *or_fac_first(const or_graph *g);
const int *or_fac_deg(const or_graph *g);
int or_var_deg(const or_graph *g, int v);
const int *or_var_edges(const or_graph *g, int v);

#ifdef __cplusplus
}
#endif
#endif
