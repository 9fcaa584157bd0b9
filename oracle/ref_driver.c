/*
 * ref_driver.c -- runs the REFERENCE engine (core/paradmm.c compiled from its
 * own source into oracle/_ref/) so that the oracle can be pinned against it.
 * TEST INFRASTRUCTURE ONLY; built only where the reference sources exist.
 *
 * Two problems:
 *   ref_square_run   -- unit-square dim-3 packing, reference engine
 *                       (addNode, updateXM/updateZ/updateUN) with the
 *                       oracle's dim-3 prox operators and seeded init;
 *   ref_triangle_run -- the reference demo's triangle packing
 *                       (packing/host.c:256-343) with the reference's own prox
 *                       operators from packing/host.c, seeded init.
 * Both record Z after every iteration.  The *_omp variants run the same
 * problems through the reference's OpenMP passes instead (the five
 * update*OpenMP calls of packing/host.c:237-244, or the one-region
 * updateGOpenMP of core/paradmm.c:949-1092), and the *_warm_start entry
 * points run initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375) on them.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include "paradmm.h"          /* the reference header (core/paradmm.h) */
#include "paradmm_oracle.h"   /* seeded generator and dim-3 operators */

/* the reference demo's operators (packing/host.c:31, 67, 83) */
void disk_plane_collision_variable_radius(double *x, double *n, double *rhos, int numvars, int numdims, void *params);
void maximize_sum_of_squared_disk_radii_alpha(double *x, double *n, double *rhos, int numvars, int numdims, void *params);
void disk_collision_variable_radius(double *x, double *n, double *rhos, int numvars, int numdims, void *params);

/* The unit-square (dim-3) graph through the reference builder with the
 * oracle's dim-3 operators.  *a must outlive the graph (the factors keep a
 * pointer to it). */
static void square_graph(graph *G, int nc, double *a)
{
    memset(G, 0, sizeof(*G));
    startG(G, 3);
    for (int i = 0; i < nc; i++) { int v = i; addNode(G, or_prox_radius3, a, 1, &v); }
    for (int i = 0; i < nc; i++) { int v = i; addNode(G, or_prox_box3, NULL, 1, &v); }
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[2] = {i, j};
            addNode(G, or_prox_pair3, NULL, 2, vv);
        }
}

/* The seeded unit-square state of or_build_square: Z drawn per circle,
 * X = N = M = Z of the edge's variable, U = 0. */
static void square_state(graph *G, int nc, uint64_t seed, double rho, double alpha)
{
    initialize_X_N_Z_M_U_zeros(G);
    initialize_RHOS_APHAS(G, rho, alpha);
    uint64_t s = seed;
    for (int i = 0; i < nc; i++) {
        G->Z[3 * i + 0] = or_uniform01(&s);
        G->Z[3 * i + 1] = or_uniform01(&s);
        G->Z[3 * i + 2] = 0.05 * or_uniform01(&s);
    }
    for (int e = 0; e < G->num_edges; e++) {
        int v = G->edgeix_to_varlistix[e];
        for (int k = 0; k < 3; k++) {
            double z = G->Z[3 * v + k];
            G->X[3 * e + k] = z; G->N[3 * e + k] = z; G->M[3 * e + k] = z; G->U[3 * e + k] = 0;
        }
    }
}

/* The reference demo's triangle graph (packing/host.c:256-343) with its own
 * operators; walls and a must outlive the graph. */
static void triangle_graph(graph *G, int nc, double walls[3][4], double *a)
{
    memset(G, 0, sizeof(*G));
    startG(G, 2);
    for (int i = 0; i < nc; i++) { int v = 2 * i + 1; addNode(G, maximize_sum_of_squared_disk_radii_alpha, a, 1, &v); }
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < nc; j++) {
            int vv[2] = {2 * j, 2 * j + 1};
            addNode(G, disk_plane_collision_variable_radius, walls[k], 2, vv);
        }
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[4] = {2 * i, 2 * i + 1, 2 * j, 2 * j + 1};
            addNode(G, disk_collision_variable_radius, NULL, 4, vv);
        }
}

static void triangle_walls(double walls[3][4])
{
    const double deg = 3.14159265358979323846 / 3.0;
    double w[3][4] = {{0., -1., 0., 0.}, {-sin(deg), cos(deg), 0., 0.}, {sin(deg), cos(deg), 1., 0.}};
    memcpy(walls, w, sizeof(w));
}

/* The seeded triangle state of or_build_triangle (the interleaving of
 * initialize_X_N_Z_M_U_rand, core/paradmm.c:247-261). */
static void triangle_state(graph *G, int nc, uint64_t seed, double lo, double hi)
{
    initialize_X_N_Z_M_U_zeros(G);
    initialize_RHOS_APHAS(G, 10.0, 0.1);
    for (int i = 0; i < nc; i++) G->RHOS[i] = 5.0;
    uint64_t s = seed;
    size_t le = (size_t)2 * G->num_edges, lv = (size_t)2 * G->num_vars;
    for (size_t i = 0; i < le; i++) {
        G->X[i] = lo + (hi - lo) * or_uniform01(&s);
        G->N[i] = lo + (hi - lo) * or_uniform01(&s);
        G->M[i] = lo + (hi - lo) * or_uniform01(&s);
        G->U[i] = lo + (hi - lo) * or_uniform01(&s);
    }
    for (size_t i = 0; i < lv; i++) G->Z[i] = lo + (hi - lo) * or_uniform01(&s);
}

/* mode 0: serial updateXM/updateZ/updateUN per iteration (packing/host.c:355-360);
 * mode 1: the five update*OpenMP calls (packing/host.c:237-244) on nthreads;
 * mode 2: updateGOpenMP(G, iters, nthreads), Z recorded only at the end. */
static void run_mode(graph *G, int iters, int mode, int nthreads, double *zhist)
{
    size_t lv = (size_t)G->num_dims * G->num_vars;
    if (mode == 2) {
        updateGOpenMP(G, iters, nthreads);
        if (zhist) memcpy(zhist, G->Z, lv * sizeof(double));
        return;
    }
    if (mode == 1) omp_set_num_threads(nthreads);
    for (int it = 0; it < iters; it++) {
        if (mode == 1) {
            updateXOpenMP(G);
            updateMOpenMP(G);
            updateZOpenMP(G);
            updateUOpenMP(G);
            updateNOpenMP(G);
        } else {
            updateXM(G);
            updateZ(G);
            updateUN(G);
        }
        if (zhist) memcpy(zhist + (size_t)it * lv, G->Z, lv * sizeof(double));
    }
}

int ref_square_run(int nc, uint64_t seed, double rho, double alpha, double a_exp, int iters, double *zhist)
{
    graph G;
    double a = a_exp;
    square_graph(&G, nc, &a);
    square_state(&G, nc, seed, rho, alpha);
    run_mode(&G, iters, 0, 1, zhist);
    closeG(&G);
    return 0;
}

int ref_square_run_omp(int nc, uint64_t seed, double rho, double alpha, double a_exp, int iters, int mode,
                       int nthreads, double *zhist)
{
    graph G;
    double a = a_exp;
    square_graph(&G, nc, &a);
    square_state(&G, nc, seed, rho, alpha);
    run_mode(&G, iters, mode, nthreads, zhist);
    closeG(&G);
    return 0;
}

int ref_triangle_run(int nc, uint64_t seed, double lo, double hi, int iters, double *zhist)
{
    double walls[3][4], a = 2.0;
    graph G;
    triangle_walls(walls);
    triangle_graph(&G, nc, walls, &a);
    triangle_state(&G, nc, seed, lo, hi);
    run_mode(&G, iters, 0, 1, zhist);
    closeG(&G);
    return 0;
}

int ref_triangle_run_omp(int nc, uint64_t seed, double lo, double hi, int iters, int mode, int nthreads,
                         double *zhist)
{
    double walls[3][4], a = 2.0;
    graph G;
    triangle_walls(walls);
    triangle_graph(&G, nc, walls, &a);
    triangle_state(&G, nc, seed, lo, hi);
    run_mode(&G, iters, mode, nthreads, zhist);
    closeG(&G);
    return 0;
}

static void copy_out(graph *G, double *X, double *M, double *N, double *U, double *Z)
{
    size_t le = (size_t)G->num_dims * G->num_edges, lv = (size_t)G->num_dims * G->num_vars;
    memcpy(X, G->X, le * sizeof(double));
    memcpy(M, G->M, le * sizeof(double));
    memcpy(N, G->N, le * sizeof(double));
    memcpy(U, G->U, le * sizeof(double));
    memcpy(Z, G->Z, lv * sizeof(double));
}

/* initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375) on the unit-square
 * graph: rho and alpha set first (the warm start reads RHOS), then the
 * reference allocates X, N, M, U, Z itself.  maxiter >= 1 (with 0 the
 * reference leaves U unset). */
int ref_square_warm_start(int nc, double rho, double a_exp, const double *Zinit, const double *Uinit, int maxiter,
                          double *X, double *M, double *N, double *U, double *Z)
{
    graph G;
    double a = a_exp;
    square_graph(&G, nc, &a);
    initialize_RHOS_APHAS(&G, rho, 1.0);
    initialize_X_N_Z_M_U_optimal(&G, (double *)Zinit, (double *)Uinit, maxiter);
    copy_out(&G, X, M, N, U, Z);
    closeG(&G);
    return 0;
}

/* the same on the triangle graph (rho 10, 5 on the radius edges, as
 * or_build_triangle) */
int ref_triangle_warm_start(int nc, const double *Zinit, const double *Uinit, int maxiter, double *X, double *M,
                            double *N, double *U, double *Z)
{
    double walls[3][4], a = 2.0;
    graph G;
    triangle_walls(walls);
    triangle_graph(&G, nc, walls, &a);
    initialize_RHOS_APHAS(&G, 10.0, 0.1);
    for (int i = 0; i < nc; i++) G.RHOS[i] = 5.0;
    initialize_X_N_Z_M_U_optimal(&G, (double *)Zinit, (double *)Uinit, maxiter);
    copy_out(&G, X, M, N, U, Z);
    closeG(&G);
    return 0;
}

/* Thin wrappers so the tests can call the reference's own prox operators. */
void ref_prox_pair2(double *x, double *n, double *r, int nv, int nd, void *p) { disk_collision_variable_radius(x, n, r, nv, nd, p); }
void ref_prox_wall2(double *x, double *n, double *r, int nv, int nd, void *p) { disk_plane_collision_variable_radius(x, n, r, nv, nd, p); }
void ref_prox_radius2(double *x, double *n, double *r, int nv, int nd, void *p) { maximize_sum_of_squared_disk_radii_alpha(x, n, r, nv, nd, p); }
