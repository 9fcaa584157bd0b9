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
 * Both record Z after every iteration.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "paradmm.h"          /* the reference header (core/paradmm.h) */
#include "paradmm_oracle.h"   /* seeded generator and dim-3 operators */

/* the reference demo's operators (packing/host.c:31, 67, 83) */
void disk_plane_collision_variable_radius(double *x, double *n, double *rhos, int numvars, int numdims, void *params);
void maximize_sum_of_squared_disk_radii_alpha(double *x, double *n, double *rhos, int numvars, int numdims, void *params);
void disk_collision_variable_radius(double *x, double *n, double *rhos, int numvars, int numdims, void *params);

static void run_and_record(graph *G, int iters, double *zhist)
{
    size_t lv = (size_t)G->num_dims * G->num_vars;
    for (int it = 0; it < iters; it++) {
        updateXM(G);
        updateZ(G);
        updateUN(G);
        if (zhist) memcpy(zhist + (size_t)it * lv, G->Z, lv * sizeof(double));
    }
}

int ref_square_run(int nc, uint64_t seed, double rho, double alpha, double a_exp, int iters, double *zhist)
{
    graph G;
    memset(&G, 0, sizeof(G));
    startG(&G, 3);
    double a = a_exp;
    for (int i = 0; i < nc; i++) { int v = i; addNode(&G, or_prox_radius3, &a, 1, &v); }
    for (int i = 0; i < nc; i++) { int v = i; addNode(&G, or_prox_box3, NULL, 1, &v); }
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[2] = {i, j};
            addNode(&G, or_prox_pair3, NULL, 2, vv);
        }
    initialize_X_N_Z_M_U_zeros(&G);
    initialize_RHOS_APHAS(&G, rho, alpha);
    uint64_t s = seed;
    for (int i = 0; i < nc; i++) {
        G.Z[3 * i + 0] = or_uniform01(&s);
        G.Z[3 * i + 1] = or_uniform01(&s);
        G.Z[3 * i + 2] = 0.05 * or_uniform01(&s);
    }
    for (int e = 0; e < G.num_edges; e++) {
        int v = G.edgeix_to_varlistix[e];
        for (int k = 0; k < 3; k++) {
            double z = G.Z[3 * v + k];
            G.X[3 * e + k] = z; G.N[3 * e + k] = z; G.M[3 * e + k] = z; G.U[3 * e + k] = 0;
        }
    }
    run_and_record(&G, iters, zhist);
    closeG(&G);
    return 0;
}

int ref_triangle_run(int nc, uint64_t seed, double lo, double hi, int iters, double *zhist)
{
    const double deg = 3.14159265358979323846 / 3.0;
    double walls[3][4] = {{0., -1., 0., 0.}, {-sin(deg), cos(deg), 0., 0.}, {sin(deg), cos(deg), 1., 0.}};
    double a = 2.0;
    graph G;
    memset(&G, 0, sizeof(G));
    startG(&G, 2);
    for (int i = 0; i < nc; i++) { int v = 2 * i + 1; addNode(&G, maximize_sum_of_squared_disk_radii_alpha, &a, 1, &v); }
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < nc; j++) {
            int vv[2] = {2 * j, 2 * j + 1};
            addNode(&G, disk_plane_collision_variable_radius, walls[k], 2, vv);
        }
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[4] = {2 * i, 2 * i + 1, 2 * j, 2 * j + 1};
            addNode(&G, disk_collision_variable_radius, NULL, 4, vv);
        }
    initialize_X_N_Z_M_U_zeros(&G);
    initialize_RHOS_APHAS(&G, 10.0, 0.1);
    for (int i = 0; i < nc; i++) G.RHOS[i] = 5.0;
    uint64_t s = seed;
    size_t le = (size_t)2 * G.num_edges, lv = (size_t)2 * G.num_vars;
    for (size_t i = 0; i < le; i++) {
        G.X[i] = lo + (hi - lo) * or_uniform01(&s);
        G.N[i] = lo + (hi - lo) * or_uniform01(&s);
        G.M[i] = lo + (hi - lo) * or_uniform01(&s);
        G.U[i] = lo + (hi - lo) * or_uniform01(&s);
    }
    for (size_t i = 0; i < lv; i++) G.Z[i] = lo + (hi - lo) * or_uniform01(&s);
    run_and_record(&G, iters, zhist);
    closeG(&G);
    return 0;
}

/* Thin wrappers so the tests can call the reference's own prox operators. */
void ref_prox_pair2(double *x, double *n, double *r, int nv, int nd, void *p) { disk_collision_variable_radius(x, n, r, nv, nd, p); }
void ref_prox_wall2(double *x, double *n, double *r, int nv, int nd, void *p) { disk_plane_collision_variable_radius(x, n, r, nv, nd, p); }
void ref_prox_radius2(double *x, double *n, double *r, int nv, int nd, void *p) { maximize_sum_of_squared_disk_radii_alpha(x, n, r, nv, nd, p); }
