/*
 * packing_ops.c -- the packing factor library of the MI355X engine.
 *
 * The prox operators are device kinds (the X pass runs them on the GPU);
 * the exported functions are the tokens a caller passes to addNode.  They
 * are not host implementations: called on the host they abort.
 *
 * Also the unit-square problem builder and its seeded initialisation.
 */
#include "paradmm.h"
#include "paradmm_gpu.h"
#include "engine_internal.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static void host_call(const char *name)
{
    fprintf(stderr, "paradmm: %s is a device prox operator; it runs inside the GPU X pass and "
                    "has no host implementation\n", name);
    abort();
}

#define TOKEN(fn)                                                                           \
    void fn(double *x, double *n, double *rhos, int nvars, int ndims, void *params)         \
    {                                                                                       \
        (void)x; (void)n; (void)rhos; (void)nvars; (void)ndims; (void)params;               \
        host_call(#fn);                                                                     \
    }

TOKEN(paradmm_prox_disk_collision_2d)
TOKEN(paradmm_prox_disk_plane_2d)
TOKEN(paradmm_prox_max_radius_2d)
TOKEN(paradmm_prox_disk_collision_3d)
TOKEN(paradmm_prox_box_3d)
TOKEN(paradmm_prox_max_radius_3d)

/* Radius exponents used by paradmm_packing_square, interned so the parameter
 * pointer stays valid for the life of the process. */
static double g_exponent[64];
static int g_nexponent;

static double *intern_exponent(double a)
{
    for (int i = 0; i < g_nexponent; i++)
        if (g_exponent[i] == a) return &g_exponent[i];
    if (g_nexponent == 64) return NULL;
    g_exponent[g_nexponent] = a;
    return &g_exponent[g_nexponent++];
}

int paradmm_packing_square(graph *G, int ncircles, double a_exp)
{
    if (!G || ncircles < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_packing_square: bad arguments");
        return PARADMM_ERR_ARG;
    }
    if (G->num_dims != 3) {
        pdm_set_error(PARADMM_ERR_SHAPE, "paradmm_packing_square: graph num_dims is %d, need 3", G->num_dims);
        return PARADMM_ERR_SHAPE;
    }
    double *a = intern_exponent(a_exp);
    if (!a) {
        pdm_set_error(PARADMM_ERR_NOMEM, "paradmm_packing_square: too many distinct exponents");
        return PARADMM_ERR_NOMEM;
    }
    /* factor order of packing/host.c:298-335: radius, container, pairs (i-major) */
    for (int i = 0; i < ncircles; i++) { int v = i; addNode(G, paradmm_prox_max_radius_3d, a, 1, &v); }
    for (int i = 0; i < ncircles; i++) { int v = i; addNode(G, paradmm_prox_box_3d, NULL, 1, &v); }
    for (int i = 0; i < ncircles - 1; i++)
        for (int j = i + 1; j < ncircles; j++) {
            int vv[2] = {i, j};
            addNode(G, paradmm_prox_disk_collision_3d, NULL, 2, vv);
        }
    return PARADMM_OK;
}

static uint64_t splitmix64(uint64_t *s)
{
    /* splitmix64 (Steele, Lea, Flood 2014; constants as published by Vigna) */
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double unit_draw(uint64_t *s)
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

double pdm_uniform01(uint64_t *state) { return unit_draw(state); }

int paradmm_packing_init_square(graph *G, uint64_t seed)
{
    if (!G || G->num_dims != 3 || G->num_edges <= 0) {
        pdm_set_error(PARADMM_ERR_STATE, "paradmm_packing_init_square: need a built dim-3 graph");
        return PARADMM_ERR_STATE;
    }
    initialize_X_N_Z_M_U_zeros(G);
    uint64_t s = seed;
    for (int i = 0; i < G->num_vars; i++) {
        G->Z[3 * i + 0] = unit_draw(&s);
        G->Z[3 * i + 1] = unit_draw(&s);
        G->Z[3 * i + 2] = 0.05 * unit_draw(&s);
    }
    for (int e = 0; e < G->num_edges; e++) {
        int v = G->edgeix_to_varlistix[e];
        for (int k = 0; k < 3; k++) {
            double z = G->Z[3 * v + k];
            G->X[3 * e + k] = z;
            G->N[3 * e + k] = z;
            G->M[3 * e + k] = z;
            G->U[3 * e + k] = 0.0;
        }
    }
    pdm_host_written(G);
    return PARADMM_OK;
}
