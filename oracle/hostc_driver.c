/*
 * hostc_driver.c -- links the reference demo packing/host.c, compiled
 * UNCHANGED against include/paradmm.h, with libparadmm (the GPU engine).
 * TEST INFRASTRUCTURE ONLY; built into oracle/_ref/ where the reference
 * sources exist.
 *
 * host.c's own main() is renamed at compile time; this main calls its two
 * demo functions (packing/host.c:137, 256) at a chosen size.  Link-time
 * wrapping (-Wl,--wrap) observes the run without touching host.c:
 *   initialize_X_N_Z_M_U_rand  -> real call, then the same draws from
 *                                 splitmix64(seed) instead of time-seeded rand()
 *   updateUN / updateNOpenMP   -> real call, then Z is recorded
 * The executable is linked with -rdynamic so that the engine identifies
 * host.c's prox callbacks by name.
 *
 * usage: hostc_link <serial|omp> <ncircles> <iters> <seed> <out.bin>
 * out.bin: iters x (num_vars*num_dims) doubles, Z after every iteration.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "paradmm.h"

void test_circlepacking_variableradius_triangle_auto(int num_balls, int maxiter);
void test_circlepacking_variableradius_triangle_auto_with_OpenMP(int num_balls, int maxiter, int nu_threads);

void __real_initialize_X_N_Z_M_U_rand(graph *G, double X_l, double X_u, double N_l, double N_u, double Z_l, double Z_u,
                                      double M_l, double M_u, double U_l, double U_u);
void __real_updateUN(graph *G);
void __real_updateNOpenMP(graph *G);

static uint64_t g_seed;
static double *g_hist;
static int g_iters, g_rec;

static double draw(uint64_t *s, double lo, double hi)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return lo + (hi - lo) * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
}

void __wrap_initialize_X_N_Z_M_U_rand(graph *G, double X_l, double X_u, double N_l, double N_u, double Z_l,
                                      double Z_u, double M_l, double M_u, double U_l, double U_u)
{
    __real_initialize_X_N_Z_M_U_rand(G, X_l, X_u, N_l, N_u, Z_l, Z_u, M_l, M_u, U_l, U_u);
    uint64_t s = g_seed;
    size_t l1 = (size_t)G->num_dims * G->num_edges, l2 = (size_t)G->num_dims * G->num_vars;
    for (size_t i = 0; i < l1; i++) {
        G->X[i] = draw(&s, X_l, X_u);
        G->N[i] = draw(&s, N_l, N_u);
        G->M[i] = draw(&s, M_l, M_u);
        G->U[i] = draw(&s, U_l, U_u);
    }
    for (size_t i = 0; i < l2; i++) G->Z[i] = draw(&s, Z_l, Z_u);
}

static void record(graph *G)
{
    size_t lv = (size_t)G->num_dims * G->num_vars;
    if (!g_hist) g_hist = (double *)calloc((size_t)g_iters * lv, sizeof(double));
    if (g_rec < g_iters) memcpy(g_hist + (size_t)g_rec * lv, G->Z, lv * sizeof(double));
    g_rec++;
}

void __wrap_updateUN(graph *G) { __real_updateUN(G); record(G); }
void __wrap_updateNOpenMP(graph *G) { __real_updateNOpenMP(G); record(G); }

int main(int argc, char **argv)
{
    if (argc != 6) {
        fprintf(stderr, "usage: %s <serial|omp> <ncircles> <iters> <seed> <out.bin>\n", argv[0]);
        return 2;
    }
    int nc = atoi(argv[2]);
    g_iters = atoi(argv[3]);
    g_seed = strtoull(argv[4], NULL, 10);
    if (!strcmp(argv[1], "serial")) test_circlepacking_variableradius_triangle_auto(nc, g_iters);
    else test_circlepacking_variableradius_triangle_auto_with_OpenMP(nc, g_iters, 4);
    if (g_rec != g_iters || !g_hist) {
        fprintf(stderr, "recorded %d of %d iterations\n", g_rec, g_iters);
        return 1;
    }
    FILE *f = fopen(argv[5], "wb");
    if (!f) return 1;
    fwrite(g_hist, sizeof(double), (size_t)g_iters * 2 * nc * 2, f);
    fclose(f);
    printf("hostc_link: %s N=%d iters=%d recorded\n", argv[1], nc, g_iters);
    return 0;
}
