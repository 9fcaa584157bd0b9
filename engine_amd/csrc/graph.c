/*
 * graph.c -- the factor-graph builder and initialisers of the C API
 * (core/paradmm.c:21-375 in the reference), host side, C99.
 *
 * Same data structure and edge numbering as the reference: a factor takes
 * the next deg consecutive edge ids in sorted-variable order, each variable
 * lists its edges in creation order, edgeix_to_varlistix maps edge -> var.
 * Differences, all invisible through the API:
 *   - arrays grow geometrically (capacity is a function of the count), not by
 *     one element per call, so 12.5M factors build in about a second;
 *   - every factor's edge list points into one shared identity table instead
 *     of one malloc per factor (the lists are always [s, s+1, ..., s+deg-1]);
 *   - the caller's variable list is sorted in a copy, not in place;
 *   - startG zeroes every pointer, so closeG never frees garbage.
 */
#define _DEFAULT_SOURCE
#include "paradmm.h"
#include "engine_internal.h"
#include "paradmm_gpu.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------ utilities */

double abs_d(double x)
{
    /* core/paradmm.c:21-25 */
    return x >= 0 ? x : -x;
}

static void sort_ints(int *a, int n)
{
    /* insertion sort for short lists, median-of-three quicksort otherwise */
    while (n > 16) {
        int m = n / 2, hi = n - 1, t;
        if (a[m] < a[0]) { t = a[m]; a[m] = a[0]; a[0] = t; }
        if (a[hi] < a[0]) { t = a[hi]; a[hi] = a[0]; a[0] = t; }
        if (a[hi] < a[m]) { t = a[hi]; a[hi] = a[m]; a[m] = t; }
        int pivot = a[m], i = 0, j = hi;
        while (i <= j) {
            while (a[i] < pivot) i++;
            while (a[j] > pivot) j--;
            if (i <= j) { t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
        }
        if (j + 1 < n - i) { sort_ints(a, j + 1); a += i; n -= i; }
        else { sort_ints(a + i, n - i); n = j + 1; }
    }
    for (int i = 1; i < n; i++) {
        int v = a[i], j = i - 1;
        while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
        a[j + 1] = v;
    }
}

void quicksort(int *array, int n)
{
    /* core/paradmm.c:28-60: sorts ascending in place (deterministic pivot here) */
    if (array && n > 1) sort_ints(array, n);
}

double get_random(double lowerbound, double upperbound)
{
    /* core/paradmm.c:63-66 */
    return ((double)rand() / RAND_MAX) * (upperbound - lowerbound) + lowerbound;
}

void *realloc2(void *ptr, size_t sizesbefore, size_t sizesnow)
{
    /* core/paradmm.c:69-73 */
    if (sizesbefore == 0) return malloc(sizesnow);
    return realloc(ptr, sizesnow);
}

static void die_oom(void)
{
    /* the reference's allocation-failure behaviour (core/paradmm.c:196-200) */
    printf("Error allocating memory!\n");
    exit(EXIT_FAILURE);
}

static void *grow(void *p, size_t bytes)
{
    void *q = realloc(p, bytes ? bytes : 1);
    if (!q) die_oom();
    return q;
}

/* capacity held by an array of `count` elements: 16, then powers of two */
static size_t cap_of(size_t count)
{
    size_t c = 16;
    while (c < count) c <<= 1;
    return c;
}

/* ------------------------------------------------------------ graph */

void startG(graph *G, int numdims)
{
    /* core/paradmm.c:79-85, with every pointer cleared */
    memset(G, 0, sizeof(*G));
    G->num_dims = numdims;
}

void closeG(graph *G)
{
    /* core/paradmm.c:91-120: frees what the library allocated; operator
     * parameters belong to the caller.  Device state goes too. */
    pdm_release_graph(G);
    free(G->X); free(G->M); free(G->N); free(G->U); free(G->Z);
    free(G->RHOS); free(G->ALPHAS);
    if (G->num_opts > 0 && G->operatorix_to_edgelistix) free(G->operatorix_to_edgelistix[0]); /* shared table */
    free(G->operatorix_to_edgelistix);
    free(G->list_of_operators_param_pointers);
    free(G->list_of_operators_param_sizes);
    free(G->operatorix_to_numedges);
    free(G->list_of_operators_functions_pointers);
    for (int v = 0; v < G->num_vars; v++) free(G->varix_to_edgelistix[v]);
    free(G->varix_to_edgelistix);
    free(G->varix_to_numedges);
    free(G->edgeix_to_varlistix);
    int d = G->num_dims;
    memset(G, 0, sizeof(*G));
    G->num_dims = d;
}

void printG(graph *G)
{
    /* core/paradmm.c:123-182 (same listing) */
    printf("number of operator in G is %d \n\n", G->num_opts);
    printf("number of global variables in G is %d \n\n", G->num_vars);
    printf("Dimension of each variable in G is %d \n\n", G->num_dims);
    printf("number of edges in G is %d \n\n", G->num_edges);
    printf("RHOS \n");
    for (int i = 0; G->RHOS && i < G->num_edges; i++) printf("%f \n", G->RHOS[i]);
    printf("\nALPHAS \n");
    for (int i = 0; G->ALPHAS && i < G->num_edges; i++) printf("%f\n", G->ALPHAS[i]);
    printf("\nList of operators_pointers \n");
    for (int i = 0; i < G->num_opts; i++) printf("%p\n", (void *)G->list_of_operators_functions_pointers[i]);
    printf("\nList of operatorix_to_numedges \n");
    for (int i = 0; i < G->num_opts; i++) printf("%d\n", G->operatorix_to_numedges[i]);
    printf("\nList of operatorix_to_edgelistin \n");
    for (int i = 0; i < G->num_opts; i++) {
        printf("edge list of operator %d : ", i);
        for (int j = 0; j < G->operatorix_to_numedges[i]; j++) printf("\t%d", G->operatorix_to_edgelistix[i][j]);
        printf("\n");
    }
    printf("\nList of varix_to_numedges \n");
    for (int i = 0; i < G->num_vars; i++) printf("%d\n", G->varix_to_numedges[i]);
    printf("\nList of varix_to_edgelistix \n");
    for (int i = 0; i < G->num_vars; i++) {
        printf("edge list of variable %d : ", i);
        for (int j = 0; j < G->varix_to_numedges[i]; j++) printf("\t%d", G->varix_to_edgelistix[i][j]);
        printf("\n");
    }
    printf("\nList of edgeix_to_varlistix: \n");
    for (int j = 0; j < G->num_edges; j++) printf("%d ", G->edgeix_to_varlistix[j]);
    printf("\n");
}

static void add_factor(graph *G, void (*op)(double *, double *, double *, int, int, void *), void *params,
                       int psize, int nvars, const int *listvarix)
{
    /* addNode / addNodeWithParamSize (core/paradmm.c:381-466, 470-558) */
    if (nvars <= 0) return;
    int small[16];
    int *vars = nvars <= 16 ? small : (int *)malloc(sizeof(int) * nvars);
    if (!vars) die_oom();
    memcpy(vars, listvarix, sizeof(int) * nvars);
    sort_ints(vars, nvars);

    const size_t nf = (size_t)G->num_opts, ne = (size_t)G->num_edges;

    /* per-factor arrays */
    if (nf + 1 > cap_of(nf) || nf == 0) {
        size_t c = cap_of(nf + 1);
        G->list_of_operators_functions_pointers = grow(G->list_of_operators_functions_pointers, c * sizeof(*G->list_of_operators_functions_pointers));
        G->list_of_operators_param_pointers = grow(G->list_of_operators_param_pointers, c * sizeof(void *));
        G->list_of_operators_param_sizes = grow(G->list_of_operators_param_sizes, c * sizeof(int));
        G->operatorix_to_numedges = grow(G->operatorix_to_numedges, c * sizeof(int));
        G->operatorix_to_edgelistix = grow(G->operatorix_to_edgelistix, c * sizeof(int *));
    }

    /* shared identity table of edge ids; factor 0's list is its start, every
     * factor's list is a window of it (factors own consecutive edges) */
    int *tab = nf > 0 ? G->operatorix_to_edgelistix[0] : NULL;
    if (ne == 0 || ne + nvars > cap_of(ne)) {
        int *old = tab;
        size_t c = cap_of(ne + nvars);
        tab = grow(old, c * sizeof(int));
        for (size_t e = ne; e < c; e++) tab[e] = (int)e;
        if (tab != old) {
            size_t s = 0;
            for (size_t f = 0; f < nf; f++) {
                G->operatorix_to_edgelistix[f] = tab + s;
                s += (size_t)G->operatorix_to_numedges[f];
            }
        }
        G->edgeix_to_varlistix = grow(G->edgeix_to_varlistix, c * sizeof(int));
    }

    G->list_of_operators_functions_pointers[nf] = op;
    G->list_of_operators_param_pointers[nf] = params;
    G->list_of_operators_param_sizes[nf] = psize;
    G->operatorix_to_numedges[nf] = nvars;
    G->operatorix_to_edgelistix[nf] = tab + ne;
    G->num_opts = (int)nf + 1;

    /* variables */
    int maxv = vars[nvars - 1];
    if (maxv + 1 > G->num_vars) {
        size_t nvold = (size_t)G->num_vars, nvnew = (size_t)maxv + 1;
        if (nvnew > cap_of(nvold) || nvold == 0) {
            size_t c = cap_of(nvnew);
            G->varix_to_numedges = grow(G->varix_to_numedges, c * sizeof(int));
            G->varix_to_edgelistix = grow(G->varix_to_edgelistix, c * sizeof(int *));
        }
        for (size_t v = nvold; v < nvnew; v++) {
            G->varix_to_numedges[v] = 0;
            G->varix_to_edgelistix[v] = NULL;
        }
        G->num_vars = (int)nvnew;
    }
    for (int i = 0; i < nvars; i++) {
        int v = vars[i];
        int e = (int)ne + i;
        size_t c = (size_t)G->varix_to_numedges[v];
        if (c == 0 || c + 1 > cap_of(c)) {
            G->varix_to_edgelistix[v] = grow(G->varix_to_edgelistix[v], cap_of(c + 1) * sizeof(int));
        }
        G->varix_to_edgelistix[v][c] = e;
        G->varix_to_numedges[v] = (int)c + 1;
        G->edgeix_to_varlistix[e] = v;
    }
    G->num_edges = (int)ne + nvars;
    if (vars != small) free(vars);
}

void addNode(graph *G, void (*op)(double *, double *, double *, int, int, void *), void *operator_para,
             int opt_numvars, int *listvarix)
{
    add_factor(G, op, operator_para, 0, opt_numvars, listvarix);
}

void addNodeWithParamSize(graph *G, void (*op)(double *, double *, double *, int, int, void *), void *operator_para,
                          int operator_para_size, int opt_numvars, int *listvarix)
{
    add_factor(G, op, operator_para, operator_para_size, opt_numvars, listvarix);
}

/* ------------------------------------------------------------ initialisers */

void initialize_RHOS_APHAS(graph *G, double rho0, double alpha0)
{
    /* core/paradmm.c:188-207 */
    size_t n = (size_t)G->num_edges;
    free(G->RHOS); free(G->ALPHAS);
    G->RHOS = (double *)malloc((n ? n : 1) * sizeof(double));
    G->ALPHAS = (double *)malloc((n ? n : 1) * sizeof(double));
    if (!G->RHOS || !G->ALPHAS) die_oom();
    for (size_t i = 0; i < n; i++) { G->RHOS[i] = rho0; G->ALPHAS[i] = alpha0; }
    pdm_host_written(G);
}

static void alloc_state(graph *G, int zero)
{
    size_t l1 = (size_t)G->num_dims * G->num_edges, l2 = (size_t)G->num_dims * G->num_vars;
    free(G->X); free(G->N); free(G->M); free(G->U); free(G->Z);
    if (zero) {
        G->X = (double *)calloc(l1 ? l1 : 1, sizeof(double));
        G->N = (double *)calloc(l1 ? l1 : 1, sizeof(double));
        G->M = (double *)calloc(l1 ? l1 : 1, sizeof(double));
        G->U = (double *)calloc(l1 ? l1 : 1, sizeof(double));
        G->Z = (double *)calloc(l2 ? l2 : 1, sizeof(double));
    } else {
        G->X = (double *)malloc((l1 ? l1 : 1) * sizeof(double));
        G->N = (double *)malloc((l1 ? l1 : 1) * sizeof(double));
        G->M = (double *)malloc((l1 ? l1 : 1) * sizeof(double));
        G->U = (double *)malloc((l1 ? l1 : 1) * sizeof(double));
        G->Z = (double *)malloc((l2 ? l2 : 1) * sizeof(double));
    }
    if (!G->X || !G->N || !G->M || !G->U || !G->Z) die_oom();
}

void initialize_X_N_Z_M_U_zeros(graph *G)
{
    /* core/paradmm.c:210-226 */
    alloc_state(G, 1);
    pdm_host_written(G);
}

void initialize_X_N_Z_M_U_rand(graph *G, double X_l, double X_u, double N_l, double N_u, double Z_l, double Z_u,
                               double M_l, double M_u, double U_l, double U_u)
{
    /* core/paradmm.c:229-262: time-seeded rand(), X/N/M/U drawn per entry in
     * that order, then Z. */
    alloc_state(G, 0);
    srand((unsigned)time(NULL));
    size_t l1 = (size_t)G->num_dims * G->num_edges, l2 = (size_t)G->num_dims * G->num_vars;
    for (size_t i = 0; i < l1; i++) {
        G->X[i] = get_random(X_l, X_u);
        G->N[i] = get_random(N_l, N_u);
        G->M[i] = get_random(M_l, M_u);
        G->U[i] = get_random(U_l, U_u);
    }
    for (size_t i = 0; i < l2; i++) G->Z[i] = get_random(Z_l, Z_u);
    pdm_host_written(G);
}

int paradmm_init_rand_seeded(graph *G, uint64_t seed, double X_l, double X_u, double N_l, double N_u, double Z_l,
                             double Z_u, double M_l, double M_u, double U_l, double U_u)
{
    /* initialize_X_N_Z_M_U_rand with the same draws in the same order (per
     * entry X, N, M, U, then Z), from a splitmix64 stream of the given seed
     * instead of the time-seeded rand(): reproducible runs */
    if (!G || G->num_edges <= 0 || G->num_vars <= 0) {
        pdm_set_error(PARADMM_ERR_STATE, "paradmm_init_rand_seeded: need a built graph");
        return PARADMM_ERR_STATE;
    }
    alloc_state(G, 0);
    uint64_t st = seed;
    size_t l1 = (size_t)G->num_dims * G->num_edges, l2 = (size_t)G->num_dims * G->num_vars;
    for (size_t i = 0; i < l1; i++) {
        G->X[i] = X_l + (X_u - X_l) * pdm_uniform01(&st);
        G->N[i] = N_l + (N_u - N_l) * pdm_uniform01(&st);
        G->M[i] = M_l + (M_u - M_l) * pdm_uniform01(&st);
        G->U[i] = U_l + (U_u - U_l) * pdm_uniform01(&st);
    }
    for (size_t i = 0; i < l2; i++) G->Z[i] = Z_l + (Z_u - Z_l) * pdm_uniform01(&st);
    pdm_host_written(G);
    return PARADMM_OK;
}

void initialize_X_N_Z_M_U_optimal(graph *G, double Zinit[], double Uinit[], int maxiter)
{
    /* core/paradmm.c:266-375: Z = Zinit, X = Z of each edge's variable, then
     * maxiter prox sweeps per factor on N = X - U.  The sweeps run on the GPU
     * (the callbacks are device kinds).  With maxiter = 0 the reference leaves U
     * uninitialised; here U = Uinit. */
    alloc_state(G, 0);
    const int d = G->num_dims;
    size_t l1 = (size_t)d * G->num_edges, l2 = (size_t)d * G->num_vars;
    for (size_t i = 0; i < l2; i++) G->Z[i] = Zinit[i];
    for (int e = 0; e < G->num_edges; e++) {
        int v = G->edgeix_to_varlistix[e];
        for (int k = 0; k < d; k++) G->X[(size_t)d * e + k] = Zinit[(size_t)d * v + k];
    }
    for (size_t i = 0; i < l1; i++) {
        G->U[i] = Uinit[i];
        G->N[i] = G->X[i] - Uinit[i];
        G->M[i] = G->X[i] + Uinit[i];
    }
    pdm_host_written(G);
    if (maxiter > 0) pdm_legacy_warm_start(G, maxiter);
}

/* ------------------------------------------------------------ bulk construction */

int paradmm_add_nodes(graph *G, void (*op)(double *, double *, double *, int, int, void *), void *params,
                      size_t param_stride, int count, int nvars, const int *vars)
{
    /* `count` factors of one operator and degree: factor i has variables
     * vars[i*nvars .. i*nvars+nvars) and parameters params + i*param_stride
     * bytes (param_stride 0: all share params).  Same edge numbering as
     * count addNode calls (core/paradmm.c:381-466). */
    if (!G || !op || count < 0 || nvars < 1 || (count > 0 && !vars)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_add_nodes: bad arguments");
        return PARADMM_ERR_ARG;
    }
    for (int i = 0; i < count; i++)
        for (int k = 0; k < nvars; k++)
            if (vars[(size_t)i * nvars + k] < 0) {
                pdm_set_error(PARADMM_ERR_ARG, "paradmm_add_nodes: factor %d has a negative variable index", i);
                return PARADMM_ERR_ARG;
            }
    for (int i = 0; i < count; i++)
        add_factor(G, op, params ? (char *)params + (size_t)i * param_stride : NULL, 0, nvars,
                   vars + (size_t)i * nvars);
    return PARADMM_OK;
}
