/*
 * paradmm_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY, never shipped).
 *
 * Restatement of the reference algorithm in double precision and in the
 * reference's loop order.  Each function cites the reference lines it
 * follows.  Functions for pieces the reference lacks say so and name the
 * test that pins them.
 */
#define _DEFAULT_SOURCE
#include "paradmm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

static void *xrealloc(void *p, size_t n)
{
    void *q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return q;
}

/* ------------------------------------------------------------------ builder */

or_graph *or_new(int ndims)
{
    /* startG (core/paradmm.c:79-85), with every pointer zeroed. */
    or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
    if (!g) abort();
    g->ndims = ndims;
    return g;
}

void or_free(or_graph *g)
{
    if (!g) return;
    free(g->fac_fn); free(g->fac_params); free(g->fac_deg); free(g->fac_first);
    for (int v = 0; v < g->nvar; v++) free(g->var_edges[v]);
    free(g->var_edges); free(g->var_deg); free(g->var_cap); free(g->edge_var);
    free(g->X); free(g->M); free(g->N); free(g->U); free(g->Z);
    free(g->RHO); free(g->ALPHA); free(g->Zprev); free(g->owned_params);
    free(g);
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

void or_add_factor(or_graph *g, or_prox_fn fn, void *params, int nvars, const int *vars_in)
{
    /* addNode (core/paradmm.c:381-466): the variable list is sorted ascending
     * (paradmm.c:383), the factor takes the next nvars consecutive edge ids
     * (paradmm.c:424-427), each edge is appended to its variable's list in
     * creation order (paradmm.c:442-453) and edge->var is recorded
     * (paradmm.c:460-465).  Unlike the reference the caller's list is not
     * modified and storage grows geometrically. */
    int vars[64];
    if (nvars > 64) { fprintf(stderr, "oracle: factor degree > 64\n"); abort(); }
    memcpy(vars, vars_in, sizeof(int) * nvars);
    qsort(vars, nvars, sizeof(int), cmp_int);

    if (g->nfac == g->cap_fac) {
        g->cap_fac = g->cap_fac ? 2 * g->cap_fac : 64;
        g->fac_fn = (or_prox_fn *)xrealloc(g->fac_fn, sizeof(or_prox_fn) * g->cap_fac);
        g->fac_params = (void **)xrealloc(g->fac_params, sizeof(void *) * g->cap_fac);
        g->fac_deg = (int *)xrealloc(g->fac_deg, sizeof(int) * g->cap_fac);
        g->fac_first = (int *)xrealloc(g->fac_first, sizeof(int) * g->cap_fac);
    }
    int f = g->nfac++;
    g->fac_fn[f] = fn;
    g->fac_params[f] = params;
    g->fac_deg[f] = nvars;
    g->fac_first[f] = g->nedge;

    int maxv = 0;
    for (int i = 0; i < nvars; i++) if (vars[i] > maxv) maxv = vars[i];
    if (maxv + 1 > g->nvar) {
        if (maxv + 1 > g->cap_var) {
            int nc = g->cap_var ? g->cap_var : 64;
            while (nc < maxv + 1) nc *= 2;
            g->var_deg = (int *)xrealloc(g->var_deg, sizeof(int) * nc);
            g->var_cap = (int *)xrealloc(g->var_cap, sizeof(int) * nc);
            g->var_edges = (int **)xrealloc(g->var_edges, sizeof(int *) * nc);
            g->cap_var = nc;
        }
        for (int v = g->nvar; v < maxv + 1; v++) {
            g->var_deg[v] = 0; g->var_cap[v] = 0; g->var_edges[v] = NULL;
        }
        g->nvar = maxv + 1;
    }
    if (g->nedge + nvars > g->cap_edge) {
        int nc = g->cap_edge ? g->cap_edge : 256;
        while (nc < g->nedge + nvars) nc *= 2;
        g->edge_var = (int *)xrealloc(g->edge_var, sizeof(int) * nc);
        g->cap_edge = nc;
    }
    for (int i = 0; i < nvars; i++) {
        int v = vars[i], e = g->nedge + i;
        if (g->var_deg[v] == g->var_cap[v]) {
            g->var_cap[v] = g->var_cap[v] ? 2 * g->var_cap[v] : 8;
            g->var_edges[v] = (int *)xrealloc(g->var_edges[v], sizeof(int) * g->var_cap[v]);
        }
        g->var_edges[v][g->var_deg[v]++] = e;
        g->edge_var[e] = v;
    }
    g->nedge += nvars;
}

void or_alloc_state(or_graph *g)
{
    size_t le = (size_t)g->ndims * g->nedge, lv = (size_t)g->ndims * g->nvar;
    g->X = (double *)calloc(le ? le : 1, sizeof(double));
    g->M = (double *)calloc(le ? le : 1, sizeof(double));
    g->N = (double *)calloc(le ? le : 1, sizeof(double));
    g->U = (double *)calloc(le ? le : 1, sizeof(double));
    g->Z = (double *)calloc(lv ? lv : 1, sizeof(double));
    g->Zprev = (double *)calloc(lv ? lv : 1, sizeof(double));
    g->RHO = (double *)calloc(g->nedge ? g->nedge : 1, sizeof(double));
    g->ALPHA = (double *)calloc(g->nedge ? g->nedge : 1, sizeof(double));
    if (!g->X || !g->M || !g->N || !g->U || !g->Z || !g->Zprev || !g->RHO || !g->ALPHA) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
}

void or_set_rho_alpha(or_graph *g, double rho, double alpha)
{
    /* initialize_RHOS_APHAS (core/paradmm.c:188-207) */
    for (int e = 0; e < g->nedge; e++) { g->RHO[e] = rho; g->ALPHA[e] = alpha; }
}

/* ------------------------------------------------------------- serial passes */

void or_update_X(or_graph *g)
{
    /* updateX (core/paradmm.c:600-623): one prox call per factor on its slice
     * of X, N and RHOS, factors in creation order. */
    const int d = g->ndims;
    for (int f = 0; f < g->nfac; f++) {
        int s = g->fac_first[f];
        g->fac_fn[f](&g->X[(size_t)d * s], &g->N[(size_t)d * s], &g->RHO[s], g->fac_deg[f], d, g->fac_params[f]);
    }
}

void or_update_M(or_graph *g)
{
    /* updateM (core/paradmm.c:625-634) */
    size_t n = (size_t)g->ndims * g->nedge;
    for (size_t i = 0; i < n; i++) g->M[i] = g->X[i] + g->U[i];
}

void or_update_XM(or_graph *g)
{
    /* updateXM (core/paradmm.c:569-597) = X pass then M pass */
    or_update_X(g);
    or_update_M(g);
}

static void z_of_var(or_graph *g, int v)
{
    /* body of updateZ (core/paradmm.c:641-669): the rho sum over the
     * variable's edges in list order, then one weighted sum per dimension. */
    const int d = g->ndims;
    const int deg = g->var_deg[v];
    const int *el = g->var_edges[v];
    double den = 0;
    for (int i = 0; i < deg; i++) den += g->RHO[el[i]];
    for (int k = 0; k < d; k++) {
        double num = 0;
        for (int i = 0; i < deg; i++) {
            int e = el[i];
            num = num + g->M[(size_t)d * e + k] * g->RHO[e];
        }
        g->Z[(size_t)d * v + k] = num / den;
    }
}

void or_update_Z(or_graph *g)
{
    for (int v = 0; v < g->nvar; v++) z_of_var(g, v);
    /* update_outgoing_rhos (core/paradmm.c:564-566, 671) is a no-op */
}

static void u_of_var(or_graph *g, int v, int do_u, int do_n)
{
    /* body of updateUN / updateU / updateN (core/paradmm.c:678-733):
     * U_e += alpha_e (X_e - Z_v), then N_e = Z_v - U_e, edges of v in list order. */
    const int d = g->ndims;
    for (int j = 0; j < g->var_deg[v]; j++) {
        int e = g->var_edges[v][j];
        double a = g->ALPHA[e];
        for (int k = 0; k < d; k++) {
            size_t ek = (size_t)e * d + k, vk = (size_t)d * v + k;
            if (do_u) g->U[ek] += a * (g->X[ek] - g->Z[vk]);
            if (do_n) g->N[ek] = g->Z[vk] - g->U[ek];
        }
    }
}

void or_update_UN(or_graph *g) { for (int v = 0; v < g->nvar; v++) u_of_var(g, v, 1, 1); }
void or_update_U(or_graph *g) { for (int v = 0; v < g->nvar; v++) u_of_var(g, v, 1, 0); }
void or_update_N(or_graph *g) { for (int v = 0; v < g->nvar; v++) u_of_var(g, v, 0, 1); }

void or_iterate(or_graph *g, int iters, double *zhist)
{
    /* updateG (core/paradmm.c:736-741) repeated, as the serial demo loop
     * (packing/host.c:355-360).  Zprev is kept for the dual residual. */
    size_t lv = (size_t)g->ndims * g->nvar;
    for (int it = 0; it < iters; it++) {
        memcpy(g->Zprev, g->Z, sizeof(double) * lv);
        or_update_XM(g);
        or_update_Z(g);
        or_update_UN(g);
        g->iter++;
        if (zhist) memcpy(zhist + (size_t)it * lv, g->Z, sizeof(double) * lv);
    }
}

/* ------------------------------------------------------------- OpenMP passes */

void or_update_X_omp(or_graph *g)
{
    /* updateXOpenMP (core/paradmm.c:812-837) */
    const int d = g->ndims;
    const int nf = g->nfac;
#pragma omp parallel for
    for (int f = 0; f < nf; f++) {
        int s = g->fac_first[f];
        g->fac_fn[f](&g->X[(size_t)d * s], &g->N[(size_t)d * s], &g->RHO[s], g->fac_deg[f], d, g->fac_params[f]);
    }
}

void or_update_M_omp(or_graph *g)
{
    /* updateMOpenMP (core/paradmm.c:839-851) */
    const long n = (long)g->ndims * g->nedge;
#pragma omp parallel for
    for (long i = 0; i < n; i++) g->M[i] = g->X[i] + g->U[i];
}

void or_update_Z_omp(or_graph *g)
{
    /* updateZOpenMP (core/paradmm.c:853-894) */
    const int nv = g->nvar;
#pragma omp parallel for
    for (int v = 0; v < nv; v++) z_of_var(g, v);
}

void or_update_U_omp(or_graph *g)
{
    /* updateUOpenMP (core/paradmm.c:896-920) */
    const int nv = g->nvar;
#pragma omp parallel for
    for (int v = 0; v < nv; v++) u_of_var(g, v, 1, 0);
}

void or_update_N_omp(or_graph *g)
{
    /* updateNOpenMP (core/paradmm.c:923-945) */
    const int nv = g->nvar;
#pragma omp parallel for
    for (int v = 0; v < nv; v++) u_of_var(g, v, 0, 1);
}

double or_iterate_omp(or_graph *g, int iters, int nthreads)
{
    /* The OpenMP demo loop (packing/host.c:230-248): five parallel regions per
     * iteration, timed with gettimeofday around the loop only. */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    for (int it = 0; it < iters; it++) {
        or_update_X_omp(g);
        or_update_M_omp(g);
        or_update_Z_omp(g);
        or_update_U_omp(g);
        or_update_N_omp(g);
    }
    gettimeofday(&t1, NULL);
    g->iter += iters;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-6 * (double)(t1.tv_usec - t0.tv_usec);
}

void or_iterate_par(or_graph *g, int iters, int nthreads, double *zhist)
{
    /* or_iterate through the OpenMP passes: each variable's sums run in one
     * thread in list order, so the results equal the serial ones bit for
     * bit (tests/test_oracle_reference.py); Zprev kept, Z recorded. */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    size_t lv = (size_t)g->ndims * g->nvar;
    for (int it = 0; it < iters; it++) {
        memcpy(g->Zprev, g->Z, sizeof(double) * lv);
        or_update_X_omp(g);
        or_update_M_omp(g);
        or_update_Z_omp(g);
        or_update_U_omp(g);
        or_update_N_omp(g);
        g->iter++;
        if (zhist) memcpy(zhist + (size_t)it * lv, g->Z, sizeof(double) * lv);
    }
}

/* ------------------------------------------------------------ prox operators */

double or_max_alpha_radius(double n, double rho, double a)
{
    /* max_alpha_radius (packing/host.c:10-29): ten Newton steps on
     * d/dr [ -r^a + (rho/2)(r-n)^2 ] / a, started at |n/2|; a negative iterate
     * returns 0 at once. */
    double r = fabs(n * 0.5);
    for (int it = 0; it < 10; it++) {
        double f = -pow(r, a - 1.0) + (rho / a) * (r - n);
        /* the curvature term vanishes identically for a = 1; the reference
         * evaluates 0 * pow(0, -1) = NaN at r = 0 (packing/host.c:18) */
        double fp = (a == 1.0 ? 0.0 : -(a - 1.0) * pow(r, a - 2.0)) + (rho / a);
        r = r - f / fp;
        if (r < 0) return 0.0;
    }
    return r;
}

void or_prox_radius2(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* maximize_sum_of_squared_disk_radii_alpha (packing/host.c:67-81): per
     * edge, the radius from the first component, written to every component. */
    double a = ((const double *)params)[0];
    for (int i = 0; i < nvars; i++) {
        double r = or_max_alpha_radius(n[i * ndims], rhos[i], a);
        for (int k = 0; k < ndims; k++) x[i * ndims + k] = r;
    }
}

void or_prox_wall2(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* disk_plane_collision_variable_radius (packing/host.c:31-65): params are
     * the unit normal v (ndims) then a point p (ndims).  Edge 0 is the centre,
     * edge 1 the radius.  If v.(c-p) + r > 0 the pair (c, r) is projected,
     * rho-weighted, onto v.(c-p) + r = 0 with the radius clamped at 0. */
    const double *v = (const double *)params;
    const double *p = v + ndims;
    double s = 0;
    for (int k = 0; k < ndims; k++) s += v[k] * (n[k] - p[k]);
    if (s + n[ndims] > 0) {
        double r = (n[ndims] * rhos[1] - s * rhos[0]) / (rhos[0] + rhos[1]);
        if (r < 0) r = 0;
        for (int k = 0; k < ndims; k++) {
            x[k] = n[k] - s * v[k] - r * v[k];
            x[ndims + k] = r;
        }
    } else {
        for (int i = 0; i < ndims * nvars; i++) x[i] = n[i];
    }
}

/* The no-overlap projection shared by the dim-2 and dim-3 pair operators
 * (packing/host.c:83-132).  Coordinates along the centre line: circle 1 at 0,
 * circle 2 at dist.  If dist^2 < (r1+r2)^2, (x1, x2, r1, r2) is projected with
 * weights (w_x1, w_r1, w_x2, w_r2) onto x2 - x1 = r1 + r2, i.e.
 *   lam = (dist - r1 - r2) / (1/w_x1 + 1/w_r1 + 1/w_x2 + 1/w_r2)
 *   x1 = lam/w_x1, x2 = dist - lam/w_x2, r1 += lam/w_r1, r2 += lam/w_r2,
 * and both centres move along the unit direction c2 - c1.  Radii are not
 * clamped.  c1, c2 point at `nc` centre components; returns 1 if active. */
static int pair_project(const double *c1, const double *c2, int nc, double r1, double r2,
                        double w_x1, double w_r1, double w_x2, double w_r2,
                        double *c1_out, double *c2_out, double *r1_out, double *r2_out)
{
    double d2 = 0;
    for (int k = 0; k < nc; k++) { double t = c1[k] - c2[k]; d2 += t * t; }
    double s = r1 + r2;
    if (!(d2 < s * s)) return 0;
    double dist = sqrt(d2);
    double lam = (dist - s) / (1.0 / w_x1 + 1.0 / w_r1 + 1.0 / w_x2 + 1.0 / w_r2);
    double x1 = lam / w_x1;
    double x2 = dist - lam / w_x2;
    *r1_out = r1 + lam / w_r1;
    *r2_out = r2 + lam / w_r2;
    for (int k = 0; k < nc; k++) {
        /* coincident centres: the reference divides 0/0 (packing/host.c:123);
         * here they separate along the first axis, the circle of the smaller
         * radius towards -x (circle 1 on equal radii), so the outcome does
         * not depend on which circle is listed first */
        double u = dist > 0 ? (c2[k] - c1[k]) / dist : (k == 0 ? (r2 < r1 ? -1.0 : 1.0) : 0.0);
        c1_out[k] = c1[k] + u * x1;
        c2_out[k] = c2[k] + u * (x2 - dist);
    }
    return 1;
}

void or_prox_pair2(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* disk_collision_variable_radius (packing/host.c:83-132): edges are
     * (centre 1, radius 1, centre 2, radius 2); radii are read from the first
     * component and written to every component. */
    (void)params;
    const int d = ndims;
    for (int i = 0; i < d * nvars; i++) x[i] = n[i];
    double r1o, r2o;
    if (pair_project(&n[0], &n[2 * d], d, n[d], n[3 * d], rhos[0], rhos[1], rhos[2], rhos[3],
                     &x[0], &x[2 * d], &r1o, &r2o)) {
        for (int k = 0; k < d; k++) { x[d + k] = r1o; x[3 * d + k] = r2o; }
    }
}

void or_prox_pair3(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* dim-3 form of the pair operator: edges are (cx1,cy1,r1), (cx2,cy2,r2)
     * with one rho per circle, i.e. the reference weights with
     * w_x1 = w_r1 = rho_0 and w_x2 = w_r2 = rho_1. */
    (void)params; (void)ndims;
    for (int i = 0; i < 3 * nvars; i++) x[i] = n[i];
    pair_project(&n[0], &n[3], 2, n[2], n[5], rhos[0], rhos[0], rhos[1], rhos[1],
                 &x[0], &x[3], &x[2], &x[5]);
}

void or_prox_box3(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* NEW (no reference function): Euclidean projection of (cx,cy,r) onto
     * {r <= cx <= 1-r, r <= cy <= 1-r, r >= 0}.  A single-edge factor with a
     * scalar rho: rho cancels.  For fixed r the centre is a clamp, so r
     * minimises the convex f(r) = (r-nr)^2 + (r-mx)+^2 + (r-my)+^2 with
     * m = min(c, 1-c), on [0, 1/2]; its stationary point is one of nr,
     * (nr+a)/2, (nr+a+b)/3 (a <= b the sorted m's), then clamped.  Pinned by
     * an active-set brute force in tests/test_oracle_prox.py. */
    (void)rhos; (void)params; (void)ndims;
    for (int i = 0; i < nvars; i++) {
        const double *ni = n + 3 * i;
        double *xi = x + 3 * i;
        double mx = ni[0] < 1.0 - ni[0] ? ni[0] : 1.0 - ni[0];
        double my = ni[1] < 1.0 - ni[1] ? ni[1] : 1.0 - ni[1];
        double a = mx < my ? mx : my, b = mx < my ? my : mx;
        double nr = ni[2], r;
        if (nr <= a) r = nr;
        else if (0.5 * (nr + a) <= b) r = 0.5 * (nr + a);
        else r = (nr + a + b) / 3.0;
        if (r < 0) r = 0;
        if (r > 0.5) r = 0.5;
        double cx = ni[0], cy = ni[1];
        if (cx < r) cx = r;
        if (cx > 1.0 - r) cx = 1.0 - r;
        if (cy < r) cy = r;
        if (cy > 1.0 - r) cy = 1.0 - r;
        xi[0] = cx; xi[1] = cy; xi[2] = r;
    }
}

void or_prox_radius3(double *x, double *n, double *rhos, int nvars, int ndims, void *params)
{
    /* dim-3 form of the radius operator (packing/host.c:67-81): the radius
     * is component 2; the centre passes through. */
    (void)ndims;
    double a = ((const double *)params)[0];
    for (int i = 0; i < nvars; i++) {
        x[3 * i + 0] = n[3 * i + 0];
        x[3 * i + 1] = n[3 * i + 1];
        x[3 * i + 2] = or_max_alpha_radius(n[3 * i + 2], rhos[i], a);
    }
}

/* ------------------------------------------------------------ seeded init */

uint64_t or_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double or_uniform01(uint64_t *s)
{
    return (double)(or_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

or_graph *or_build_square(int nc, uint64_t seed, double rho, double alpha, double a_exp)
{
    or_graph *g = or_new(3);
    g->owned_params = (double *)xmalloc(sizeof(double));
    g->owned_params[0] = a_exp;
    for (int i = 0; i < nc; i++) or_add_factor(g, or_prox_radius3, g->owned_params, 1, &i);
    for (int i = 0; i < nc; i++) or_add_factor(g, or_prox_box3, NULL, 1, &i);
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[2] = {i, j};
            or_add_factor(g, or_prox_pair3, NULL, 2, vv);
        }
    or_alloc_state(g);
    or_set_rho_alpha(g, rho, alpha);
    uint64_t s = seed;
    for (int i = 0; i < nc; i++) {
        g->Z[3 * i + 0] = or_uniform01(&s);
        g->Z[3 * i + 1] = or_uniform01(&s);
        g->Z[3 * i + 2] = 0.05 * or_uniform01(&s);
    }
    for (int e = 0; e < g->nedge; e++) {
        int v = g->edge_var[e];
        for (int k = 0; k < 3; k++) {
            double z = g->Z[3 * v + k];
            g->X[3 * e + k] = z; g->N[3 * e + k] = z; g->M[3 * e + k] = z; g->U[3 * e + k] = 0;
        }
    }
    memcpy(g->Zprev, g->Z, sizeof(double) * 3 * g->nvar);
    return g;
}

or_graph *or_build_triangle(int nc, uint64_t seed, double lo, double hi)
{
    /* Factor order and parameters of packing/host.c:256-343. */
    or_graph *g = or_new(2);
    const double deg = 3.14159265358979323846 / 3.0;
    g->owned_params = (double *)xmalloc(sizeof(double) * 13);
    double *a = g->owned_params;
    a[0] = 2.0;
    double *w = a + 1; /* 3 walls x (normal, point) */
    double wall[3][4] = {{0., -1., 0., 0.}, {-sin(deg), cos(deg), 0., 0.}, {sin(deg), cos(deg), 1., 0.}};
    memcpy(w, wall, sizeof(wall));
    for (int i = 0; i < nc; i++) { int v = 2 * i + 1; or_add_factor(g, or_prox_radius2, a, 1, &v); }
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < nc; j++) {
            int vv[2] = {2 * j, 2 * j + 1};
            or_add_factor(g, or_prox_wall2, w + 4 * k, 2, vv);
        }
    for (int i = 0; i < nc - 1; i++)
        for (int j = i + 1; j < nc; j++) {
            int vv[4] = {2 * i, 2 * i + 1, 2 * j, 2 * j + 1};
            or_add_factor(g, or_prox_pair2, NULL, 4, vv);
        }
    or_alloc_state(g);
    or_set_rho_alpha(g, 10.0, 0.1);
    for (int i = 0; i < nc; i++) g->RHO[i] = 5.0;
    uint64_t s = seed;
    size_t le = (size_t)2 * g->nedge, lv = (size_t)2 * g->nvar;
    for (size_t i = 0; i < le; i++) {
        g->X[i] = lo + (hi - lo) * or_uniform01(&s);
        g->N[i] = lo + (hi - lo) * or_uniform01(&s);
        g->M[i] = lo + (hi - lo) * or_uniform01(&s);
        g->U[i] = lo + (hi - lo) * or_uniform01(&s);
    }
    for (size_t i = 0; i < lv; i++) g->Z[i] = lo + (hi - lo) * or_uniform01(&s);
    memcpy(g->Zprev, g->Z, sizeof(double) * lv);
    return g;
}

/* ------------------------------------------------------------ warm start */

void or_warm_start(or_graph *g, const double *Zinit, const double *Uinit, int maxiter)
{
    /* initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375): Z = Zinit, X the
     * Z of each edge's variable, then per factor maxiter sweeps of
     *   Xit = prox(N);  U = Xit - N;  N = X - U   (N starts at X - Uinit),
     * finally M = X + U.  With maxiter = 0 the reference leaves U unset; here
     * U = Uinit. */
    const int d = g->ndims;
    size_t le = (size_t)d * g->nedge, lv = (size_t)d * g->nvar;
    memcpy(g->Z, Zinit, sizeof(double) * lv);
    for (int e = 0; e < g->nedge; e++)
        for (int k = 0; k < d; k++) g->X[(size_t)d * e + k] = Zinit[(size_t)d * g->edge_var[e] + k];
    double *xit = (double *)xmalloc(sizeof(double) * (le ? le : 1));
    double *rho = (double *)xmalloc(sizeof(double) * (g->nedge ? g->nedge : 1));
    memcpy(rho, g->RHO, sizeof(double) * g->nedge);
    for (size_t i = 0; i < le; i++) { g->U[i] = Uinit[i]; g->N[i] = g->X[i] - Uinit[i]; xit[i] = g->X[i]; }
    for (int f = 0; f < g->nfac; f++) {
        size_t s = (size_t)d * g->fac_first[f];
        int len = d * g->fac_deg[f];
        for (int it = 0; it < maxiter; it++) {
            g->fac_fn[f](&xit[s], &g->N[s], &rho[g->fac_first[f]], g->fac_deg[f], d, g->fac_params[f]);
            for (int j = 0; j < len; j++) {
                g->U[s + j] = xit[s + j] - g->N[s + j];
                g->N[s + j] = g->X[s + j] - g->U[s + j];
            }
        }
    }
    for (size_t i = 0; i < le; i++) g->M[i] = g->X[i] + g->U[i];
    free(xit);
    free(rho);
}

/* ------------------------------------------------------------ residuals */

void or_residuals(const or_graph *g, double *primal2, double *dual2)
{
    /* NEW (the reference computes no residual): primal^2 = sum over edges of
     * ||X_e - Z_v(e)||^2 and dual^2 = sum over edges of rho_e^2 ||Z_v - Zprev_v||^2. */
    const int d = g->ndims;
    double p = 0, q = 0;
    for (int e = 0; e < g->nedge; e++) {
        int v = g->edge_var[e];
        double dz2 = 0;
        for (int k = 0; k < d; k++) {
            double t = g->X[(size_t)d * e + k] - g->Z[(size_t)d * v + k];
            p += t * t;
            double u = g->Z[(size_t)d * v + k] - g->Zprev[(size_t)d * v + k];
            dz2 += u * u;
        }
        q += g->RHO[e] * g->RHO[e] * dz2;
    }
    *primal2 = p;
    *dual2 = q;
}

void or_square_quality(const or_graph *g, int nc, double *max_overlap, double *max_box, double *sum_r)
{
    const double *Z = g->Z;
    double mo = 0, mb = 0, sr = 0;
    for (int i = 0; i < nc; i++) {
        double cx = Z[3 * i], cy = Z[3 * i + 1], r = Z[3 * i + 2];
        sr += r;
        double viol[5] = {r - cx, cx + r - 1.0, r - cy, cy + r - 1.0, -r};
        for (int k = 0; k < 5; k++) if (viol[k] > mb) mb = viol[k];
        for (int j = i + 1; j < nc; j++) {
            double dx = cx - Z[3 * j], dy = cy - Z[3 * j + 1];
            double ov = r + Z[3 * j + 2] - sqrt(dx * dx + dy * dy);
            if (ov > mo) mo = ov;
        }
    }
    *max_overlap = mo;
    *max_box = mb;
    *sum_r = sr;
}

/* ------------------------------------------------------------ accessors */

int or_nvar(const or_graph *g) { return g->nvar; }
int or_nedge(const or_graph *g) { return g->nedge; }
int or_nfac(const or_graph *g) { return g->nfac; }
int or_ndims(const or_graph *g) { return g->ndims; }
const int *or_edge_var(const or_graph *g) { return g->edge_var; }
const int *or_fac_first(const or_graph *g) { return g->fac_first; }
const int *or_fac_deg(const or_graph *g) { return g->fac_deg; }
int or_var_deg(const or_graph *g, int v) { return g->var_deg[v]; }
const int *or_var_edges(const or_graph *g, int v) { return g->var_edges[v]; }

double *or_ptr(or_graph *g, int which)
{
    switch (which) {
    case 0: return g->X;
    case 1: return g->M;
    case 2: return g->N;
    case 3: return g->U;
    case 4: return g->Z;
    case 5: return g->RHO;
    case 6: return g->ALPHA;
    case 7: return g->Zprev;
    default: return NULL;
    }
}
