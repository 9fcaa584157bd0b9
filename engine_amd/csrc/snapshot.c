/*
 * snapshot.c -- on-disk snapshot of a graph's iteration state.
 *
 * The reference keeps X, M, N, U, Z, RHOS and ALPHAS in host arrays of the
 * graph struct (core/paradmm.h:53-61) and has no file format.  A snapshot
 * stores exactly those arrays so that a solve can be resumed, compared or
 * used as a fixture:
 *
 *   offset 0   char[8]   "PDMSNAP1"
 *          8   int32     num_dims, num_vars, num_opts, num_edges
 *         24   uint64    structure hash: FNV-1a over operatorix_to_numedges
 *                        and edgeix_to_varlistix (the edge numbering)
 *         32   float64   X, M, N, U [num_edges * num_dims] each,
 *                        Z [num_vars * num_dims], RHOS, ALPHAS [num_edges]
 *
 * Little-endian, as written by the host.  Saving first brings the host
 * arrays up to date with the device; loading marks them as written so the
 * device state is refreshed before the next device call.
 */
#include "engine_internal.h"
#include "paradmm.h"
#include "paradmm_gpu.h"

#include <stdint.h>
#include <stdio.h>
#include <string.h>

static const char MAGIC[8] = {'P', 'D', 'M', 'S', 'N', 'A', 'P', '1'};

static uint64_t fnv1a(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t structure_hash(const graph *G)
{
    uint64_t h = 14695981039346656037ull;
    h = fnv1a(h, G->operatorix_to_numedges, sizeof(int) * (size_t)G->num_opts);
    h = fnv1a(h, G->edgeix_to_varlistix, sizeof(int) * (size_t)G->num_edges);
    return h;
}

static int arrays_ready(const graph *G, const char *who)
{
    if (!G || G->num_edges <= 0 || !G->X || !G->M || !G->N || !G->U || !G->Z || !G->RHOS || !G->ALPHAS) {
        pdm_set_error(PARADMM_ERR_STATE, "%s: graph state not allocated", who);
        return PARADMM_ERR_STATE;
    }
    return PARADMM_OK;
}

int paradmm_save_state(graph *G, const char *path)
{
    int rc = arrays_ready(G, "paradmm_save_state");
    if (rc) return rc;
    rc = pdm_sync_if_device_newer(G);
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_save_state: cannot open %s", path);
        return PARADMM_ERR_ARG;
    }
    const int32_t hdr[4] = {G->num_dims, G->num_vars, G->num_opts, G->num_edges};
    const uint64_t h = structure_hash(G);
    const size_t ed = (size_t)G->num_edges * G->num_dims, vd = (size_t)G->num_vars * G->num_dims;
    int ok = fwrite(MAGIC, 1, 8, f) == 8 && fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(&h, sizeof h, 1, f) == 1;
    const double *arrs[4] = {G->X, G->M, G->N, G->U};
    for (int a = 0; ok && a < 4; a++) ok = fwrite(arrs[a], sizeof(double), ed, f) == ed;
    ok = ok && fwrite(G->Z, sizeof(double), vd, f) == vd;
    ok = ok && fwrite(G->RHOS, sizeof(double), (size_t)G->num_edges, f) == (size_t)G->num_edges;
    ok = ok && fwrite(G->ALPHAS, sizeof(double), (size_t)G->num_edges, f) == (size_t)G->num_edges;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_save_state: write to %s failed", path);
        return PARADMM_ERR_ARG;
    }
    return PARADMM_OK;
}

int paradmm_load_state(graph *G, const char *path)
{
    int rc = arrays_ready(G, "paradmm_load_state");
    if (rc) return rc;
    FILE *f = fopen(path, "rb");
    if (!f) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_load_state: cannot open %s", path);
        return PARADMM_ERR_ARG;
    }
    char magic[8];
    int32_t hdr[4];
    uint64_t h = 0;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, MAGIC, 8) != 0 || fread(hdr, sizeof hdr, 1, f) != 1 ||
        fread(&h, sizeof h, 1, f) != 1) {
        fclose(f);
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_load_state: %s is not a snapshot", path);
        return PARADMM_ERR_ARG;
    }
    if (hdr[0] != G->num_dims || hdr[1] != G->num_vars || hdr[2] != G->num_opts || hdr[3] != G->num_edges ||
        h != structure_hash(G)) {
        fclose(f);
        pdm_set_error(PARADMM_ERR_SHAPE, "paradmm_load_state: %s was saved from another graph structure", path);
        return PARADMM_ERR_SHAPE;
    }
    const size_t ed = (size_t)G->num_edges * G->num_dims, vd = (size_t)G->num_vars * G->num_dims;
    const size_t ne = (size_t)G->num_edges;
    /* the payload must be complete before any array of the graph is touched:
     * a short file leaves the graph as it was */
    const long payload_at = ftell(f);
    const long long need = (long long)sizeof(double) * (long long)(4 * ed + vd + 2 * ne);
    long end = -1;
    if (payload_at >= 0 && fseek(f, 0, SEEK_END) == 0) end = ftell(f);
    if (end < 0 || (long long)(end - payload_at) != need || fseek(f, payload_at, SEEK_SET) != 0) {
        fclose(f);
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_load_state: %s is truncated or has trailing bytes", path);
        return PARADMM_ERR_ARG;
    }
    double *arrs[4] = {G->X, G->M, G->N, G->U};
    int ok = 1;
    for (int a = 0; ok && a < 4; a++) ok = fread(arrs[a], sizeof(double), ed, f) == ed;
    ok = ok && fread(G->Z, sizeof(double), vd, f) == vd;
    ok = ok && fread(G->RHOS, sizeof(double), ne, f) == ne;
    ok = ok && fread(G->ALPHAS, sizeof(double), ne, f) == ne;
    fclose(f);
    pdm_host_written(G);
    if (!ok) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_load_state: reading %s failed", path);
        return PARADMM_ERR_ARG;
    }
    return PARADMM_OK;
}
