/*
 * legacy.c -- the update entry points of the reference API
 * (core/paradmm.c:564-741 serial, 812-1092 OpenMP), routed to the GPU.
 *
 * Each call runs its passes on the device in the reference order
 * X -> M -> Z -> U -> N and leaves the host arrays exactly as the reference
 * would (uploading the pass inputs, downloading the pass outputs), so a
 * caller that reads or writes G.X, G.Z, G.RHOS ... between calls sees the
 * reference semantics.  The OpenMP variants are the same device passes;
 * their thread count is ignored.
 */
#include "paradmm.h"
#include "engine_internal.h"

#include <stdio.h>
#include <stdlib.h>

static void run(graph *G, int passes)
{
    int rc = pdm_legacy_passes(G, passes);
    if (rc) pdm_legacy_fail(rc);
}

void update_outgoing_rhos(graph *G)
{
    /* core/paradmm.c:564-566: the reference hook is empty.  Rho adaptation
     * runs inside the device iteration (paradmm_gpu.h rho policy). */
    (void)G;
}

void updateXM(graph *G) { run(G, PDM_PASS_X | PDM_PASS_M); }   /* core/paradmm.c:569-597 */
void updateX(graph *G) { run(G, PDM_PASS_X); }                 /* core/paradmm.c:600-623 */
void updateM(graph *G) { run(G, PDM_PASS_M); }                 /* core/paradmm.c:625-634 */
void updateZ(graph *G) { run(G, PDM_PASS_Z); update_outgoing_rhos(G); } /* core/paradmm.c:637-672 */
void updateUN(graph *G) { run(G, PDM_PASS_U | PDM_PASS_N); }   /* core/paradmm.c:675-694 */
void updateU(graph *G) { run(G, PDM_PASS_U); }                 /* core/paradmm.c:696-715 */
void updateN(graph *G) { run(G, PDM_PASS_N); }                 /* core/paradmm.c:718-734 */
void updateG(graph *G) { run(G, PDM_PASS_X | PDM_PASS_M | PDM_PASS_Z | PDM_PASS_U | PDM_PASS_N); } /* core/paradmm.c:736-741 */

/* core/paradmm.c:812-945 */
void updateXOpenMP(graph *G) { run(G, PDM_PASS_X); }
void updateMOpenMP(graph *G) { run(G, PDM_PASS_M); }
void updateZOpenMP(graph *G) { run(G, PDM_PASS_Z); update_outgoing_rhos(G); }
void updateUOpenMP(graph *G) { run(G, PDM_PASS_U); }
void updateNOpenMP(graph *G) { run(G, PDM_PASS_N); }

void updateGOpenMP(graph *G, int maxiter, int numthreads)
{
    /* core/paradmm.c:949-1092: maxiter whole iterations in one call.  On the
     * GPU they run device-resident with the host synchronised once before and
     * once after. */
    (void)numthreads;
    int rc = pdm_legacy_iterations(G, maxiter);
    if (rc) pdm_legacy_fail(rc);
}
