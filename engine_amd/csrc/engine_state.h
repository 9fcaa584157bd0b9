// engine_state.h -- device state of one graph, shared by engine.hip and the
// layout-specific files.  Internal.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "paradmm.h"

struct paradmm_rho_mlp;

namespace pdm {

// device parameter vector of a rho network (mlp.hip)
const float* mlp_device_params(paradmm_rho_mlp* m);

struct KindInfo {
    int deg;      // edges per factor; 0 = any degree, each edge independent
    int dims;     // num_dims the operator is defined for
    int nparams;  // doubles read from the factor's parameter pointer
    const char* name;
};
const KindInfo& kind_info(int kind);

struct KindTable {
    int kind = 0;
    int nitems = 0;
    int* d_items = nullptr;     // [deg][nitems] device positions
    float* d_params = nullptr;  // [nitems][nparams]
};

struct DenseState;  // dense.hip

// Factors whose callback has no device kind: their prox runs on the host,
// in FP64, between a device gather of (N, X, rho) and a scatter of (X, rho).
struct HostProx {
    std::vector<void*> fn;      // the callbacks (reference signature)
    std::vector<void*> params;  // their parameter pointers
    std::vector<int> first;     // first staged edge of each factor
    std::vector<int> deg;       // its degree
    int nedges = 0;
    int maxdeg = 0;
    int* d_pos = nullptr;       // [nedges] device positions, factor order
    float* d_buf = nullptr;     // [nedges][2D+1]: N, X, rho
    float* h_buf = nullptr;     // pinned host mirror
};

// One timed kernel launch: events recorded on the engine stream around it.
struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

struct DevGraph {
    graph* G = nullptr;
    int E = 0, V = 0, F = 0, D = 0;
    int layout = 0;   // 1 generic, 2 dense pair tiles
    long stride = 0;  // elements per component of the edge SoA buffers
    long npos = 0;

    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> allocs;

    // maps
    int* d_pos = nullptr;           // api edge -> device position
    int* d_edge_var_api = nullptr;  // api edge -> variable
    int* d_edge_var_dev = nullptr;  // device position -> variable (generic)
    int* d_var_off = nullptr;       // [V+1] CSR offsets (generic)
    std::vector<KindTable> kinds;   // generic X pass tables

    // state
    float *d_X = nullptr, *d_U = nullptr, *d_N = nullptr, *d_M = nullptr;
    float *d_rho = nullptr, *d_alpha = nullptr;
    float *d_Z = nullptr, *d_Zprev = nullptr;
    double* d_res_part = nullptr;  // residual partials
    double* d_res = nullptr;       // [2] primal^2, dual^2
    double* d_stage = nullptr;     // host transfer staging (fp64)
    double* d_quality = nullptr;   // packing quality partials
    unsigned char* d_edge_type_api = nullptr;  // api edge -> PDM_EDGE_* (pair, container, radius)
    float* d_ckpt = nullptr;       // saved X, U, rho, Z, Zprev (paradmm_gpu_state_save)
    bool ckpt_n_from_buffer = false;
    bool ckpt_rho_uni = false;
    float ckpt_rho_u = 0.0f;
    int ckpt_raised = 0;

    // coherence
    bool host_newer = true;    // host arrays hold data the device has not seen
    bool dev_newer = false;    // device holds iterations the host has not seen
    bool n_from_buffer = true; // the next X pass must read N from the N buffer
    bool m_valid = false;      // the M buffer holds the M of the last iteration
    bool rho_changed = true;   // rho/alpha uploaded since the last per-variable sums
    bool rho_device_written = false;
    // every edge holds the same rho / alpha (set at upload, cleared by a
    // per-edge change on the device): the dense kernels then skip those streams
    bool rho_uni = false, alpha_uni = false;
    float rho_u = 0.0f, alpha_u = 0.0f;
    bool iters_done = false;
    bool dense_unary_stale = false;

    // captured runs of iterations: a run of graph_iters is replayed as often
    // as it fits, the rest of a call is one run of its own length; runs are
    // kept per (length, kernel variant), a few at a time
    struct CapturedRun {
        int iters = 0, variant = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    int graph_iters = 64;
    std::vector<CapturedRun> runs;

    DenseState* dense = nullptr;
    HostProx* host = nullptr;  // generic layout only
    // generic layout, dim-2 all-pairs packing of the reference demo: every
    // pair factor (pos_i, rad_i, pos_j, rad_j), i < j, has its edges in
    // rows -- circle c's pos and rad edges of its pairs at rowP[c] + col
    // and rowR[c] + col, col = the other circle's index, less one past c --
    // so the pair X pass runs as coalesced tiles (k_xpass_pair2d_rows)
    int zpass_wg = 0;           // generic layout: edges per thread of k_zpass_fused_wg (0: every variable's
                                // edges do not fit its registers)
    int rows2d_n = 0;           // circles (0: no row structure)
    int* d_rowP = nullptr;      // [rows2d_n]
    int* d_rowR = nullptr;

    // rho policy applied by paradmm_gpu_solve every policy_every iterations;
    // iter_frac = iterations done / max_iters.  Set by the learned-rho module.
    int (*rho_policy)(DevGraph* s, float iter_frac) = nullptr;
    void* policy_ctx = nullptr;
    int policy_per_type = 0;           // 0 per edge, 1 per-type mean, 2 one mean over all edges,
                                       // 3 that mean plus the edges the network raises
    float* d_policy_y = nullptr;       // per-slot network output (mode 3)
    int policy_raised = 0;             // edges raised by the last mode-3 application
    bool policy_average = false;       // average the estimates over the applications so far
    bool policy_in_solve = false;      // applied by paradmm_gpu_solve, with iterations to follow
    void* d_policy_part = nullptr;     // per-type partial sums of that mean
    int* d_slot_var = nullptr;         // device slot -> variable (-1: padding), policy kernels
    unsigned char* d_slot_type = nullptr;

    // kernel timing (paradmm_gpu_profile_kernels): eager launches only
    bool prof = false;
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    size_t prof_used = 0;
};

// Brackets one kernel launch with events when profiling (no-op otherwise).
void prof_begin(DevGraph* s, const char* name);
void prof_end(DevGraph* s);
// For a launch through hipExtLaunchKernelGGL, whose dispatch records the
// events at the kernel's own start and end (no event-record overhead in
// the figure): both null when not profiling.
void prof_dispatch(DevGraph* s, const char* name, hipEvent_t& start, hipEvent_t& stop);

int hip_fail(hipError_t e, const char* what, const char* file, int line);
// device state of G, finalised; with upload, host writes since are uploaded
int pdm_get_state(graph* G, DevGraph** out, bool upload);
// kind of every factor; with allow_host, a callback without a device kind
// gets kind -1 (its prox runs on the host) instead of an error
int pdm_factor_kinds(graph* G, std::vector<int>& fkind, bool allow_host);
// the dense pair-tile slot of every edge for tile side T; -1 when the graph
// is not the all-pairs unit-square packing
int dense_structure(graph* G, const std::vector<int>& fkind, int T, std::vector<int>& pos, std::vector<float>& aexp,
                    int& nb, int& nt);
int dmalloc(DevGraph* s, void** p, size_t bytes);
int ensure_aux(DevGraph* s);
int launch_mrecon(DevGraph* s);
int launch_npass(DevGraph* s);
int residual_reduce(DevGraph* s);
// unit-square packing read from Z: out = {max overlap, max container violation, sum r}
int square_quality(DevGraph* s, double out[3]);
int warm_start_device(DevGraph* s, int maxiter);
void free_dev(DevGraph* s);

// dense.hip: returns -1 if the graph is not an all-pairs dim-3 packing,
// 0 when the dense layout was built, >0 on error.
int dense_try_finalize(DevGraph* s, graph* G, const std::vector<int>& fkind);
int dense_alloc_state(DevGraph* s);
// mphase: DENSE_PLAIN, or the place of the iteration in a captured run with
// alpha = 1 (enter, middle, exit of the M state)
// or DENSE_CONTACT: a contact-state iteration that skips the residual sums
// (in the contact state every other phase is a contact iteration with them)
enum { DENSE_PLAIN = 0, DENSE_M_ENTER = 1, DENSE_M_MID = 2, DENSE_M_EXIT = 3, DENSE_CONTACT = 4 };
int dense_iteration(DevGraph* s, int from_n, int mphase = DENSE_PLAIN);
int dense_legacy_xpass(DevGraph* s);
int dense_legacy_zpass(DevGraph* s);
int dense_residual_reduce(DevGraph* s);
void dense_free(DevGraph* s);
// true when the next dense iteration is the steady-state (captured) kind
bool dense_steady(DevGraph* s);
// device copy {rho_u, alpha_u} read by the uniform kernel variant (dense only)
float* dense_uni_ptr(DevGraph* s);
// bit of a dense launch-order tile entry's J: the tile holds edges whose rho
// differs from the common one (a rho policy raised them), so a launch of
// the uniform-rho variant takes that tile through the per-edge streams
constexpr int TILE_RAISED = 1 << 30;

// dense layout: the launch-order tile list (whose TILE_RAISED bits flag
// tiles with edges of their own rho), the layout -> launch position map,
// the tile count and the slots per tile; null for other layouts
int2* dense_tile_flags(DevGraph* s, const int** tile_pos, int* nt, int* tile_slots);
// writes s->rho_u, s->alpha_u to that copy (stream-ordered, synchronous)
int sync_uniform(DevGraph* s);
// the device state was replaced from the host: U is current
void dense_reset(DevGraph* s);
// applies the pending U update (and its primal residual) so U is current
// (leaves the contact state first)
int dense_finish(DevGraph* s);
// contact state (alpha = 1): only the pairs whose prox is active keep
// per-edge values; PARADMM_CONTACT=0 disables it
bool dense_contact_ok(DevGraph* s);
bool dense_in_contact(DevGraph* s);
// the contact state's kernels read per-edge rho (chosen when it is entered)
bool dense_contact_rho_per_edge(DevGraph* s);
// from the steady state: enters the contact state and runs one iteration in it
int dense_contact_enter(DevGraph* s);
// back to the steady X, U state (no-op outside the contact state)
int dense_contact_leave(DevGraph* s);
// which contact list the last iteration wrote: contact kernels take it as an
// argument, so a captured run is bound to its start parity and leaves the
// parity flipped by its length
int dense_contact_parity(DevGraph* s);
// whether the contact state can take a rho policy application itself
// (every circle's unlisted pair edges share one rho)
bool dense_contact_policy_ok(DevGraph* s);
// one application of the rho policy (mode 2 or 3) inside the contact state;
// valid only when a contact iteration follows before the state is left
int dense_contact_policy(DevGraph* s, const float* params, float frac, int mode, bool average, double* hist);
// contact tile side and pairs in the last list (0, 0 outside the contact state)
int dense_contact_info(DevGraph* s, int* tile_side, long* listed_pairs);
void dense_contact_set_parity(DevGraph* s, int p);

}  // namespace pdm
