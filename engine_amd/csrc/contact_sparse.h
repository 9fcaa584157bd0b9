// contact_sparse.h -- device code of the spatial form of the contact state
// (dense.hip).  Internal.
//
// The id-order contact state tests every pair of the graph every iteration:
// its tiles are blocks of consecutive circle ids, which lie anywhere in the
// square, so every tile holds a few contacts and must run.  The spatial
// form numbers the circles in the Morton order of their centres when the
// state is entered (contact index c, circle perm[c]); a block of CS
// consecutive contact indices then covers a small patch of the square and
// most pairs of blocks are too far apart to touch.  An iteration is two
// launches:
//   k_scs_circles  one thread per circle: Z_k from the previous launch's
//                  sums, (Z_k, c) for the tiles, the launch's sum buffer
//                  cleared, and each block's box of the circles' common N,
//                  widened by a rounding margin;
//   k_scs_rows     a few workgroups per block row P, each listing its share
//                  of the tiles (P, Q >= P) that run (the two boxes meet, or
//                  the previous list holds pairs) in Q order in LDS; block
//                  P's box and radius edges; then its wavefronts take the
//                  listed tiles in turn (the id-order form's tile body).
// A pair whose blocks' boxes do not meet cannot pass the prox's collision
// test (the margin covers the test's roundings), and an inactive pair that
// is in no list adds nothing to any sum, so the skipped tiles change no
// value: the circle sums are the same integers as the id-order form's,
// and Z, X and U are the same bits.
#pragma once
#include <hip/hip_runtime.h>

#include "contact_device.h"

namespace pdm {

// 32-bit Morton code of a centre, each coordinate quantised to 16 bits over
// [0, 1] (clamped: a circle outside the square only loses locality)
__device__ __forceinline__ unsigned scs_spread16(unsigned v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ unsigned scs_morton(float x, float y)
{
    const float qx = fminf(fmaxf(x, 0.0f), 1.0f) * 65535.0f, qy = fminf(fmaxf(y, 0.0f), 1.0f) * 65535.0f;
    // NaN compares false in both clamps' sense: fmin/fmax return the number
    const unsigned ix = (unsigned)qx, iy = (unsigned)qy;
    return scs_spread16(ix) | (scs_spread16(iy) << 1);
}

// Per circle: Z_k from the previous launch's sums (or the canonical arrays
// at a run's first launch), stored to the run's ring and, with
// c = Z_k - Z_{k-1}, to the table the tiles read (the same operations as
// the id-order form's contact_zc, so the same bits); the launch's sum
// buffer cleared.  Returns the circle's common N as the tiles form it.
// A lane past the last circle (valid false) calls it with c = N - 1 and
// stores nothing: the loads then need no branch, so they issue together
// with the kernel argument loads they depend on.
template <bool RES>
__device__ __forceinline__ void scs_circle(const CArgs& a, const int c, const bool valid, float (&n)[3])
{
    constexpr int Q = RES ? CQ : 3;
    const int N = a.N;
    float z[3], cz[3];
    if (a.Ap) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float zk = a.Zin[3 * c + k];
            z[k] = zk + csum_f32(a.Ap[(long)k * N + c]) / (float)(N + 1);
            cz[k] = z[k] - zk;
        }
    } else {
        const int pc = a.perm[c];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            z[k] = a.Z[3 * pc + k];
            cz[k] = z[k] - a.Zp[3 * pc + k];
        }
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.Zo[3 * c + k] = z[k];
        a.zcv[2 * c] = make_float4(z[0], z[1], z[2], cz[0]);
        a.zcv[2 * c + 1] = make_float4(cz[1], cz[2], 0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < Q; ++q) a.Ac[(long)q * N + c] = 0ull;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) n[k] = z[k] + cz[k];
}

// The box and radius edges of block P's circles (lane e < CS: contact
// circle P CS + e): the id-order form's contact_unary on the contact-order
// compact array, reading (Z_k, c) from the table.
template <int CS, bool RES>
__device__ __forceinline__ void scs_unary(const CArgs& a, const int P)
{
    constexpr int Q = RES ? CQ : 3;
    const int e = (int)(threadIdx.x & 63);
    const int c = P * CS + e;
    if (e >= CS || c >= a.N) return;
    const long NL = a.N;
    const float4 u4 = a.zcv[2 * c], v4 = a.zcv[2 * c + 1];
    const float z[3] = {u4.x, u4.y, u4.z}, cz[3] = {u4.w, v4.x, v4.y};
    long long s[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) s[q] = 0;
    float g2 = 0.0f;
    unsigned bad = 0u;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        float* xu = a.Xu + 6 * side * NL + c;
        const float r = a.uni[0];
        float x[3], u[3], nn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = xu[k * NL];
            u[k] = (x[k] + xu[(3 + k) * NL]) - z[k];
            nn[k] = z[k] - u[k];
        }
        if (side == 0) {
            box3(x[0], x[1], x[2], nn[0], nn[1], nn[2]);
        } else {
            x[0] = nn[0];
            x[1] = nn[1];
            x[2] = max_alpha_radius(nn[2], r, a.aexp[a.perm[c]]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xu[k * NL] = x[k];
            xu[(3 + k) * NL] = u[k];
            s[k] += cfix((x[k] + u[k]) - z[k], bad);
            if constexpr (RES) {
                const float g = (x[k] - z[k]) - cz[k];
                s[3 + k] += cfix(g, bad);
                g2 += g * g;
            }
        }
        if constexpr (RES) {
            long long hi, lo;
            cfix2(g2, hi, lo, bad);
            s[6] += hi;
            s[7] += lo;
            g2 = 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (s[q]) atomicAdd(&a.Ac[(long)q * NL + c], (unsigned long long)s[q]);
    if (bad) atomicOr(a.flag, 1u);
}

// v from another lane of its 16-lane row by DPP control CTRL (a VALU
// operand modifier: no LDS round trip).  Every lane of the row is active.
template <int CTRL>
__device__ __forceinline__ float scs_dpp(float v)
{
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false));
}

// min (MAX false) or max of v over each aligned group of G lanes (G = 16 or
// 32), every lane of the group receiving it.  Within a 16-lane row: the
// quad swaps [1,0,3,2] and [2,3,0,1] leave each lane its quad's value, the
// half-row mirror (lane i of 8 <- lane 7 - i) its 8 lanes', the row mirror
// (i <- 15 - i) its row's; min and max do not depend on the order.  Across
// the two rows of 32, one shuffle.
template <int G, bool MAX>
__device__ __forceinline__ float scs_group_extreme(float v)
{
    static_assert(G == 16 || G == 32, "groups of 16 or 32 lanes");
    auto op = [](float x, float y) { return MAX ? fmaxf(x, y) : fminf(x, y); };
    v = op(v, scs_dpp<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, scs_dpp<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, scs_dpp<0x141>(v));  // row_half_mirror
    v = op(v, scs_dpp<0x140>(v));  // row_mirror
    if constexpr (G == 32) v = op(v, __shfl_xor(v, 16, 64));
    return v;
}

// The box of a block's common N, x and y each widened by |r|, then by
// 1e-5 (|x| + |y| + |r|)max: the collision test d2 < (r1 + r2)^2 rounds a
// few times by 2^-24 relative, so a pair that passes it has
// |x1 - x2| <= |r1| + |r2| up to far less than that margin, and so for y.
// Lanes of CS consecutive contact indices (CS divides 64) reduce by
// scs_group_extreme; a lane without a circle adds an empty box.
template <int CS>
__device__ __forceinline__ void scs_block_box(const CArgs& a, int c, bool valid, const float (&n)[3])
{
    const float ar = fabsf(n[2]);
    const float inf = __builtin_inff();
    float lx = valid ? n[0] - ar : inf, hx = valid ? n[0] + ar : -inf;
    float ly = valid ? n[1] - ar : inf, hy = valid ? n[1] + ar : -inf;
    float mg = valid ? fabsf(n[0]) + fabsf(n[1]) + ar : 0.0f;
    lx = scs_group_extreme<CS, false>(lx);
    hx = scs_group_extreme<CS, true>(hx);
    ly = scs_group_extreme<CS, false>(ly);
    hy = scs_group_extreme<CS, true>(hy);
    mg = scs_group_extreme<CS, true>(mg);
    if ((c % CS) == 0 && c < a.N) {
        const float d = 1e-5f * mg;
        a.bb[c / CS] = make_float4(lx - d, hx + d, ly - d, hy + d);
    }
}

// Whether tile (P, Qb) runs: its blocks' boxes meet, or its previous list
// holds pairs (their residual terms, and their prox if they still touch).
// A tile that does not run gets an empty new list (its mask may stay
// stale: a list of length 0 is read as empty).
// (The box, both counts are loaded together and combined without
// short-circuit operators: otherwise each is a load of its own behind a
// branch, and the test waits for them one after another.)
__device__ __forceinline__ bool scs_tile_runs(const CArgs& a, const float4 p, int P, int Qb)
{
    const int t = P * a.nbs - P * (P - 1) / 2 + (Qb - P);
    const long to = (long)a.par * a.nts + t, tn = (long)(a.par ^ 1) * a.nts + t;
    const float4 q = a.bb[Qb];
    const int c_old = a.cnt[to], c_new = a.cnt[tn];
    const bool meet = (p.x <= q.y) & (q.x <= p.y) & (p.z <= q.w) & (q.z <= p.w);
    const bool run = meet | (c_old > 0);
    if (!run & (c_new != 0)) a.cnt[tn] = 0;
    return run;
}

// The end of a run: as contact_zreduce_circle, with the run's Z in contact
// order (Zin = the last launch's Z_k, Zpin = the one before, or null: the
// canonical Z_prev) and the canonical arrays and residuals in circle order.
template <bool RES>
__device__ __forceinline__ void scs_zreduce_circle(const CArgs& a, const int c)
{
    constexpr int Q = RES ? CQ : 3;
    const int N = a.N;
    const int pc = a.perm[c];
    unsigned long long sa[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) sa[k] = a.Ac[(long)k * N + c];
    double s[Q];
#pragma unroll
    for (int q = 3; q < Q; ++q) s[q] = (double)(long long)a.Ac[(long)q * N + c] * (q == 7 ? CFIX2_INV : CFIX_INV);
    float zk[3], zp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zk[k] = a.Zin[3 * c + k];
        zp[k] = a.Zpin ? a.Zpin[3 * c + k] : a.Zp[3 * pc + k];
    }
    // every sum buffer is zero for a later id-order state
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
        a.Ac[(long)q * N + c] = 0ull;
        a.Acl[(long)q * N + c] = 0ull;
    }
    const float deg = (float)(N + 1);
    double p = 0.0, dz2 = 0.0;
    if constexpr (RES) p = s[6] + s[7];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float zn = zk[k] + csum_f32(sa[k]) / deg;
        if constexpr (RES) {
            const float dz = zn - zk[k];
            const double e = (double)(dz - (zk[k] - zp[k]));
            p += -2.0 * e * s[3 + k] + (double)deg * e * e;
            dz2 += (double)dz * dz;
        }
        a.Zpp[3 * pc + k] = zp[k];
        a.Zp[3 * pc + k] = zk[k];
        a.Z[3 * pc + k] = zn;
    }
    if constexpr (RES) {
        const float rho = a.uni[0];
        a.Rpc[pc] = p > 0.0 ? p : 0.0;
        a.Rd[pc] = (double)rho * rho * (double)deg * dz2;
    }
}

// contact tile of a pair of contact indices lo < hi
__device__ __forceinline__ void scs_tile_of(int lo, int hi, int CS, int nbs, int& t, int& idx)
{
    const int P = lo / CS, Qb = hi / CS;
    t = P * nbs - P * (P - 1) / 2 + (Qb - P);
    idx = (lo % CS) * CS + hi % CS;
}

// dense-layout slots of the edges of circles i and j in their pair factor
// (i != j): the edge of i first
__device__ __forceinline__ void scs_pair_slots(int i, int j, int nb, int T, long& qi, long& qj)
{
    const long TT = (long)T * T;
    if (i < j) {
        qi = dense_pair_slot(i, j, nb, T);
        qj = qi + TT;
    } else {
        qj = dense_pair_slot(j, i, nb, T);
        qi = qj + TT;
    }
}

}  // namespace pdm
