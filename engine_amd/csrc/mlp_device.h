// mlp_device.h -- device side of the learned rho network, shared by the
// training / policy kernels (mlp.hip) and the batched solver (batch.hip).
//
// 6 -> 32 -> 32 -> 1 ReLU MLP on bf16 MFMA (v_mfma_f32_16x16x16_bf16) with
// fp32 accumulation.  MFMA 16x16x16 operand layout (wave64, 4 elements per
// lane):
//   A[i][k]: lane i + 16*(k/4), element k%4      (i row, k reduction)
//   B[k][j]: lane j + 16*(k/4), element k%4      (j column)
//   D[i][j]: lane j + 16*(i/4), element i%4
// A wave pushes tiles of 16 rows through the network; activations move from
// the D layout of one layer to the A layout of the next through LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <stdint.h>

#include "mlp.h"

namespace pdm {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round-to-nearest-even fp32 -> bf16 (NaN stays NaN): the gfx950
// conversion instruction (v_cvt_pk_bf16_f32)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, __float2bfloat16(f)); }
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Weights staged in LDS by every block: bf16 operands and fp32 biases.
struct MlpShared {
    short w1[MLP_NH][16];  // K padded 6 -> 16
    short w2[MLP_NH][MLP_NH];
    short w3[MLP_NH];
    float b1[MLP_NH], b2[MLP_NH], b3;
};

__device__ void stage_weights(const float* __restrict__ p, MlpShared& w)
{
    for (int q = threadIdx.x; q < MLP_NH * 16; q += blockDim.x) {
        const int u = q / 16, k = q % 16;
        w.w1[u][k] = k < MLP_NIN ? f2bf(p[MLP_OFF_W1 + u * MLP_NIN + k]) : (short)0;
    }
    for (int q = threadIdx.x; q < MLP_NH * MLP_NH; q += blockDim.x) w.w2[q / MLP_NH][q % MLP_NH] = f2bf(p[MLP_OFF_W2 + q]);
    for (int q = threadIdx.x; q < MLP_NH; q += blockDim.x) {
        w.w3[q] = f2bf(p[MLP_OFF_W3 + q]);
        w.b1[q] = p[MLP_OFF_B1 + q];
        w.b2[q] = p[MLP_OFF_B2 + q];
    }
    if (threadIdx.x == 0) w.b3 = p[MLP_OFF_B3];
}

// Per-wave scratch: one 16-row tile.
struct TileScratch {
    short f[16][16];     // features (bf16), K padded
    short h1[16][MLP_NH];
    short h2[MLP_NH > 16 ? 16 : 16][MLP_NH];
    float y[16];
    short dz[16][MLP_NH];  // back-propagated pre-activation gradient (bf16)
};

// Forward of one 16-row tile whose features are in ts.f.  Returns in z1, z2
// (D layout, two 16-column tiles each) the pre-activations; ts.h1, ts.h2 get
// the bf16 activations; ts.y the outputs (written by lanes with l%16 == 0).
__device__ __forceinline__ void tile_forward(const MlpShared& w, TileScratch& ts, v4f (&z1)[2], v4f (&z2)[2])
{
    const int l = threadIdx.x & 63;
    const int li = l & 15, lq = l >> 4;
    // layer 1: A = features [16 x 16], B = W1^T
    v4s a;
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] = ts.f[li][4 * lq + e];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        v4s b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = w.w1[nt * 16 + li][4 * lq + e];
        z1[nt] = mfma16(a, b, v4f{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            z1[nt][e] += w.b1[nt * 16 + li];
            ts.h1[4 * lq + e][nt * 16 + li] = f2bf(fmaxf(z1[nt][e], 0.0f));
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 2: A = h1 [16 x 32] (two K steps), B = W2^T
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            v4s aa, bb;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                aa[e] = ts.h1[li][ks * 16 + 4 * lq + e];
                bb[e] = w.w2[nt * 16 + li][ks * 16 + 4 * lq + e];
            }
            acc = mfma16(aa, bb, acc);
        }
        z2[nt] = acc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            z2[nt][e] += w.b2[nt * 16 + li];
            ts.h2[4 * lq + e][nt * 16 + li] = f2bf(fmaxf(z2[nt][e], 0.0f));
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 3: A = h2, B = W3^T in column 0
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        v4s aa, bb;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            aa[e] = ts.h2[li][ks * 16 + 4 * lq + e];
            bb[e] = li == 0 ? w.w3[ks * 16 + 4 * lq + e] : (short)0;
        }
        acc = mfma16(aa, bb, acc);
    }
    if (li == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) ts.y[4 * lq + e] = acc[e] + w.b3;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

// Features of one edge (shared by the training-data extraction and the
// in-solver policy): log10 of the primal residual |X_e - Z_v|, of the
// rho-weighted dual residual rho_e |Z_v - Zprev_v| and of rho_e, scaled to
// O(1); the factor type as two indicator slots (pair, radius; the container
// is 0,0); the fraction of the iteration budget done.
__device__ __forceinline__ void edge_features(float r, float s, float rho, int type, float frac, float (&f)[MLP_NIN])
{
    f[0] = log10f(fmaxf(r, 1e-10f)) * 0.2f;
    f[1] = log10f(fmaxf(s, 1e-10f)) * 0.2f;
    f[2] = log10f(fmaxf(rho, 1e-10f)) * 0.5f;
    f[3] = type == PDM_EDGE_PAIR ? 1.0f : 0.0f;
    f[4] = type == PDM_EDGE_RADIUS ? 1.0f : 0.0f;
    f[5] = frac;
}

// Network output of 64 edges per wave-step (one per lane), pushed through
// the network as four 16-row MFMA tiles; y = clamped log10 rho per lane.
__device__ __forceinline__ float policy_wave_eval(const MlpShared& w, TileScratch& ts, float (&yv)[64], const float (&f)[MLP_NIN],
                                                  const int l)
{
    for (int t = 0; t < 4; ++t) {
        // rows of tile t are lanes 16t .. 16t+15
        if ((l >> 4) == t) {
#pragma unroll
            for (int k = 0; k < 16; ++k) ts.f[l & 15][k] = k < MLP_NIN ? f2bf(f[k]) : (short)0;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        v4f z1[2], z2[2];
        tile_forward(w, ts, z1, z2);
        if (l < 16) yv[16 * t + l] = ts.y[l];
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
    const float y = fminf(fmaxf(yv[l], MLP_LOG10_RHO_MIN), MLP_LOG10_RHO_MAX);
    __builtin_amdgcn_wave_barrier();
    return y;
}

}  // namespace pdm
