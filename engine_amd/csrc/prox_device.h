// prox_device.h -- fp32 device forms of the packing prox operators.
//
// Each operator is the per-factor proximal map that the reference calls
// through a function pointer from its X pass (core/paradmm.c:590,
// core/paradmm-gpu.cu:45).  Here they are inline device functions selected
// statically per factor kind, so the X pass has no indirect calls.
#pragma once
#include <hip/hip_runtime.h>

namespace pdm {

// r^e, evaluated exactly for the small integer exponents the radius
// objective produces (a = 1, 2 give e in {-1, 0, 1}); powf otherwise.
// Same values as powf there, including r = 0 (1/0 = inf) and e = 0 (1).
__device__ __forceinline__ float pow_exp(float r, float e)
{
    if (e == 0.0f) return 1.0f;
    if (e == 1.0f) return r;
    if (e == -1.0f) return 1.0f / r;
    if (e == 2.0f) return r * r;
    return powf(r, e);
}

// max_alpha_radius (packing/host.c:10-29): ten Newton steps on
// -r^(a-1) + (rho/a)(r - n) = 0 from |n/2|; a negative iterate returns 0.
__device__ __forceinline__ float max_alpha_radius(float n, float rho, float a)
{
#pragma clang fp contract(off)
    float r = fabsf(n * 0.5f);
#pragma unroll 1
    for (int it = 0; it < 10; ++it) {
        const float f = -pow_exp(r, a - 1.0f) + (rho / a) * (r - n);
        // the curvature term vanishes identically for a = 1 (the reference
        // evaluates 0 * pow(0, -1) = NaN at r = 0, packing/host.c:18)
        const float fp = (a == 1.0f ? 0.0f : -(a - 1.0f) * pow_exp(r, a - 2.0f)) + (rho / a);
        r = r - f / fp;
        if (r < 0.0f) return 0.0f;
    }
    return r;
}

// The operators are evaluated without contracting a * b + c into an FMA:
// every inlined copy (the contact kernel tests a pair in one place and
// projects it in another) then rounds the same way.
//
// the collision test of pair_project: |c1 - c2|^2 < (r1 + r2)^2
__device__ __forceinline__ bool pair_hits(float c1x, float c1y, float r1, float c2x, float c2y, float r2)
{
#pragma clang fp contract(off)
    const float dx = c1x - c2x, dy = c1y - c2y;
    const float d2 = fmaf(dx, dx, dy * dy);
    const float s = r1 + r2;
    return d2 < s * s;
}

// No-overlap projection shared by the dim-2 and dim-3 pair operators
// (packing/host.c:83-132).  Along the centre line circle 1 sits at 0 and
// circle 2 at dist; when dist^2 < (r1+r2)^2 the tuple (x1, x2, r1, r2) is
// projected with weights (wx1, wx2, wr1, wr2) onto x2 - x1 = r1 + r2:
//   lam = (dist - r1 - r2) / (1/wx1 + 1/wr1 + 1/wx2 + 1/wr2),
// centre 1 moves by lam/wx1 and centre 2 by -lam/wx2 along c2 - c1, radii
// by lam/wr.  Radii are not clamped.  Returns false when not overlapping.
__device__ __forceinline__ bool pair_project(float c1x, float c1y, float r1, float c2x, float c2y, float r2,
                                             float wx1, float wr1, float wx2, float wr2,
                                             float& o1x, float& o1y, float& or1,
                                             float& o2x, float& o2y, float& or2)
{
#pragma clang fp contract(off)
    if (!pair_hits(c1x, c1y, r1, c2x, c2y, r2)) return false;
    const float dx = c1x - c2x, dy = c1y - c2y;
    const float d2 = dx * dx + dy * dy;
    const float s = r1 + r2;
    const float dist = sqrtf(d2);
    const float ix1 = 1.0f / wx1, ir1 = 1.0f / wr1, ix2 = 1.0f / wx2, ir2 = 1.0f / wr2;
    const float lam = (dist - s) / (ix1 + ir1 + ix2 + ir2);
    // coincident centres have no centre line (the reference divides 0/0,
    // packing/host.c:123): they are separated along x, the circle of the
    // smaller radius towards -x (circle 1 on equal radii), whichever of the
    // two is listed first
    const bool same = !(dist > 0.0f);
    const float idist = same ? 0.0f : 1.0f / dist;
    const float ux = same ? (r2 < r1 ? -1.0f : 1.0f) : -dx * idist, uy = same ? 0.0f : -dy * idist;
    const float m1 = lam * ix1, m2 = lam * ix2;
    o1x = c1x + ux * m1;
    o1y = c1y + uy * m1;
    o2x = c2x - ux * m2;
    o2y = c2y - uy * m2;
    or1 = r1 + lam * ir1;
    or2 = r2 + lam * ir2;
    return true;
}

// disk_collision_variable_radius (packing/host.c:83-132): 4 edges of dim 2,
// (centre 1, radius 1, centre 2, radius 2); radii read from component 0 and
// written to both components.
__device__ __forceinline__ void prox_pair2(float (&x)[4][2], const float (&n)[4][2], const float (&rho)[4])
{
#pragma unroll
    for (int s = 0; s < 4; ++s) { x[s][0] = n[s][0]; x[s][1] = n[s][1]; }
    float a0, a1, r1, b0, b1, r2;
    if (pair_project(n[0][0], n[0][1], n[1][0], n[2][0], n[2][1], n[3][0], rho[0], rho[1], rho[2], rho[3],
                     a0, a1, r1, b0, b1, r2)) {
        x[0][0] = a0; x[0][1] = a1; x[1][0] = r1; x[1][1] = r1;
        x[2][0] = b0; x[2][1] = b1; x[3][0] = r2; x[3][1] = r2;
    }
}

// disk_plane_collision_variable_radius (packing/host.c:31-65): edge 0 the
// centre, edge 1 the radius; params = unit normal v, point p.  If
// v.(c-p) + r > 0 the rho-weighted projection onto v.(c-p) + r = 0 with the
// radius clamped at 0.
__device__ __forceinline__ void prox_wall2(float (&x)[2][2], const float (&n)[2][2], const float (&rho)[2],
                                           const float (&prm)[4])
{
    const float s = prm[0] * (n[0][0] - prm[2]) + prm[1] * (n[0][1] - prm[3]);
    if (s + n[1][0] > 0.0f) {
        float r = (n[1][0] * rho[1] - s * rho[0]) / (rho[0] + rho[1]);
        if (r < 0.0f) r = 0.0f;
        x[0][0] = n[0][0] - s * prm[0] - r * prm[0];
        x[0][1] = n[0][1] - s * prm[1] - r * prm[1];
        x[1][0] = r;
        x[1][1] = r;
    } else {
        x[0][0] = n[0][0]; x[0][1] = n[0][1]; x[1][0] = n[1][0]; x[1][1] = n[1][1];
    }
}

// maximize_sum_of_squared_disk_radii_alpha (packing/host.c:67-81), one edge
// of dim 2: radius from component 0 written to both.
__device__ __forceinline__ void prox_radius2(float (&x)[1][2], const float (&n)[1][2], const float (&rho)[1],
                                             const float (&prm)[1])
{
    const float r = max_alpha_radius(n[0][0], rho[0], prm[0]);
    x[0][0] = r;
    x[0][1] = r;
}

// dim-3 pair: edges (cx1,cy1,r1), (cx2,cy2,r2), one rho per circle.
__device__ __forceinline__ void prox_pair3(float (&x)[2][3], const float (&n)[2][3], const float (&rho)[2])
{
#pragma unroll
    for (int s = 0; s < 2; ++s) { x[s][0] = n[s][0]; x[s][1] = n[s][1]; x[s][2] = n[s][2]; }
    pair_project(n[0][0], n[0][1], n[0][2], n[1][0], n[1][1], n[1][2], rho[0], rho[0], rho[1], rho[1],
                 x[0][0], x[0][1], x[0][2], x[1][0], x[1][1], x[1][2]);
}

// Unit-square container (no reference function): Euclidean projection of
// (cx,cy,r) onto {r <= cx <= 1-r, r <= cy <= 1-r, r >= 0}.  For fixed r the
// centre is a clamp; r minimises (r-nr)^2 + (r-mx)+^2 + (r-my)+^2 with
// m = min(c, 1-c), whose stationary point is nr, (nr+a)/2 or (nr+a+b)/3
// (a <= b the sorted m's), clamped to [0, 1/2].
__device__ __forceinline__ void box3(float& cx, float& cy, float& r, float ncx, float ncy, float nr)
{
    const float mx = fminf(ncx, 1.0f - ncx), my = fminf(ncy, 1.0f - ncy);
    const float a = fminf(mx, my), b = fmaxf(mx, my);
    float t;
    if (nr <= a) t = nr;
    else if (0.5f * (nr + a) <= b) t = 0.5f * (nr + a);
    else t = (nr + a + b) * (1.0f / 3.0f);
    t = fminf(fmaxf(t, 0.0f), 0.5f);
    r = t;
    cx = fminf(fmaxf(ncx, t), 1.0f - t);
    cy = fminf(fmaxf(ncy, t), 1.0f - t);
}

__device__ __forceinline__ void prox_box3(float (&x)[1][3], const float (&n)[1][3])
{
    box3(x[0][0], x[0][1], x[0][2], n[0][0], n[0][1], n[0][2]);
}

// dim-3 radius operator: component 2 is the radius, the centre passes.
__device__ __forceinline__ void prox_radius3(float (&x)[1][3], const float (&n)[1][3], const float (&rho)[1],
                                             const float (&prm)[1])
{
    x[0][0] = n[0][0];
    x[0][1] = n[0][1];
    x[0][2] = max_alpha_radius(n[0][2], rho[0], prm[0]);
}

}  // namespace pdm
