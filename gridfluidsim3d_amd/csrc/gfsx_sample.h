// gfsx_sample.h -- MAC-grid sampling shared by the forward G2P/RK kernels
// and their adjoints: cell coordinates, trilinear and tricubic samples with
// the reference's out-of-range rule, and the in-grid test.
#pragma once

#include "gfsx_internal.h"

namespace gfsx {

// Per-axis cell coordinates for the unshifted (a) and half-shifted (b)
// variants: u samples (a,b,b), v (b,a,b), w (b,b,a)
// (macvelocityfield.cpp:355-366,389-400,423-434).
struct AxisCoords {
    int ia, ib;
    float ta, tb;
};

// Cell index and fraction along one axis for the unshifted (a) and the
// half-shifted (b) samples, x >= 0.  With ga = x/dx and r = ga - floor(ga)
// (exact), gb = ga - 1/2 has floor(gb) = floor(ga) - (r < 1/2) and
// gb - floor(gb) = r - 1/2 or r + 1/2.  For ga >= 1/4 this is the FP64 value
// the reference computes (ga - 1/2 is exact there); below 1/4 its two
// roundings (ga - 1/2, then + 1) may differ from this one by an FP64 ulp,
// far under FP32 resolution.
__device__ __forceinline__ AxisCoords axis_coords(float x, double inv_dx) {
    AxisCoords c;
    const double ga = (double)x * inv_dx;
    const double fa = floor(ga);
    const double r = ga - fa;
    const bool low = r < 0.5;
    c.ia = (int)fa;
    c.ib = c.ia - (int)low;
    c.ta = (float)r;
    c.tb = (float)(low ? r + 0.5 : r - 0.5);
    return c;
}

// axis_coords for x >= 0 when inv_dx is a power of two, in FP32 and with
// the same bits.  x * inv_dx is then exact in FP32 (a change of exponent),
// so are floor(ga) and r = ga - floor(ga) (r keeps a subset of ga's bits,
// or is ga itself below 1), and ta = r as in FP64.  tb: r - 1/2 (r >= 1/2)
// is exact; r + 1/2 (r < 1/2) is one rounding of the exact sum in FP32, as
// (float)(r + 0.5) is one rounding of the exact FP64 sum.
__device__ __forceinline__ AxisCoords axis_coords_p2(float x, float inv_dx) {
    AxisCoords c;
    const float ga = x * inv_dx;
    const float fa = floorf(ga);
    const float r = ga - fa;
    const bool low = r < 0.5f;
    c.ia = (int)fa;
    c.ib = c.ia - (int)low;
    c.ta = r;
    c.tb = low ? r + 0.5f : r - 0.5f;
    return c;
}

// axis_coords, or axis_coords_p2 when P2 (the caller's guarantee: inv_dx is
// a power of two and x >= 0).
template <bool P2>
__device__ __forceinline__ AxisCoords axis_coords_t(float x, double inv_dx) {
    return P2 ? axis_coords_p2(x, (float)inv_dx) : axis_coords(x, inv_dx);
}

__device__ __forceinline__ float ld_or_zero(const float *__restrict__ f,
                                            bool ok, uint32_t idx) {
    float v = f[idx];
    return ok ? v : 0.0f;
}

// Trilinear sample of one component array with dims (w,h,d); out-of-range
// corners read 0 (macvelocityfield.cpp:470-478).
__device__ __forceinline__ float trilinear(const float *__restrict__ f, int w,
                                           int h, int d, int i, int j, int k,
                                           float tx, float ty, float tz) {
    const bool vi0 = (unsigned)i < (unsigned)w, vi1 = (unsigned)(i + 1) < (unsigned)w;
    const bool vj0 = (unsigned)j < (unsigned)h, vj1 = (unsigned)(j + 1) < (unsigned)h;
    const bool vk0 = (unsigned)k < (unsigned)d, vk1 = (unsigned)(k + 1) < (unsigned)d;
    const uint32_t i0 = (uint32_t)min(max(i, 0), w - 1), i1 = (uint32_t)min(max(i + 1, 0), w - 1);
    const uint32_t j0 = (uint32_t)min(max(j, 0), h - 1), j1 = (uint32_t)min(max(j + 1, 0), h - 1);
    const uint32_t k0 = (uint32_t)min(max(k, 0), d - 1), k1 = (uint32_t)min(max(k + 1, 0), d - 1);
    const uint32_t r00 = (uint32_t)w * (j0 + (uint32_t)h * k0);
    const uint32_t r10 = (uint32_t)w * (j1 + (uint32_t)h * k0);
    const uint32_t r01 = (uint32_t)w * (j0 + (uint32_t)h * k1);
    const uint32_t r11 = (uint32_t)w * (j1 + (uint32_t)h * k1);
    const float v000 = ld_or_zero(f, vi0 && vj0 && vk0, r00 + i0);
    const float v100 = ld_or_zero(f, vi1 && vj0 && vk0, r00 + i1);
    const float v010 = ld_or_zero(f, vi0 && vj1 && vk0, r10 + i0);
    const float v110 = ld_or_zero(f, vi1 && vj1 && vk0, r10 + i1);
    const float v001 = ld_or_zero(f, vi0 && vj0 && vk1, r01 + i0);
    const float v101 = ld_or_zero(f, vi1 && vj0 && vk1, r01 + i1);
    const float v011 = ld_or_zero(f, vi0 && vj1 && vk1, r11 + i0);
    const float v111 = ld_or_zero(f, vi1 && vj1 && vk1, r11 + i1);
    const float c00 = fmaf(tx, v100 - v000, v000);
    const float c10 = fmaf(tx, v110 - v010, v010);
    const float c01 = fmaf(tx, v101 - v001, v001);
    const float c11 = fmaf(tx, v111 - v011, v011);
    const float c0 = fmaf(ty, c10 - c00, c00);
    const float c1 = fmaf(ty, c11 - c01, c01);
    return fmaf(tz, c1 - c0, c0);
}

// The Catmull-Rom cubic (Keys a = -1/2) between p1 and p2 of
// Interpolation::cubicInterpolate (interpolation.cpp:44-46) as weights of
// p0..p3, its coefficients expanded: w0 = (-t + 2t^2 - t^3)/2,
// w1 = 1 - 5t^2/2 + 3t^3/2, w2 = (t + 4t^2 - 3t^3)/2, w3 = (t^3 - t^2)/2.
// A tricubic sample then costs 84 multiply-adds on top of three weight
// sets, instead of 21 cubic evaluations.
__device__ __forceinline__ void cubic_weights(float t, float w[4]) {
    const float t2 = t * t;
    w[0] = 0.5f * t * fmaf(2.0f - t, t, -1.0f);
    w[1] = fmaf(t2, fmaf(1.5f, t, -2.5f), 1.0f);
    w[2] = 0.5f * t * fmaf(t, fmaf(-3.0f, t, 4.0f), 1.0f);
    w[3] = 0.5f * t2 * (t - 1.0f);
}

// Sum over a 4x4x4 stencil of at(i,j,k) times the weights, x innermost, in
// one fixed order (every tricubic sampler uses it, so they agree bit for
// bit).
template <class F>
__device__ __forceinline__ float tricubic_sum(F &&at, float tx, float ty, float tz) {
    float wx[4], wy[4], wz[4];
    cubic_weights(tx, wx);
    cubic_weights(ty, wy);
    cubic_weights(tz, wz);
    float acc = 0.0f;
#pragma unroll
    for (int mk = 0; mk < 4; mk++) {
        float pl = 0.0f;
#pragma unroll
        for (int mj = 0; mj < 4; mj++) {
            float row = wx[0] * at(0, mj, mk);
#pragma unroll
            for (int mi = 1; mi < 4; mi++) row = fmaf(wx[mi], at(mi, mj, mk), row);
            pl = fmaf(wy[mj], row, pl);
        }
        acc = fmaf(wz[mk], pl, acc);
    }
    return acc;
}

// Tricubic sample over the 4x4x4 stencil at (i-1..i+2, j-1..j+2,
// k-1..k+2) (interpolation.cpp:26-42); samples outside the array read 0
// (macvelocityfield.cpp:372-378).
__device__ __forceinline__ float tricubic(const float *__restrict__ f, int w,
                                          int h, int d, int i, int j, int k,
                                          float tx, float ty, float tz) {
    uint32_t ic[4], jr[4], kr[4];
    bool vi[4], vj[4], vk[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int a = i - 1 + m, b = j - 1 + m, c = k - 1 + m;
        vi[m] = (unsigned)a < (unsigned)w;
        vj[m] = (unsigned)b < (unsigned)h;
        vk[m] = (unsigned)c < (unsigned)d;
        ic[m] = (uint32_t)min(max(a, 0), w - 1);
        jr[m] = (uint32_t)min(max(b, 0), h - 1);
        kr[m] = (uint32_t)min(max(c, 0), d - 1);
    }
    return tricubic_sum(
        [&](int mi, int mj, int mk) {
            const uint32_t row = (uint32_t)w * (jr[mj] + (uint32_t)h * kr[mk]);
            return ld_or_zero(f, vj[mj] && vk[mk] && vi[mi], row + ic[mi]);
        },
        tx, ty, tz);
}

template <int ST>
__device__ __forceinline__ float sample_comp(const float *__restrict__ f,
                                             int w, int h, int d, int i, int j,
                                             int k, float tx, float ty,
                                             float tz) {
    if (ST == 0) return trilinear(f, w, h, d, i, j, k, tx, ty, tz);
    return tricubic(f, w, h, d, i, j, k, tx, ty, tz);
}

// Velocity of grid g (and, when OLD, of the same-shaped grid o0,o1,o2) at
// (x,y,z): MACVelocityField::evaluateVelocityAtPosition
// {,Linear} (macvelocityfield.cpp:549-575), (0,0,0) outside the grid.
template <int ST, bool OLD>
__device__ __forceinline__ void sample_mac(const MacGrid &g, const float *o0,
                                           const float *o1, const float *o2,
                                           float x, float y, float z,
                                           float out[3], float out_old[3]) {
    if (!(x >= 0.0f && y >= 0.0f && z >= 0.0f && x < g.limf[0] && y < g.limf[1] &&
          z < g.limf[2])) {
        out[0] = out[1] = out[2] = 0.0f;
        if (OLD) out_old[0] = out_old[1] = out_old[2] = 0.0f;
        return;
    }
    const AxisCoords cx = axis_coords(x, g.inv_dx);
    const AxisCoords cy = axis_coords(y, g.inv_dx);
    const AxisCoords cz = axis_coords(z, g.inv_dx);
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    out[0] = sample_comp<ST>(g.f[0], ni + 1, nj, nk, cx.ia, cy.ib, cz.ib, cx.ta, cy.tb, cz.tb);
    out[1] = sample_comp<ST>(g.f[1], ni, nj + 1, nk, cx.ib, cy.ia, cz.ib, cx.tb, cy.ta, cz.tb);
    out[2] = sample_comp<ST>(g.f[2], ni, nj, nk + 1, cx.ib, cy.ib, cz.ia, cx.tb, cy.tb, cz.ta);
    if (OLD) {
        out_old[0] = sample_comp<ST>(o0, ni + 1, nj, nk, cx.ia, cy.ib, cz.ib, cx.ta, cy.tb, cz.tb);
        out_old[1] = sample_comp<ST>(o1, ni, nj + 1, nk, cx.ib, cy.ia, cz.ib, cx.tb, cy.ta, cz.tb);
        out_old[2] = sample_comp<ST>(o2, ni, nj, nk + 1, cx.ib, cy.ib, cz.ia, cx.tb, cy.tb, cz.ta);
    }
}

// ParticleAdvector::_validateOutput (particleadvector.cpp:1139-1149)
__device__ __forceinline__ void validate3(float v[3]) {
    if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2])))
        v[0] = v[1] = v[2] = 0.0f;
}

}  // namespace gfsx
