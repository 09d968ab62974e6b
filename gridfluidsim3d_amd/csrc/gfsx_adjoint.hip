// gfsx_adjoint.hip -- gfx950 adjoints of the FLIP transfer + advection step.
//
// The plain step (trilinear stencil) maps the particles (x, v) of one scene
// to (x', v'):
//
//   S_f = sum_p K(|c_f - x_p|^2) v_p,   W_f = sum_p K(...)        (P2G)
//   n_f = S_f / W_f where W_f > 0, else 0                     (normalise)
//   g_f = n_f + m_f * body_force * dt   (m_f: face borders a cell holding
//                                        a particle; constant in x)
//   v'  = r_pic I(g, x) + r_flip (v + I(g, x) - I(n, x))       (PIC/FLIP)
//   x'  = RK4 of x through I(g, .)                              (advect)
//
// with I the MAC trilinear sample.  Given dL/dx' and dL/dv' the adjoint
// gives dL/dx and dL/dv in three passes over the particles:
//
//   1. k_adj_advect_brick (or k_adj_advect): recompute the RK4 stage
//      positions from (x, g), walk the stages backwards, and scatter the
//      sample adjoints onto the faces (dL/dg from the stages and the
//      PIC/FLIP blend, dL/dn from the FLIP difference).  The brick kernel
//      sums in 64-bit fixed point (deterministic); k_adj_advect adds floats
//      atomically (the order of the FP32 additions may change between runs).
//   2. k_adj_p2g_step: a gather over each particle's faces through the
//      quotient rule of the normalisation,
//        dL/dv_p += K_pf lam_f / W_f,
//        dL/dx_p += lam_f (v_p - n_f) / W_f * dK_pf/dx_p,
//      with lam_f = dL/dg_f + dL/dn_f (the body force is additive).  A pass
//      over the faces first forms A_f = lam_f / W_f and B_f = A_f n_f, so
//      the gather reads two arrays per face: lam_f (v_p - n_f) / W_f =
//      v_p A_f - B_f.
//   3. (scalar splat) k_adj_p2g_scalar: the adjoint of the unnormalised
//      S = sum K val, W = sum K of gfsx_p2g for any radius.
//
// The forward maths being differentiated follow ScalarField::addPointValue
// and applyWeightField (scalarfield.cpp:90-106,167-201),
// MACVelocityField::_interpolateLinearU/V/W (macvelocityfield.cpp:452-543),
// ParticleAdvector::_RK4 (particleadvector.cpp:1045-1054) and the PIC/FLIP
// update of FluidSimulation (fluidsimulation.cpp:3115-3127).  The fluid mask
// and the cell indices are piecewise constant in x: their derivative is zero
// almost everywhere, which is the derivative the adjoint returns.

#include "gfsx_internal.h"
#include "gfsx_sample.h"

#include <algorithm>

namespace gfsx {

namespace {

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>((n + kBlock - 1) / kBlock, 1);
}

// dK/dt of the Wyvill kernel K(t) = (1-t)^2 (5/9 + 4(1-t)/9), t = r^2/R^2:
// -a (10 + 12 a) / 9 with a = 1 - t, zero for t >= 1 (K is C1 at the cutoff).
__device__ __forceinline__ float wyvill_dt(float t) {
    const float a = fmaxf(1.0f - t, 0.0f);
    return -a * fmaf(12.0f / 9.0f, a, 10.0f / 9.0f);
}

// One MAC component's trilinear sample at cell (i,j,k), fractions t: adds
// lam * dI/dt to dt_out and hands lam * (corner weight) for every in-range
// corner to sink(corner i, j, k, flat index, value).  Out-of-range corners
// read 0 and receive nothing (macvelocityfield.cpp:470-478).
template <class Sink>
__device__ __forceinline__ void tri_adjoint(const float *__restrict__ f, int w, int h, int d,
                                            int i, int j, int k, float tx, float ty, float tz,
                                            float lam, float dt_out[3], Sink &&sink) {
    if (lam == 0.0f) return;
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
    const bool o000 = vi0 && vj0 && vk0, o100 = vi1 && vj0 && vk0;
    const bool o010 = vi0 && vj1 && vk0, o110 = vi1 && vj1 && vk0;
    const bool o001 = vi0 && vj0 && vk1, o101 = vi1 && vj0 && vk1;
    const bool o011 = vi0 && vj1 && vk1, o111 = vi1 && vj1 && vk1;
    const float v000 = ld_or_zero(f, o000, r00 + i0), v100 = ld_or_zero(f, o100, r00 + i1);
    const float v010 = ld_or_zero(f, o010, r10 + i0), v110 = ld_or_zero(f, o110, r10 + i1);
    const float v001 = ld_or_zero(f, o001, r01 + i0), v101 = ld_or_zero(f, o101, r01 + i1);
    const float v011 = ld_or_zero(f, o011, r11 + i0), v111 = ld_or_zero(f, o111, r11 + i1);
    const float sx = 1.0f - tx, sy = 1.0f - ty, sz = 1.0f - tz;
    dt_out[0] += lam * (sy * sz * (v100 - v000) + ty * sz * (v110 - v010) +
                        sy * tz * (v101 - v001) + ty * tz * (v111 - v011));
    dt_out[1] += lam * (sx * sz * (v010 - v000) + tx * sz * (v110 - v100) +
                        sx * tz * (v011 - v001) + tx * tz * (v111 - v101));
    dt_out[2] += lam * (sx * sy * (v001 - v000) + tx * sy * (v101 - v100) +
                        sx * ty * (v011 - v010) + tx * ty * (v111 - v110));
    if (o000) sink(i, j, k, r00 + i0, lam * (sx * sy * sz));
    if (o100) sink(i + 1, j, k, r00 + i1, lam * (tx * sy * sz));
    if (o010) sink(i, j + 1, k, r10 + i0, lam * (sx * ty * sz));
    if (o110) sink(i + 1, j + 1, k, r10 + i1, lam * (tx * ty * sz));
    if (o001) sink(i, j, k + 1, r01 + i0, lam * (sx * sy * tz));
    if (o101) sink(i + 1, j, k + 1, r01 + i1, lam * (tx * sy * tz));
    if (o011) sink(i, j + 1, k + 1, r11 + i0, lam * (sx * ty * tz));
    if (o111) sink(i + 1, j + 1, k + 1, r11 + i1, lam * (tx * ty * tz));
}

// Adjoint of the MAC sample I(f, q) at q: lq += (dI/dq)^T lam, and lam
// scattered through sink(array, i, j, k, flat, value) with array = base + c
// (c the component).  Zero outside the grid, where the sample is the
// constant 0 (macvelocityfield.cpp:550-552).
template <class Sink>
__device__ __forceinline__ void mac_adjoint(const StepAdjoint &a, const float *const f[3], int base,
                                            const float q[3], const float lam[3], float lq[3],
                                            Sink &&sink) {
    if (!(q[0] >= 0.0f && q[1] >= 0.0f && q[2] >= 0.0f && q[0] < a.mac.limf[0] &&
          q[1] < a.mac.limf[1] && q[2] < a.mac.limf[2]))
        return;
    const AxisCoords cx = axis_coords(q[0], a.mac.inv_dx);
    const AxisCoords cy = axis_coords(q[1], a.mac.inv_dx);
    const AxisCoords cz = axis_coords(q[2], a.mac.inv_dx);
    const int ni = a.mac.ni, nj = a.mac.nj, nk = a.mac.nk;
    float dt[3] = {0.0f, 0.0f, 0.0f};
    tri_adjoint(f[0], ni + 1, nj, nk, cx.ia, cy.ib, cz.ib, cx.ta, cy.tb, cz.tb, lam[0], dt,
                [&](int i, int j, int k, uint32_t o, float v) { sink(base, i, j, k, o, v); });
    tri_adjoint(f[1], ni, nj + 1, nk, cx.ib, cy.ia, cz.ib, cx.tb, cy.ta, cz.tb, lam[1], dt,
                [&](int i, int j, int k, uint32_t o, float v) { sink(base + 1, i, j, k, o, v); });
    tri_adjoint(f[2], ni, nj, nk + 1, cx.ib, cy.ib, cz.ia, cx.tb, cy.tb, cz.ta, lam[2], dt,
                [&](int i, int j, int k, uint32_t o, float v) { sink(base + 2, i, j, k, o, v); });
    // the cell fractions move by 1/dx per unit of x
    const float inv_dx = (float)a.mac.inv_dx;
    for (int c = 0; c < 3; c++) lq[c] += dt[c] * inv_dx;
}

// The RK4 + PIC/FLIP adjoint of one particle: from x0 and the output
// adjoints ax (of x') and av (of v'), returns lx (of x0).  The policy
// samples the grid after the body force (pol.sample(q, out)) and
// back-propagates a sample (pol.adjoint(base, q, lam, lq): base 0 for the
// grid after the body force, whose adjoint is dL/dg, 3 for the normalised
// P2G result, whose adjoint is dL/dn).
template <class Policy>
__device__ __forceinline__ void adj_particle(const StepAdjoint &a, const float x0[3],
                                             const float ax[3], const float av[3], float lx[3],
                                             Policy &&pol) {
    // forward stages, as the fused step computes them (particleadvector.cpp:1045-1054)
    float k1[3], k2[3], k3[3], q2[3], q3[3], q4[3];
    pol.sample(x0, k1);
    for (int c = 0; c < 3; c++) q2[c] = x0[c] + a.h_half * k1[c];
    pol.sample(q2, k2);
    for (int c = 0; c < 3; c++) q3[c] = x0[c] + a.h_half * k2[c];
    pol.sample(q3, k3);
    for (int c = 0; c < 3; c++) q4[c] = x0[c] + a.h_full * k3[c];
    // x' = x + sixth*(((k1 + 2k2) + 2k3) + k4), walked backwards
    float lk[3], lq[3];
    for (int c = 0; c < 3; c++) {
        lx[c] = ax[c];
        lk[c] = a.h_sixth * ax[c];               // k4
        lq[c] = 0.0f;
    }
    pol.adjoint(0, q4, lk, lq);
    for (int c = 0; c < 3; c++) {
        lx[c] += lq[c];
        lk[c] = 2.0f * a.h_sixth * ax[c] + a.h_full * lq[c];    // k3
        lq[c] = 0.0f;
    }
    pol.adjoint(0, q3, lk, lq);
    for (int c = 0; c < 3; c++) {
        lx[c] += lq[c];
        lk[c] = 2.0f * a.h_sixth * ax[c] + a.h_half * lq[c];    // k2
        lq[c] = 0.0f;
    }
    pol.adjoint(0, q2, lk, lq);
    // k1 and the PIC/FLIP samples share the point x: I(g, x) enters v' with
    // r_pic + r_flip, I(n, x) with -r_flip
    float lgx[3], lnx[3];
    for (int c = 0; c < 3; c++) {
        lx[c] += lq[c];
        lgx[c] = a.h_sixth * ax[c] + a.h_half * lq[c] + (a.r_pic + a.r_flip) * av[c];
        lnx[c] = -a.r_flip * av[c];
    }
    pol.adjoint(0, x0, lgx, lx);
    pol.adjoint(3, x0, lnx, lx);
}

// Samples and adjoints straight from and to global memory.
struct GlobalPolicy {
    const StepAdjoint &a;
    __device__ void sample(const float q[3], float out[3]) const {
        MacGrid g = a.mac;
        for (int c = 0; c < 3; c++) g.f[c] = a.g[c];
        float unused[3];
        sample_mac<0, false>(g, nullptr, nullptr, nullptr, q[0], q[1], q[2], out, unused);
    }
    __device__ void adjoint(int base, const float q[3], const float lam[3], float lq[3]) const {
        mac_adjoint(a, base == 0 ? a.g : a.old, base, q, lam, lq,
                    [&](int arr, int, int, int, uint32_t o, float v) {
                        unsafeAtomicAdd((arr < 3 ? a.lg[arr] : a.ln[arr - 3]) + o, v);
                    });
    }
};

// One thread per particle in input order; grid adjoints straight to global
// memory.
__global__ __launch_bounds__(kBlock) void k_adj_advect(StepAdjoint a) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= a.n) return;
    float x0[3], ax[3], av[3], lx[3];
    for (int c = 0; c < 3; c++) {
        x0[c] = a.x[c][p];
        ax[c] = a.ax[c][p];
        av[c] = a.av[c][p];
    }
    adj_particle(a, x0, ax, av, lx, GlobalPolicy{a});
    for (int c = 0; c < 3; c++) {
        a.gx[c][p] = lx[c];
        a.gv[c][p] = a.r_flip * av[c];
    }
}

// The P2G bricks (4x4x4 cells) of the forward step's brick sort, one
// work-group each, over the brick's particles (one contiguous range of
// sorted slots).  A tile of every face index within M
// cells of the brick's
// trilinear footprint holds the grid after the body force and the
// normalised P2G result (0 for faces outside the arrays, which is what an
// out-of-range sample reads).  Samples whose corners leave the tile (a stage
// moved more than M cells) use global memory.  Work-groups past the bricks
// take the particles of the overflow bin (outside the bin grid).
//
// Every sample's adjoint is computed, dL/dx and dL/dv are written, and
// lam * weight is added as a 64-bit fixed-point integer, scale
// 2^61 / (5 n B) with B a bound on |sample adjoint| (adj_bound), into an LDS
// tile (ds_add_u64; LDS float atomics run at a small fraction of the integer
// rate on gfx950) and then once per face into global accumulators.  Integer
// sums do not depend on the order of the adds, so the grid adjoints are the
// same on every run; no face can overflow: it receives at most 5n terms of
// magnitude at most 2^61 / (5n).
namespace abrick {
constexpr int BX = 4, BY = 4, BZ = 4;   // the P2G bricks of the forward's brick sort
constexpr int M = 1;
constexpr int TX = BX + 2 + 2 * M, TY = BY + 2 + 2 * M, TZ = BZ + 2 + 2 * M;
constexpr int TN = TX * TY * TZ;
constexpr int NEXTRA = 16;
}  // namespace abrick

// An upper bound B on |lam| over every sample adjoint of adj_particle, from
// bound[0] = max |dL/dx'|, bound[1] = max |dL/dv'| (over particles and
// components) and bound[2] = max |g| over the faces (float bits).  A
// trilinear sample's derivative along an axis is a convex combination of
// differences of two grid values, at most 2 max|g| per cell, so a stage's
// position adjoint is at most kappa = 3 * 2 max|g| / dx times its largest
// sample adjoint; the walk of adj_particle then gives, with h = dt and
// A, V the two maxima,
//   k4: h/6 A,   k3: 2h/6 A + h kappa L4,   k2: 2h/6 A + h/2 kappa L3,
//   x:  h/6 A + h/2 kappa L2 + |r_pic + r_flip| V,   n: |r_flip| V,
// and B is their maximum, widened by 1% for the FP32 rounding of the
// samples.  It needs no pass over the samples themselves.
__device__ __forceinline__ double adj_bound(const uint32_t *bound, const StepAdjoint &a) {
    const double A = __uint_as_float(bound[0]), V = __uint_as_float(bound[1]);
    const double kappa = 6.0 * (double)__uint_as_float(bound[2]) * a.mac.inv_dx;
    const double h6 = fabs((double)a.h_sixth), hh = fabs((double)a.h_half);
    const double hf = fabs((double)a.h_full);
    const double l4 = h6 * A;
    const double l3 = 2.0 * h6 * A + hf * kappa * l4;
    const double l2 = 2.0 * h6 * A + hh * kappa * l3;
    const double l1 = h6 * A + hh * kappa * l2 + fabs((double)(a.r_pic + a.r_flip)) * V;
    const double ln = fabs((double)a.r_flip) * V;
    return 1.01 * fmax(fmax(fmax(l4, l3), fmax(l2, l1)), ln);
}

// Fixed-point scale for adjoint terms of magnitude at most B, n particles:
// min(2^61 / (5n), 2^50) / B, so that a face's sum of at most 5n terms stays
// below 2^61 and every term below 2^50 (to_fixed's range).
__device__ __forceinline__ double fixed_scale(const uint32_t *bound, const StepAdjoint &a) {
    const double B = adj_bound(bound, a);
    if (!(B > 0.0) || !(B < 1.0e300)) return 1.0;
    const double cap = 2305843009213693952.0 / (5.0 * (double)(a.n > 0 ? a.n : 1));   // 2^61
    return fmin(cap, 1125899906842624.0) / B;   // 2^50
}

// round(v * scale) as a two's-complement 64-bit integer, for
// |v * scale| < 2^51: one FP64 FMA onto 1.5 * 2^52 rounds the product to an
// integer (the sum lies in [2^52, 2^53), where doubles are the integers), and
// there the bit patterns differ by the integers' difference.
__device__ __forceinline__ unsigned long long to_fixed(float v, double scale) {
    constexpr double kMagic = 6755399441055744.0;   // 1.5 * 2^52
    const double y = fma((double)v, scale, kMagic);
    return (unsigned long long)(__double_as_longlong(y) - __double_as_longlong(kMagic));
}

__global__ __launch_bounds__(kBlock) void k_adj_advect_brick(StepAdjoint a, AdjointSort so,
                                                             uint32_t nbrick) {
    using namespace abrick;
    __shared__ float gt[6 * TN];                 // g (0-2) and n (3-5) on the tile
    __shared__ unsigned long long at[6 * TN];    // fixed-point dL/dg, dL/dn
    const int tid = threadIdx.x;
    const uint32_t bid = blockIdx.x;
    const double scale = fixed_scale(so.bound, a);
    auto fixed_sink = [&](int arr, int, int, int, uint32_t o, float v) {
        atomicAdd(so.fx + (arr < 3 ? a.lg[arr] - a.lg[0] : a.ln[arr - 3] - a.lg[0]) + o,
                  to_fixed(v, scale));
    };
    // global path of one sample (outside the tile, or an overflow particle)
    auto global_adjoint = [&](int base, const float q[3], const float lam[3], float lq[3]) {
        mac_adjoint(a, base == 0 ? a.g : a.old, base, q, lam, lq, fixed_sink);
    };
    struct GlobalFixed {
        const GlobalPolicy &gp;
        decltype(global_adjoint) &ga;
        __device__ void sample(const float q[3], float out[3]) const { gp.sample(q, out); }
        __device__ void adjoint(int base, const float q[3], const float lam[3], float lq[3]) const {
            ga(base, q, lam, lq);
        }
    };
    const GlobalPolicy gpol{a};
    const GlobalFixed gfix{gpol, global_adjoint};
    auto finish = [&](uint32_t q, const float lx[3], const float av[3]) {
        const uint32_t p = so.perm[q];
        for (int c = 0; c < 3; c++) {
            a.gx[c][p] = lx[c];
            a.gv[c][p] = a.r_flip * av[c];
        }
    };
    if (bid >= nbrick) {
        // overflow bin: grid-stride over its slots
        const uint32_t q0 = so.pstart[so.nbb], q1 = so.pstart[so.nbb + 1];
        for (uint32_t q = q0 + (bid - nbrick) * kBlock + tid; q < q1; q += NEXTRA * kBlock) {
            const uint32_t p = so.perm[q];
            float x0[3], ax[3], av[3], lx[3];
            for (int c = 0; c < 3; c++) {
                x0[c] = so.xs[c][q];
                ax[c] = a.ax[c][p];
                av[c] = a.av[c][p];
            }
            adj_particle(a, x0, ax, av, lx, gfix);
            finish(q, lx, av);
        }
        return;
    }
    // brick bid: bin bid of the sort, one contiguous range of slots
    const uint32_t q0 = so.pstart[bid], total = so.pstart[bid + 1] - q0;
    if (total == 0) return;   // uniform across the work-group
    const uint32_t g = bid >> 3, sub = bid & 7u;
    const int bx = (int)(2 * (g % (uint32_t)so.gbx) + (sub & 1u));
    const int by = (int)(2 * ((g / (uint32_t)so.gbx) % (uint32_t)so.gby) + ((sub >> 1) & 1u));
    const int bz = (int)(2 * (g / ((uint32_t)so.gbx * (uint32_t)so.gby)) + (sub >> 2));
    // tile origin in face-index space: a particle in cell c touches faces
    // c - 1 .. c + 1 per axis, stages move up to M cells
    const int ox = bx * BX - 1 - M, oy = by * BY - 1 - M, oz = bz * BZ - 1 - M;
    const int ni = a.mac.ni, nj = a.mac.nj, nk = a.mac.nk;
    {
        // every global load of this thread's share goes out before the
        // first LDS store (one memory round trip per work-group)
        constexpr int PER = (6 * TN + kBlock - 1) / kBlock;
        float v[PER];
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = tid + m * kBlock;
            const int arr = e / TN, l = e - arr * TN;
            const int c = arr % 3;
            const int i = ox + l % TX, j = oy + (l / TX) % TY, k = oz + l / (TX * TY);
            const int w = ni + (c == 0), h = nj + (c == 1), d = nk + (c == 2);
            const bool in = e < 6 * TN && (unsigned)i < (unsigned)w && (unsigned)j < (unsigned)h &&
                            (unsigned)k < (unsigned)d;
            const uint32_t o =
                in ? (uint32_t)i + (uint32_t)w * ((uint32_t)j + (uint32_t)h * (uint32_t)k) : 0u;
            const float *src = c == 0 ? (arr < 3 ? a.g[0] : a.old[0])
                             : c == 1 ? (arr < 3 ? a.g[1] : a.old[1]) : (arr < 3 ? a.g[2] : a.old[2]);
            v[m] = gld(src, o);
            v[m] = in ? v[m] : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = tid + m * kBlock;
            if (e < 6 * TN) {
                gt[e] = v[m];
                at[e] = 0ull;
            }
        }
    }
    __syncthreads();

    // local flat index of lower corner (i,j,k), or -1 when a corner leaves the tile
    auto local = [&](int i, int j, int k) -> int {
        const int li = i - ox, lj = j - oy, lk = k - oz;
        return ((unsigned)li < (unsigned)(TX - 1) && (unsigned)lj < (unsigned)(TY - 1) &&
                (unsigned)lk < (unsigned)(TZ - 1)) ? li + TX * (lj + TY * lk) : -1;
    };
    auto in_grid = [&](const float q[3]) {
        return q[0] >= 0.0f && q[1] >= 0.0f && q[2] >= 0.0f && q[0] < a.mac.limf[0] &&
               q[1] < a.mac.limf[1] && q[2] < a.mac.limf[2];
    };
    // trilinear value from the tile, in the order of the forward sampler
    auto tile_value = [&](int arr, int l, float tx, float ty, float tz) {
        const float *f = gt + arr * TN + l;
        const float v000 = f[0], v100 = f[1], v010 = f[TX], v110 = f[TX + 1];
        const float v001 = f[TX * TY], v101 = f[TX * TY + 1];
        const float v011 = f[TX * TY + TX], v111 = f[TX * TY + TX + 1];
        const float c00 = fmaf(tx, v100 - v000, v000);
        const float c10 = fmaf(tx, v110 - v010, v010);
        const float c01 = fmaf(tx, v101 - v001, v001);
        const float c11 = fmaf(tx, v111 - v011, v011);
        const float c0 = fmaf(ty, c10 - c00, c00);
        const float c1 = fmaf(ty, c11 - c01, c01);
        return fmaf(tz, c1 - c0, c0);
    };
    auto tile_adjoint = [&](int arr, int l, float tx, float ty, float tz, float lam, float dt[3]) {
        if (lam == 0.0f) return;
        const float *f = gt + arr * TN + l;
        const float v000 = f[0], v100 = f[1], v010 = f[TX], v110 = f[TX + 1];
        const float v001 = f[TX * TY], v101 = f[TX * TY + 1];
        const float v011 = f[TX * TY + TX], v111 = f[TX * TY + TX + 1];
        const float sx = 1.0f - tx, sy = 1.0f - ty, sz = 1.0f - tz;
        dt[0] += lam * (sy * sz * (v100 - v000) + ty * sz * (v110 - v010) +
                        sy * tz * (v101 - v001) + ty * tz * (v111 - v011));
        dt[1] += lam * (sx * sz * (v010 - v000) + tx * sz * (v110 - v100) +
                        sx * tz * (v011 - v001) + tx * tz * (v111 - v101));
        dt[2] += lam * (sx * sy * (v001 - v000) + tx * sy * (v101 - v100) +
                        sx * ty * (v011 - v010) + tx * ty * (v111 - v110));
        unsigned long long *t = at + arr * TN + l;
        atomicAdd(t, to_fixed(lam * (sx * sy * sz), scale));
        atomicAdd(t + 1, to_fixed(lam * (tx * sy * sz), scale));
        atomicAdd(t + TX, to_fixed(lam * (sx * ty * sz), scale));
        atomicAdd(t + TX + 1, to_fixed(lam * (tx * ty * sz), scale));
        atomicAdd(t + TX * TY, to_fixed(lam * (sx * sy * tz), scale));
        atomicAdd(t + TX * TY + 1, to_fixed(lam * (tx * sy * tz), scale));
        atomicAdd(t + TX * TY + TX, to_fixed(lam * (sx * ty * tz), scale));
        atomicAdd(t + TX * TY + TX + 1, to_fixed(lam * (tx * ty * tz), scale));
    };
    struct TilePolicy {
        const StepAdjoint &a;
        const GlobalFixed &gf;
        decltype(local) &loc;
        decltype(in_grid) &ing;
        decltype(tile_value) &tv;
        decltype(tile_adjoint) &tadj;
        __device__ void sample(const float q[3], float out[3]) const {
            if (!ing(q)) {
                out[0] = out[1] = out[2] = 0.0f;
                return;
            }
            const AxisCoords cx = axis_coords(q[0], a.mac.inv_dx);
            const AxisCoords cy = axis_coords(q[1], a.mac.inv_dx);
            const AxisCoords cz = axis_coords(q[2], a.mac.inv_dx);
            const int l0 = loc(cx.ia, cy.ib, cz.ib), l1 = loc(cx.ib, cy.ia, cz.ib);
            const int l2 = loc(cx.ib, cy.ib, cz.ia);
            if (l0 < 0 || l1 < 0 || l2 < 0) {
                gf.sample(q, out);
                return;
            }
            out[0] = tv(0, l0, cx.ta, cy.tb, cz.tb);
            out[1] = tv(1, l1, cx.tb, cy.ta, cz.tb);
            out[2] = tv(2, l2, cx.tb, cy.tb, cz.ta);
        }
        __device__ void adjoint(int base, const float q[3], const float lam[3], float lq[3]) const {
            if (!ing(q)) return;
            const AxisCoords cx = axis_coords(q[0], a.mac.inv_dx);
            const AxisCoords cy = axis_coords(q[1], a.mac.inv_dx);
            const AxisCoords cz = axis_coords(q[2], a.mac.inv_dx);
            const int l0 = loc(cx.ia, cy.ib, cz.ib), l1 = loc(cx.ib, cy.ia, cz.ib);
            const int l2 = loc(cx.ib, cy.ib, cz.ia);
            if (l0 < 0 || l1 < 0 || l2 < 0) {
                gf.adjoint(base, q, lam, lq);
                return;
            }
            float dt[3] = {0.0f, 0.0f, 0.0f};
            tadj(base + 0, l0, cx.ta, cy.tb, cz.tb, lam[0], dt);
            tadj(base + 1, l1, cx.tb, cy.ta, cz.tb, lam[1], dt);
            tadj(base + 2, l2, cx.tb, cy.tb, cz.ta, lam[2], dt);
            const float inv_dx = (float)a.mac.inv_dx;
            for (int c = 0; c < 3; c++) lq[c] += dt[c] * inv_dx;
        }
    };
    const TilePolicy tpol{a, gfix, local, in_grid, tile_value, tile_adjoint};
    // blocked mapping: thread t takes items [t*per, (t+1)*per), so the lanes
    // of a wave sit in different cells (fewer same-address LDS atomics)
    const uint32_t per = (total + kBlock - 1) / kBlock;
    for (uint32_t it = tid * per, end = min(total, (uint32_t)(tid + 1) * per); it < end; it++) {
        const uint32_t q = q0 + it;
        const uint32_t p = so.perm[q];
        float x0[3], ax[3], av[3], lx[3];
        for (int c = 0; c < 3; c++) {
            x0[c] = so.xs[c][q];
            ax[c] = a.ax[c][p];
            av[c] = a.av[c][p];
        }
        adj_particle(a, x0, ax, av, lx, tpol);
        finish(q, lx, av);
    }
    __syncthreads();
    for (int e = tid; e < 6 * TN; e += kBlock) {
        const unsigned long long v = at[e];
        if (v == 0ull) continue;
        const int arr = e / TN, l = e - arr * TN;
        const int c = arr % 3;
        const int i = ox + l % TX, j = oy + (l / TX) % TY, k = oz + l / (TX * TY);
        const int w = ni + (c == 0), h = nj + (c == 1), d = nk + (c == 2);
        // tile entries outside the arrays collect nothing that exists
        if (!((unsigned)i < (unsigned)w && (unsigned)j < (unsigned)h && (unsigned)k < (unsigned)d))
            continue;
        const uint32_t o = (uint32_t)i + (uint32_t)w * ((uint32_t)j + (uint32_t)h * (uint32_t)k);
        const float *base = c == 0 ? (arr < 3 ? a.lg[0] : a.ln[0])
                          : c == 1 ? (arr < 3 ? a.lg[1] : a.ln[1]) : (arr < 3 ? a.lg[2] : a.ln[2]);
        atomicAdd(so.fx + (base - a.lg[0]) + o, v);
    }
}

struct FaceCount {
    int64_t nf[3], tot;   // faces per component, all faces
};

// The two factors the P2G adjoint gathers per face, from the face's grid
// adjoints lg = dL/dg, ln = dL/dn, weight w and normalised value nv:
// A = (lg + ln) / w and B = A nv where w > 0, else 0 (the face took no
// weight and its value was set, not normalised).
__device__ __forceinline__ void face_coeffs(float lg, float ln, float w, float nv, float &A,
                                            float &B) {
    const bool on = w > 0.0f;
    const float q = (lg + ln) * (1.0f / w);
    A = on ? q : 0.0f;
    B = on ? q * nv : 0.0f;
}

// the face of flat index f < tot (u faces, then v, then w): component, index
__device__ __forceinline__ int face_comp(int64_t f, const int64_t nf[3], int64_t &l) {
    const int c = f < nf[0] ? 0 : f < nf[0] + nf[1] ? 1 : 2;
    l = f - (c == 0 ? 0 : c == 1 ? nf[0] : nf[0] + nf[1]);
    return c;
}

// dL/dg, dL/dn from the fixed-point sums, turned into A and B (written over
// lg and ln) for one thread per face.
__global__ __launch_bounds__(kBlock) void k_fixed_to_coeffs(const unsigned long long *fx,
                                                            StepAdjoint a, FaceCount fc,
                                                            const uint32_t *bound) {
    const int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (f >= fc.tot) return;
    const double scale = fixed_scale(bound, a);
    const float lg = (float)((double)(long long)fx[f] / scale);
    const float ln = (float)((double)(long long)fx[fc.tot + f] / scale);
    int64_t l;
    const int c = face_comp(f, fc.nf, l);
    face_coeffs(lg, ln, a.w[c][l], a.old[c][l], a.lg[c][l], a.ln[c][l]);
}

// bound[0..2] = max |dL/dx'|, max |dL/dv'| over the particles and max |g|
// over the faces (adj_bound's inputs), as float bits: non-negative floats
// order like their bit patterns.  NaN entries are passed over, as fmaxf
// does.
__global__ __launch_bounds__(kBlock) void k_adj_bound(StepAdjoint a, FaceCount fc,
                                                      uint32_t *bound) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    float m[3] = {0.0f, 0.0f, 0.0f};
    for (int64_t p = i0; p < a.n; p += stride)
        for (int c = 0; c < 3; c++) {
            m[0] = fmaxf(m[0], fabsf(a.ax[c][p]));
            m[1] = fmaxf(m[1], fabsf(a.av[c][p]));
        }
    for (int64_t f = i0; f < fc.tot; f += stride) {
        int64_t l;
        const int c = face_comp(f, fc.nf, l);
        m[2] = fmaxf(m[2], fabsf(a.g[c][l]));
    }
    // one atomic per work-group and word at most, and none once the word
    // holds as much (same-address atomics serialise in L2)
    __shared__ float s_m[3][kBlock / 64];
    for (int k = 0; k < 3; k++) {
        float v = m[k];
        for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
        if ((threadIdx.x & 63) == 0) s_m[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        float v = s_m[k][0];
        for (int w = 1; w < kBlock / 64; w++) v = fmaxf(v, s_m[k][w]);
        const uint32_t bits = __float_as_uint(v);
        if (v > 0.0f && bits > __hip_atomic_load(bound + k, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(bound + k, bits);
    }
}

// the same from float grid adjoints (the atomic-add variant)
__global__ __launch_bounds__(kBlock) void k_face_coeffs(StepAdjoint a, FaceCount fc) {
    const int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (f >= fc.tot) return;
    int64_t l;
    const int c = face_comp(f, fc.nf, l);
    face_coeffs(a.lg[c][l], a.ln[c][l], a.w[c][l], a.old[c][l], a.lg[c][l], a.ln[c][l]);
}

// Faces of component c within the kernel radius R = dx of x: the node
// (i,j,k) of component c sits at (i,j,k)*dx + off_c in the frame where the
// reference splats (fluidsimulation.cpp:2526-2570).  A particle at cell
// coordinate s along an axis reaches nodes in (s - 1, s + 1): floor(s) and
// floor(s) + 1, 2x2x2 candidates per component; the kernel's support
// decides which of them take weight.
__global__ __launch_bounds__(kBlock) void k_adj_p2g_step(StepAdjoint a) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= a.n) return;
    const float dxf_half = a.half_dx;
    float lx[3] = {0.0f, 0.0f, 0.0f};
    float lv[3] = {0.0f, 0.0f, 0.0f};
    float xp[3], vp[3];
    for (int c = 0; c < 3; c++) {
        xp[c] = a.x[c][p];
        vp[c] = a.v[c][p];
    }
    const double dx = a.mac.dx;
    for (int c = 0; c < 3; c++) {
        const int dims[3] = {a.mac.ni + (c == 0), a.mac.nj + (c == 1), a.mac.nk + (c == 2)};
        float ps[3];
        int lo[3], hi[3];
        for (int ax = 0; ax < 3; ax++) {
            // p -= offset (scalarfield.cpp:168)
            ps[ax] = xp[ax] - (ax == c ? 0.0f : dxf_half);
            // the nodes within R = dx of ps along this axis are floor(ps/dx)
            // and the one after it (a node exactly dx away has t = 1 and
            // takes no weight)
            const double s = floor((double)ps[ax] * a.mac.inv_dx);
            // false for NaN; when true, s lies in [-1, dims - 1] and converts safely
            const bool any = s + 1.0 >= 0.0 && s <= (double)dims[ax] - 1.0;
            lo[ax] = any ? (int)fmax(s, 0.0) : 0;
            hi[ax] = any ? (int)fmin(s + 1.0, (double)dims[ax] - 1.0) : -1;
        }
        const float *__restrict__ A = a.lg[c];   // per-face coefficients (face_coeffs)
        const float *__restrict__ B = a.ln[c];
        for (int k = lo[2]; k <= hi[2]; k++) {
            const float gz = (float)((float)k * dx);
            for (int j = lo[1]; j <= hi[1]; j++) {
                const float gy = (float)((float)j * dx);
                for (int i = lo[0]; i <= hi[0]; i++) {
                    const float gx = (float)((float)i * dx);
                    const float vx = gx - ps[0], vy = gy - ps[1], vz = gz - ps[2];
                    const float d2 = vx * vx + vy * vy + vz * vz;
                    const float t = d2 * a.inv_r2;
                    if (!(t < 1.0f)) continue;
                    const uint32_t f = (uint32_t)i + (uint32_t)dims[0] * ((uint32_t)j + (uint32_t)dims[1] * (uint32_t)k);
                    const float af = A[f], bf = B[f];
                    lv[c] += wyvill(t) * af;
                    // d(d2)/dx_p = -2 (g - p); lam (v_p - n_f) / W_f = v_p A - B
                    const float coef = fmaf(vp[c], af, -bf) * wyvill_dt(t) * a.inv_r2 * -2.0f;
                    lx[0] += coef * vx;
                    lx[1] += coef * vy;
                    lx[2] += coef * vz;
                }
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        a.gx[c][p] += lx[c];
        a.gv[c][p] += lv[c];
    }
}

struct ScalarAdjoint {
    const float *pos, *val;
    int64_t n;
    float off[3];
    double dx, inv_dx;
    int dims[3];
    int reach;      // ceil(R/dx) + 1
    float inv_r2;
    const float *gS, *gW;
    float *gpos, *gval;
};

__global__ __launch_bounds__(kBlock) void k_adj_p2g_scalar(ScalarAdjoint a) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= a.n) return;
    float ps[3];
    int lo[3], hi[3];
    for (int ax = 0; ax < 3; ax++) {
        ps[ax] = a.pos[3 * p + ax] - a.off[ax];
        const double s = floor((double)ps[ax] * a.inv_dx);
        const bool any = s + a.reach + 1.0 >= 0.0 && s - a.reach <= (double)a.dims[ax] - 1.0;
        lo[ax] = any ? (int)fmax(s - a.reach, 0.0) : 0;
        hi[ax] = any ? (int)fmin(s + a.reach + 1.0, (double)a.dims[ax] - 1.0) : -1;
    }
    const float val = a.val ? a.val[p] : 1.0f;
    float lx[3] = {0.0f, 0.0f, 0.0f};
    float lval = 0.0f;
    for (int k = lo[2]; k <= hi[2]; k++) {
        const float gz = (float)((float)k * a.dx);
        for (int j = lo[1]; j <= hi[1]; j++) {
            const float gy = (float)((float)j * a.dx);
            for (int i = lo[0]; i <= hi[0]; i++) {
                const float gx = (float)((float)i * a.dx);
                const float vx = gx - ps[0], vy = gy - ps[1], vz = gz - ps[2];
                const float t = (vx * vx + vy * vy + vz * vz) * a.inv_r2;
                if (!(t < 1.0f)) continue;
                const uint32_t f = (uint32_t)i + (uint32_t)a.dims[0] * ((uint32_t)j + (uint32_t)a.dims[1] * (uint32_t)k);
                const float gs = a.gS ? a.gS[f] : 0.0f;
                const float gw = a.gW ? a.gW[f] : 0.0f;
                lval += wyvill(t) * gs;
                const float coef = (gs * val + gw) * wyvill_dt(t) * a.inv_r2 * -2.0f;
                lx[0] += coef * vx;
                lx[1] += coef * vy;
                lx[2] += coef * vz;
            }
        }
    }
    for (int ax = 0; ax < 3; ax++) a.gpos[3 * p + ax] = lx[ax];
    if (a.gval) a.gval[p] = lval;
}

}  // namespace

void flip_step_adjoint(const StepAdjoint &a, const AdjointSort *so, hipStream_t s) {
    int64_t nf[3], tot = 0;
    for (int c = 0; c < 3; c++)
        tot += (nf[c] = (int64_t)(a.mac.ni + (c == 0)) * (a.mac.nj + (c == 1)) * (a.mac.nk + (c == 2)));
    FaceCount fc;
    for (int c = 0; c < 3; c++) fc.nf[c] = nf[c];
    fc.tot = tot;
    if (so) {
        // the grid adjoints are one block: u,v,w of dL/dg, then of dL/dn
        if (a.lg[1] != a.lg[0] + nf[0] || a.lg[2] != a.lg[1] + nf[1] ||
            a.ln[0] != a.lg[2] + nf[2] || a.ln[1] != a.ln[0] + nf[0] ||
            a.ln[2] != a.ln[1] + nf[1])
            throw Error("flip_step_adjoint: grid adjoint arrays must be contiguous");
        fill_u32((uint32_t *)so->fx, 0u, 4 * tot, s);
        fill_u32(so->bound, 0u, 3, s);
        // the fixed-point scale's bound, then one work-group per bin of the
        // brick sort and the overflow bin's
        hipLaunchKernelGGL(k_adj_bound, dim3((unsigned)std::min<int64_t>(
                               blocks_for(std::max<int64_t>(a.n, tot)), 1024)),
                           dim3(kBlock), 0, s, a, fc, so->bound);
        GFSX_LAUNCH_CHECK();
        if (a.n > 0) {
            const uint32_t nbrick = so->nbb;
            hipLaunchKernelGGL(k_adj_advect_brick, dim3(nbrick + abrick::NEXTRA), dim3(kBlock),
                               0, s, a, *so, nbrick);
            GFSX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_fixed_to_coeffs, dim3(blocks_for(tot)), dim3(kBlock), 0, s,
                           (const unsigned long long *)so->fx, a, fc,
                           (const uint32_t *)so->bound);
        GFSX_LAUNCH_CHECK();
    } else {
        for (int c = 0; c < 3; c++) {
            fill_f32(a.lg[c], 0.0f, nf[c], s);
            fill_f32(a.ln[c], 0.0f, nf[c], s);
        }
        if (a.n > 0) {
            hipLaunchKernelGGL(k_adj_advect, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a);
            GFSX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_face_coeffs, dim3(blocks_for(tot)), dim3(kBlock), 0, s, a, fc);
        GFSX_LAUNCH_CHECK();
    }
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_adj_p2g_step, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a);
    GFSX_LAUNCH_CHECK();
}

void p2g_adjoint(const float *pos, const float *val, int64_t n, double radius,
                 const float off[3], double dx, int ni, int nj, int nk, const float *gS,
                 const float *gW, float *gpos, float *gval, hipStream_t s) {
    if (n <= 0) return;
    ScalarAdjoint a;
    a.pos = pos;
    a.val = val;
    a.n = n;
    for (int ax = 0; ax < 3; ax++) a.off[ax] = off[ax];
    a.dx = dx;
    a.inv_dx = 1.0 / dx;
    a.dims[0] = ni;
    a.dims[1] = nj;
    a.dims[2] = nk;
    a.reach = (int)std::ceil(radius / dx) + 1;
    a.inv_r2 = (float)(1.0 / (radius * radius));
    a.gS = gS;
    a.gW = gW;
    a.gpos = gpos;
    a.gval = gval;
    hipLaunchKernelGGL(k_adj_p2g_scalar, dim3(blocks_for(n)), dim3(kBlock), 0, s, a);
    GFSX_LAUNCH_CHECK();
}

}  // namespace gfsx
