// gfsx_device.h -- device helpers shared by the kernels of the fused FLIP
// step: launch sizing, wave prefix sums, scene lookup in a batched scene
// table, register sorting networks, the G2P view of a scene with its LDS
// grid boxes and particle store, and the per-face outputs of P2G.
#pragma once

#include "gfsx_internal.h"
#include "gfsx_sample.h"

#include <algorithm>

namespace gfsx {

static inline unsigned grid_for(int64_t n, int64_t per_block) {
    int64_t b = (n + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(b, 1);
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive prefix of v over the 256-thread block; total receives the sum.
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v,
                                                        uint32_t *wsum,
                                                        uint32_t &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_sum(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
        uint32_t s = wsum[w];
        woff += (w < wid) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return woff + inc - v;
}

// Bitonic sort of the 64*K keys held by a wave, element e = lane*K + r in
// v[r]: pairs inside a lane swap registers, pairs across lanes exchange
// through a lane shuffle and keep the minimum or the maximum.
template <int K>
__device__ __forceinline__ void wave_bitonic(uint64_t (&v)[K]) {
    constexpr int N = 64 * K;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < K) {
#pragma unroll
                for (int r = 0; r < K; r++) {
                    const int l = r ^ j;
                    if (l > r) {
                        const int e = lane * K + r;
                        const bool up = (e & k) == 0;
                        const uint64_t a = v[r], b = v[l];
                        const bool sw = up ? (a > b) : (a < b);
                        v[r] = sw ? b : a;
                        v[l] = sw ? a : b;
                    }
                }
            } else {
                const int lx = j / K;
#pragma unroll
                for (int r = 0; r < K; r++) {
                    const int e = lane * K + r;
                    const uint32_t lo = __shfl_xor((uint32_t)v[r], lx, 64);
                    const uint32_t hi = __shfl_xor((uint32_t)(v[r] >> 32), lx, 64);
                    const uint64_t o = ((uint64_t)hi << 32) | lo;
                    const bool up = (e & k) == 0, lower = (e & j) == 0;
                    const bool keep_min = up == lower;
                    const uint64_t mn = v[r] < o ? v[r] : o, mx = v[r] < o ? o : v[r];
                    v[r] = keep_min ? mn : mx;
                }
            }
        }
    }
}

// Index of the scene that owns value v of a per-scene ascending offset
// (p0, b0, blk0, k0, kb0, g0 or f0): the last scene whose offset is <= v.
template <class F>
__device__ __forceinline__ int scene_of(const SceneDesc *__restrict__ sc, int nsc, uint64_t v,
                                        F offset) {
    int lo = 0, hi = nsc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((uint64_t)offset(sc[mid]) <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Bitonic sorting network over N keys in registers, ascending.
template <int N, class T>
__device__ __forceinline__ void bitonic_sort(T (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const T a = v[i], b = v[l];
                    const T lo = a < b ? a : b, hi = a < b ? b : a;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
        }
    }
}

struct RkConsts {
    float half, three_q, full, sixth, ninth;
};

// ParticleAdvector::_RK1.._RK4 (particleadvector.cpp:1045-1078) in FP32
// vec3 arithmetic; the step sizes are the reference's FP32 casts.
template <int ST>
__device__ __forceinline__ void rk_step(const MacGrid &g, int order,
                                        const RkConsts &c, const float x0[3],
                                        const float k1[3], float x1[3]) {
    float q[3], k2[3], k3[3], k4[3], unused[3];
    if (order == 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) x1[a] = x0[a] + c.full * k1[a];
        return;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) q[a] = x0[a] + c.half * k1[a];
    sample_mac<ST, false>(g, nullptr, nullptr, nullptr, q[0], q[1], q[2], k2, unused);
    if (order == 2) {
#pragma unroll
        for (int a = 0; a < 3; a++) x1[a] = x0[a] + c.full * k2[a];
        return;
    }
    if (order == 3) {
#pragma unroll
        for (int a = 0; a < 3; a++) q[a] = x0[a] + c.three_q * k2[a];
        sample_mac<ST, false>(g, nullptr, nullptr, nullptr, q[0], q[1], q[2], k3, unused);
#pragma unroll
        for (int a = 0; a < 3; a++)
            x1[a] = x0[a] + c.ninth * ((2.0f * k1[a] + 3.0f * k2[a]) + 4.0f * k3[a]);
        return;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) q[a] = x0[a] + c.half * k2[a];
    sample_mac<ST, false>(g, nullptr, nullptr, nullptr, q[0], q[1], q[2], k3, unused);
#pragma unroll
    for (int a = 0; a < 3; a++) q[a] = x0[a] + c.full * k3[a];
    sample_mac<ST, false>(g, nullptr, nullptr, nullptr, q[0], q[1], q[2], k4, unused);
#pragma unroll
    for (int a = 0; a < 3; a++)
        x1[a] = x0[a] + c.sixth * (((k1[a] + 2.0f * k2[a]) + 2.0f * k3[a]) + k4[a]);
}

// What the epilogue of a face of component c writes and reads of the scene:
// selected per lane from the three components' fields, which are uniform
// (scalar loads).  Indexing the scene's arrays with a per-lane component
// instead makes every field a dependent vector load from the scene table.
struct FaceOut {
    float *grid, *old, *wgt;
    int n0, n1;          // nodes of component c along i and j
    int nax;             // cells along axis c
    int grav_on;
    float grav_delta;
};

__device__ __forceinline__ FaceOut face_out(const StepP2G &a, int c) {
    FaceOut o;
    o.grid = c == 0 ? a.grid[0] : (c == 1 ? a.grid[1] : a.grid[2]);
    o.old = c == 0 ? a.grid_old[0] : (c == 1 ? a.grid_old[1] : a.grid_old[2]);
    o.wgt = c == 0 ? a.wgt[0] : (c == 1 ? a.wgt[1] : a.wgt[2]);
    o.n0 = a.ni + (c == 0);
    o.n1 = a.nj + (c == 1);
    o.nax = c == 0 ? a.ni : (c == 1 ? a.nj : a.nk);
    o.grav_on = c == 0 ? a.grav_on[0] : (c == 1 ? a.grav_on[1] : a.grav_on[2]);
    o.grav_delta = c == 0 ? a.grav_delta[0] : (c == 1 ? a.grav_delta[1] : a.grav_delta[2]);
    return o;
}

struct StepG2P {
    const float *xin[3], *vin[3];   // in sort order (all scenes), or the scene's inputs
    const uint32_t *gperm;          // when set: slot q holds input particle gperm[q] - p0
    float *xout[3], *vout[3];       // the scene's outputs
    const uint32_t *id_in;
    uint32_t *id_out;
    const float *old[3];
    RkConsts rk;
    float r_pic, r_flip;
    int keep_order;
    int64_t n;
    uint32_t p0;                    // the scene's first sort slot and particle index
};

__host__ __device__ inline StepG2P g2p_view(const SceneDesc &sd, const Sorted &so, int keep_order) {
    StepG2P b;
    b.gperm = so.perm;
    for (int c = 0; c < 3; c++) {
        b.xin[c] = so.perm ? sd.xin[c] : so.x[c];
        b.vin[c] = so.perm ? sd.vin[c] : so.v[c];
        b.xout[c] = sd.xout[c];
        b.vout[c] = sd.vout[c];
        b.old[c] = sd.grid_old[c];
    }
    b.id_in = sd.id_in;
    b.id_out = sd.id_out;
    b.rk.half = sd.h_half;
    b.rk.three_q = 0.0f;
    b.rk.full = sd.h_full;
    b.rk.sixth = sd.h_sixth;
    b.rk.ninth = 0.0f;
    b.r_pic = sd.r_pic;
    b.r_flip = sd.r_flip;
    b.keep_order = keep_order;
    b.n = sd.n;
    b.p0 = (uint32_t)sd.p0;
    return b;
}

__host__ __device__ inline MacGrid mac_view(const SceneDesc &sd) {
    MacGrid g;
    for (int c = 0; c < 3; c++) {
        g.f[c] = sd.grid[c];
        g.lim[c] = sd.lim[c];
        g.limf[c] = sd.limf[c];
    }
    g.ni = sd.ni;
    g.nj = sd.nj;
    g.nk = sd.nk;
    g.dx = sd.dx;
    g.inv_dx = 1.0 / sd.dx;
    return g;
}

// Write the new state of the particle in sort slot q: at q (sorted order)
// or at its original index perm[q] (kept order), with its id, all relative
// to the scene's first particle.
// Index into a.xin / a.vin of the particle in sort slot q.
__device__ __forceinline__ uint32_t particle_at(const StepG2P &a, uint32_t q) {
    return a.gperm ? gld(a.gperm, q) - a.p0 : q;
}

// Position and velocity of the particle in sort slot q.
__device__ __forceinline__ void load_particle(const StepG2P &a, uint32_t q, float x[3],
                                              float v[3]) {
    const uint32_t i = particle_at(a, q);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        x[c] = gld(a.xin[c], i);
        v[c] = gld(a.vin[c], i);
    }
}

__device__ __forceinline__ void store_particle(const StepG2P &a,
                                               const uint32_t *__restrict__ perm,
                                               uint32_t q, const float x1[3],
                                               const float vnew[3]) {
    uint32_t o = q - a.p0, p = o;
    if (a.keep_order || a.id_out) p = perm[q] - a.p0;
    if (a.keep_order) o = p;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        gst(a.xout[c], o, x1[c]);
        gst(a.vout[c], o, vnew[c]);
    }
    if (a.id_out) gst(a.id_out, o, a.id_in ? gld(a.id_in, p) : p);
}

namespace gbrick {
constexpr int GB = 8;        // cells per brick edge
constexpr int GH = 1;        // halo cells for RK stage samples
constexpr int NT = 256;      // threads per work-group
constexpr int MAXROW = (GB + 2) * (GB + 2);   // (y,z) bin rows, pads included
constexpr int NEXTRA = 64;   // work-groups for particles outside the bin grid
// Boxes of a brick of GBE^3 cells (GBE = GB by default).
template <int ST, int GBE = GB> struct Box {
    static constexpr int S = ST;                         // stencil reach below the cell: 0 or 1
    static constexpr int E = GBE + 2 + 2 * GH + 2 * S;   // new-field box edge
    static constexpr int LO = 1 + GH + S;                // new-field box origin: C0 - LO
    static constexpr int EO = GBE + 2 + 2 * S;           // snapshot box edge
    static constexpr int LOO = 1 + S;                    // snapshot box origin: C0 - LOO
};
static_assert(MAXROW <= 2 * 64, "two waves scan the rows");
}  // namespace gbrick

// Box sample of one component; b points at the stencil's base cell
// (i,j,k) in a box of edge E, and the stencil (i-S..i+1+S per axis) lies
// inside the box.
template <int ST, int E>
__device__ __forceinline__ float box_comp_at(const float *__restrict__ b, float tx, float ty,
                                             float tz) {
    typedef const __attribute__((address_space(3))) float LdsFloat;
    if (ST == 0) {
        // b as an opaque LDS base: the 8 taps are immediate offsets of at
        // most E^2 + E + 1 < 256 dwords from it (ds_read2's range), where
        // the compiler would otherwise fold the box's own LDS offset into
        // each read and add per read.  b points into an LDS box.
        static_assert(E * E + E + 1 < 256, "box too large for one base");
        LdsFloat *o = (LdsFloat *)b;
        __asm__("" : "+v"(o));
        const float v000 = o[0], v100 = o[1], v010 = o[E], v110 = o[E + 1];
        const float v001 = o[E * E], v101 = o[E * E + 1];
        const float v011 = o[E * E + E], v111 = o[E * E + E + 1];
        const float c00 = fmaf(tx, v100 - v000, v000);
        const float c10 = fmaf(tx, v110 - v010, v010);
        const float c01 = fmaf(tx, v101 - v001, v001);
        const float c11 = fmaf(tx, v111 - v011, v011);
        const float c0 = fmaf(ty, c10 - c00, c00);
        const float c1 = fmaf(ty, c11 - c01, c01);
        return fmaf(tz, c1 - c0, c0);
    }
    // Two LDS bases, the stencil's lowest entry and the one two planes up,
    // kept opaque to the compiler: every tap is then a non-negative
    // immediate offset of at most 2E^2 + 3E + 3 < 256 dwords from one of
    // them (ds_read2's offset range) instead of a negative or out-of-range
    // offset that costs an address add per read.  b points into an LDS box
    // (every caller's boxes are __shared__ arrays).
    static_assert(E * E + 3 * E + 3 < 256, "box too large for two plane bases");
    LdsFloat *o0 = (LdsFloat *)(b - (1 + E + E * E));
    LdsFloat *o2 = o0 + 2 * E * E;
    __asm__("" : "+v"(o0));
    __asm__("" : "+v"(o2));
    return tricubic_sum(
        [&](int mi, int mj, int mk) {
            return mk < 2 ? o0[mi + E * (mj + E * mk)] : o2[mi + E * (mj + E * (mk - 2))];
        },
        tx, ty, tz);
}

// The same with the base cell given relative to the box origin.
template <int ST, int E>
__device__ __forceinline__ float box_comp(const float *__restrict__ box, int i,
                                          int j, int k, float tx, float ty,
                                          float tz) {
    return box_comp_at<ST, E>(box + (i + E * (j + E * k)), tx, ty, tz);
}

// True when the stencils of all three components at cell coordinates
// (cx,cy,cz) lie in the box of edge E whose origin is at cell lo[].  Per
// axis the components use base cells ia (own axis) and ib <= ia (others),
// so the lowest cell read is ib - S and the highest ia + 1 + S.
template <int ST, int E>
__device__ __forceinline__ bool in_box(const AxisCoords &cx, const AxisCoords &cy,
                                       const AxisCoords &cz, int lx, int ly, int lz) {
    constexpr int S = ST;
    return (unsigned)(cx.ib - S - lx) < (unsigned)E && (unsigned)(cx.ia + 1 + S - lx) < (unsigned)E &&
           (unsigned)(cy.ib - S - ly) < (unsigned)E && (unsigned)(cy.ia + 1 + S - ly) < (unsigned)E &&
           (unsigned)(cz.ib - S - lz) < (unsigned)E && (unsigned)(cz.ia + 1 + S - lz) < (unsigned)E;
}

// sample_mac with the LDS boxes as the first choice.  cur[c] / old[c] are
// the staged boxes (edges E, EO; origins l*, lo*); g and o* are the global
// arrays used when a stencil leaves its box.  P2: g.inv_dx is a power of two
// (cell coordinates in FP32, the same bits).
template <int ST, bool OLD, int GBE = gbrick::GB, bool P2 = false>
__device__ __forceinline__ void brick_sample(const MacGrid &g, const float *o0,
                                             const float *o1, const float *o2,
                                             const float *__restrict__ cur0,
                                             const float *__restrict__ cur1,
                                             const float *__restrict__ cur2,
                                             const float *__restrict__ old0,
                                             const float *__restrict__ old1,
                                             const float *__restrict__ old2,
                                             int lx, int ly, int lz,
                                             int lox, int loy, int loz,
                                             float x, float y, float z,
                                             float out[3], float out_old[3]) {
    using B = gbrick::Box<ST, GBE>;
    if (!(x >= 0.0f && y >= 0.0f && z >= 0.0f && x < g.limf[0] && y < g.limf[1] &&
          z < g.limf[2])) {
        out[0] = out[1] = out[2] = 0.0f;
        if (OLD) out_old[0] = out_old[1] = out_old[2] = 0.0f;
        return;
    }
    const AxisCoords cx = axis_coords_t<P2>(x, g.inv_dx);
    const AxisCoords cy = axis_coords_t<P2>(y, g.inv_dx);
    const AxisCoords cz = axis_coords_t<P2>(z, g.inv_dx);
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    if (in_box<ST, B::E>(cx, cy, cz, lx, ly, lz)) {
        out[0] = box_comp<ST, B::E>(cur0, cx.ia - lx, cy.ib - ly, cz.ib - lz, cx.ta, cy.tb, cz.tb);
        out[1] = box_comp<ST, B::E>(cur1, cx.ib - lx, cy.ia - ly, cz.ib - lz, cx.tb, cy.ta, cz.tb);
        out[2] = box_comp<ST, B::E>(cur2, cx.ib - lx, cy.ib - ly, cz.ia - lz, cx.tb, cy.tb, cz.ta);
    } else {
        out[0] = sample_comp<ST>(g.f[0], ni + 1, nj, nk, cx.ia, cy.ib, cz.ib, cx.ta, cy.tb, cz.tb);
        out[1] = sample_comp<ST>(g.f[1], ni, nj + 1, nk, cx.ib, cy.ia, cz.ib, cx.tb, cy.ta, cz.tb);
        out[2] = sample_comp<ST>(g.f[2], ni, nj, nk + 1, cx.ib, cy.ib, cz.ia, cx.tb, cy.tb, cz.ta);
    }
    if (OLD) {
        if (in_box<ST, B::EO>(cx, cy, cz, lox, loy, loz)) {
            out_old[0] = box_comp<ST, B::EO>(old0, cx.ia - lox, cy.ib - loy, cz.ib - loz, cx.ta, cy.tb, cz.tb);
                out_old[1] = box_comp<ST, B::EO>(old1, cx.ib - lox, cy.ia - loy, cz.ib - loz, cx.tb, cy.ta, cz.tb);
                out_old[2] = box_comp<ST, B::EO>(old2, cx.ib - lox, cy.ib - loy, cz.ia - loz, cx.tb, cy.tb, cz.ta);
        } else {
            out_old[0] = sample_comp<ST>(o0, ni + 1, nj, nk, cx.ia, cy.ib, cz.ib, cx.ta, cy.tb, cz.tb);
            out_old[1] = sample_comp<ST>(o1, ni, nj + 1, nk, cx.ib, cy.ia, cz.ib, cx.tb, cy.ta, cz.tb);
            out_old[2] = sample_comp<ST>(o2, ni, nj, nk + 1, cx.ib, cy.ib, cz.ia, cx.tb, cy.tb, cz.ta);
        }
    }
}

// brick_sample for a position whose stencils are known to lie in the
// boxes: one base index per box from the half-shifted cells, the unshifted
// axis of each component one step further.  The same arithmetic as the
// in-box branch of brick_sample, so the results are the same bits.
// INGRID: the position is known to lie in the grid (no test).
template <int ST, bool OLD, int GBE = gbrick::GB, bool P2 = false, bool INGRID = false>
__device__ __forceinline__ void brick_sample_in(const MacGrid &g,
                                                const float *__restrict__ cur0,
                                                const float *__restrict__ cur1,
                                                const float *__restrict__ cur2,
                                                const float *__restrict__ old0,
                                                const float *__restrict__ old1,
                                                const float *__restrict__ old2,
                                                int lx, int ly, int lz,
                                                int lox, int loy, int loz,
                                                float x, float y, float z,
                                                float out[3], float out_old[3]) {
    using B = gbrick::Box<ST, GBE>;
    if (!INGRID && !(x >= 0.0f && y >= 0.0f && z >= 0.0f && x < g.limf[0] && y < g.limf[1] &&
                     z < g.limf[2])) {
        out[0] = out[1] = out[2] = 0.0f;
        if (OLD) out_old[0] = out_old[1] = out_old[2] = 0.0f;
        return;
    }
    const AxisCoords cx = axis_coords_t<P2>(x, g.inv_dx);
    const AxisCoords cy = axis_coords_t<P2>(y, g.inv_dx);
    const AxisCoords cz = axis_coords_t<P2>(z, g.inv_dx);
    const int dx = cx.ia - cx.ib, dy = cy.ia - cy.ib, dz = cz.ia - cz.ib;
    constexpr int E = B::E;
    const int b = (cx.ib - lx) + E * ((cy.ib - ly) + E * (cz.ib - lz));
    out[0] = box_comp_at<ST, E>(cur0 + b + dx, cx.ta, cy.tb, cz.tb);
    out[1] = box_comp_at<ST, E>(cur1 + b + E * dy, cx.tb, cy.ta, cz.tb);
    out[2] = box_comp_at<ST, E>(cur2 + b + E * E * dz, cx.tb, cy.tb, cz.ta);
    if (OLD) {
        constexpr int EO = B::EO;
        const int bo = (cx.ib - lox) + EO * ((cy.ib - loy) + EO * (cz.ib - loz));
        out_old[0] = box_comp_at<ST, EO>(old0 + bo + dx, cx.ta, cy.tb, cz.tb);
        out_old[1] = box_comp_at<ST, EO>(old1 + bo + EO * dy, cx.tb, cy.ta, cz.tb);
        out_old[2] = box_comp_at<ST, EO>(old2 + bo + EO * EO * dz, cx.tb, cy.tb, cz.ta);
    }
}

// One thread's share of a box of edge E with origin (ox,oy,oz) in a component
// array with dims (w,h,d): the loads are issued here, stored by
// BoxShare::store.  Cells outside the array are 0.
template <int E>
struct BoxShare {
    static constexpr int N = E * E * E;
    static constexpr int PER = (N + gbrick::NT - 1) / gbrick::NT;
    float v[PER];
    __device__ __forceinline__ void load(const float *__restrict__ f, int w, int h, int d, int ox,
                                         int oy, int oz) {
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = (int)threadIdx.x + m * gbrick::NT;
            const int ii = e % E, jj = (e / E) % E, kk = e / (E * E);
            const int gi = ox + ii, gj = oy + jj, gk = oz + kk;
            const bool in = e < N && (unsigned)gi < (unsigned)w && (unsigned)gj < (unsigned)h &&
                            (unsigned)gk < (unsigned)d;
            // an element outside the array reads element 0 and is replaced by 0
            const size_t idx =
                in ? (size_t)gi + (size_t)w * ((size_t)gj + (size_t)h * (size_t)gk) : 0;
            v[m] = gld(f, idx);
            v[m] = in ? v[m] : 0.0f;
        }
    }
    // Stores into the LDS box; returns the largest |value| stored (NaN if a
    // NaN was).
    __device__ __forceinline__ float store(float *__restrict__ box) const {
        float mx = 0.0f;
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = (int)threadIdx.x + m * gbrick::NT;
            if (e < N) {
                box[e] = v[m];
                // NaN stays: fmaxf would drop it
                mx = (fabsf(v[m]) > mx || v[m] != v[m]) ? fabsf(v[m]) : mx;
            }
        }
        return mx;
    }
};

// Copy the box of edge E with origin (ox,oy,oz) of a component array with
// dims (w,h,d) into LDS; cells outside the array are 0.  Returns the largest
// |value| this thread staged (NaN if it staged a NaN).  Every global load a
// thread makes goes out before the first LDS store.
template <int E>
__device__ __forceinline__ float stage_box(float *__restrict__ box,
                                           const float *__restrict__ f, int w,
                                           int h, int d, int ox, int oy, int oz) {
    BoxShare<E> b;
    b.load(f, w, h, d, ox, oy, oz);
    return b.store(box);
}

// The three component boxes of a MAC grid (dims (ni+1,nj,nk), (ni,nj+1,nk),
// (ni,nj,nk+1)) with one origin, every load of the three issued first: one
// memory round trip for all three.  m[c] receives stage_box's maximum.
template <int E>
__device__ __forceinline__ void stage_mac_boxes(float *const box[3], const float *const f[3],
                                                int ni, int nj, int nk, int ox, int oy, int oz,
                                                float m[3]) {
    BoxShare<E> b0, b1, b2;
    b0.load(f[0], ni + 1, nj, nk, ox, oy, oz);
    b1.load(f[1], ni, nj + 1, nk, ox, oy, oz);
    b2.load(f[2], ni, nj, nk + 1, ox, oy, oz);
    m[0] = b0.store(box[0]);
    m[1] = b1.store(box[1]);
    m[2] = b2.store(box[2]);
}

// PIC/FLIP blend and RK4 for one particle, given its new-field and
// snapshot velocities at the start position.
__device__ __forceinline__ void flip_blend(const StepG2P &a, const float v0[3],
                                           const float k1[3], const float vo_in[3],
                                           float vnew[3]) {
    float vn[3] = {k1[0], k1[1], k1[2]};
    float vo[3] = {vo_in[0], vo_in[1], vo_in[2]};
    validate3(vn);
    validate3(vo);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float vflip = (v0[c] + vn[c]) - vo[c];
        vnew[c] = a.r_pic * vn[c] + a.r_flip * vflip;
    }
}

}  // namespace gfsx
