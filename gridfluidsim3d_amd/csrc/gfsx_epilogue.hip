// gfsx_epilogue.hip -- FluidSimulation's stages around the transfer: the
// face fill after P2G and the velocity extrapolation, on gfx950.
//
// After the P2G gather has written the normalised values n and the weights
// W of every face (no body force yet):
//
//   1. k_fs_layers_init / k_fs_dilate: the cell layer grid of
//      MACVelocityField::extrapolateVelocityField (macvelocityfield.cpp:
//      577-664): fluid cells (a particle, not solid; fluidsimulation.cpp:
//      1990-2017) 0, then one 6-neighbour dilation per layer through
//      non-solid cells; solid cells are marked -2 so that no test matches
//      them.
//   2. k_fs_fill: FluidSimulation::_advectVelocityFieldU/V/W
//      (fluidsimulation.cpp:2597-2730): faces bordering fluid keep n where
//      W > 1e-9 (isValueSet, :2575-2585), else the FP64 average of their 26
//      neighbours (|n| > 0 for U, :2630; isValueSet for V, W, :2675,2720);
//      other faces 0.  The snapshot takes that field; the field gets the
//      body force on faces bordering fluid (:2766-2805, 3306).
//   3. k_fs_extrapolate, one launch per layer L: on both the snapshot and
//      the field (:3307, 3334), faces that border layer L, not layer L-1 and
//      no solid cell take the FP64 average of their 6 same-component
//      neighbours that border layer L-1, or 0 when that sum is 0
//      (macvelocityfield.cpp:616-760).  Within one layer the faces written
//      never border layer L-1 and the faces read always do, so the launch
//      reads only values of earlier layers, as the reference's in-place loop
//      does.
//
// Everything here is integer work or a fixed-order FP64 sum per face:
// results are bit-identical to the sequential restatement.

#include "gfsx_internal.h"

#include <algorithm>

namespace gfsx {

namespace {

constexpr int8_t kSolid = -2;

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>((n + kBlock - 1) / kBlock, 1);
}

__device__ __forceinline__ int8_t layer_at(const FsEpilogue &e, int i, int j, int k) {
    return e.layer[(uint32_t)i + (uint32_t)e.ni * ((uint32_t)j + (uint32_t)e.nj * (uint32_t)k)];
}

// the face of component c at (i,j,k) borders a cell whose layer is v
// (FluidMaterialGrid::isFaceBorderingMaterialU/V/W, fluidmaterialgrid.cpp:119-143)
__device__ __forceinline__ bool face_borders(const FsEpilogue &e, int c, int i, int j, int k,
                                             int v) {
    const int a = c == 0 ? i : (c == 1 ? j : k);
    const int na = c == 0 ? e.ni : (c == 1 ? e.nj : e.nk);
    const int li = i - (c == 0), lj = j - (c == 1), lk = k - (c == 2);
    if (a == na) return layer_at(e, li, lj, lk) == v;
    if (a > 0) return layer_at(e, i, j, k) == v || layer_at(e, li, lj, lk) == v;
    return layer_at(e, i, j, k) == v;
}

// component and index of flattened face f (u faces, then v, then w)
__device__ __forceinline__ void face_of(const FsEpilogue &e, int64_t f, int &c, int &i, int &j,
                                        int &k, int d[3]) {
    const int64_t f0 = (int64_t)(e.ni + 1) * e.nj * e.nk;
    const int64_t f1 = (int64_t)e.ni * (e.nj + 1) * e.nk;
    c = f < f0 ? 0 : (f < f0 + f1 ? 1 : 2);
    const int64_t l = c == 0 ? f : (c == 1 ? f - f0 : f - f0 - f1);
    d[0] = e.ni + (c == 0);
    d[1] = e.nj + (c == 1);
    d[2] = e.nk + (c == 2);
    i = (int)(l % d[0]);
    j = (int)((l / d[0]) % d[1]);
    k = (int)(l / ((int64_t)d[0] * d[1]));
}

__global__ __launch_bounds__(kBlock) void k_fs_layers_init(FsEpilogue e) {
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t ncell = (int64_t)e.ni * e.nj * e.nk;
    if (t >= ncell) return;
    const int i = (int)(t % e.ni), j = (int)((t / e.ni) % e.nj), k = (int)(t / ((int64_t)e.ni * e.nj));
    int8_t v;
    if (e.solid && e.solid[t]) {
        v = kSolid;
    } else {
        // the cell's bin in the sort's padded bin grid
        const uint32_t b = (uint32_t)(i + 1) + (uint32_t)e.b[0] *
                           ((uint32_t)(j + 1) + (uint32_t)e.b[1] * (uint32_t)(k + 1));
        v = e.start[b + 1] > e.start[b] ? 0 : -1;
    }
    e.layer[t] = v;
}

__global__ __launch_bounds__(kBlock) void k_fs_dilate(FsEpilogue e, int L) {
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t ncell = (int64_t)e.ni * e.nj * e.nk;
    if (t >= ncell) return;
    if (e.layer[t] != -1) return;
    const int i = (int)(t % e.ni), j = (int)((t / e.ni) % e.nj), k = (int)(t / ((int64_t)e.ni * e.nj));
    const int8_t src = (int8_t)(L - 1);
    // neighbours set in this launch hold L, never L - 1, whichever value is read
    bool hit = (i > 0 && layer_at(e, i - 1, j, k) == src) ||
               (i + 1 < e.ni && layer_at(e, i + 1, j, k) == src) ||
               (j > 0 && layer_at(e, i, j - 1, k) == src) ||
               (j + 1 < e.nj && layer_at(e, i, j + 1, k) == src) ||
               (k > 0 && layer_at(e, i, j, k - 1) == src) ||
               (k + 1 < e.nk && layer_at(e, i, j, k + 1) == src);
    if (hit) e.layer[t] = (int8_t)L;
}

__global__ __launch_bounds__(kBlock) void k_fs_fill(FsEpilogue e, int64_t nface) {
    const int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (f >= nface) return;
    int c, i, j, k, d[3];
    face_of(e, f, c, i, j, k, d);
    const uint32_t o = (uint32_t)i + (uint32_t)d[0] * ((uint32_t)j + (uint32_t)d[1] * (uint32_t)k);
    float *grid = c == 0 ? e.grid[0] : (c == 1 ? e.grid[1] : e.grid[2]);
    float *old = c == 0 ? e.old[0] : (c == 1 ? e.old[1] : e.old[2]);
    if (!face_borders(e, c, i, j, k, 0)) {
        grid[o] = 0.0f;
        old[o] = 0.0f;
        return;
    }
    const float *n = c == 0 ? e.n[0] : (c == 1 ? e.n[1] : e.n[2]);
    const float *w = c == 0 ? e.w[0] : (c == 1 ? e.w[1] : e.w[2]);
    float val;
    if ((double)w[o] > 1e-9) {
        val = n[o];
    } else {
        double avg = 0.0, weight = 0.0;
        for (int nk = k - 1; nk <= k + 1; nk++)
            for (int nj = j - 1; nj <= j + 1; nj++)
                for (int ni = i - 1; ni <= i + 1; ni++) {
                    if (ni == i && nj == j && nk == k) continue;
                    if ((unsigned)ni >= (unsigned)d[0] || (unsigned)nj >= (unsigned)d[1] ||
                        (unsigned)nk >= (unsigned)d[2])
                        continue;
                    const uint32_t g = (uint32_t)ni + (uint32_t)d[0] * ((uint32_t)nj + (uint32_t)d[1] * (uint32_t)nk);
                    const float nv = n[g];
                    const bool use = c == 0 ? fabsf(nv) > 0.0f : (double)w[g] > 1e-9;
                    if (use) {
                        avg += (double)nv;
                        weight += 1.0;
                    }
                }
        val = weight > 0.0 ? (float)(avg / weight) : 0.0f;
    }
    old[o] = val;
    const bool on = c == 0 ? e.grav_on[0] : (c == 1 ? e.grav_on[1] : e.grav_on[2]);
    const float gd = c == 0 ? e.grav_delta[0] : (c == 1 ? e.grav_delta[1] : e.grav_delta[2]);
    grid[o] = on ? val + gd : val;
}

__global__ __launch_bounds__(kBlock) void k_fs_extrapolate(FsEpilogue e, int64_t nface, int L) {
    const int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (f >= nface) return;
    int c, i, j, k, d[3];
    face_of(e, f, c, i, j, k, d);
    if (!(face_borders(e, c, i, j, k, L) && !face_borders(e, c, i, j, k, L - 1) &&
          !face_borders(e, c, i, j, k, kSolid)))
        return;
    float *grid = c == 0 ? e.grid[0] : (c == 1 ? e.grid[1] : e.grid[2]);
    float *old = c == 0 ? e.old[0] : (c == 1 ? e.old[1] : e.old[2]);
    double sg = 0.0, so = 0.0, cnt = 0.0;
    const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                          {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    for (int q = 0; q < 6; q++) {
        const int a = nb[q][0], b = nb[q][1], cc = nb[q][2];
        if ((unsigned)a >= (unsigned)d[0] || (unsigned)b >= (unsigned)d[1] ||
            (unsigned)cc >= (unsigned)d[2])
            continue;
        if (!face_borders(e, c, a, b, cc, L - 1)) continue;
        const uint32_t g = (uint32_t)a + (uint32_t)d[0] * ((uint32_t)b + (uint32_t)d[1] * (uint32_t)cc);
        sg += (double)grid[g];
        so += (double)old[g];
        cnt += 1.0;
    }
    const uint32_t o = (uint32_t)i + (uint32_t)d[0] * ((uint32_t)j + (uint32_t)d[1] * (uint32_t)k);
    grid[o] = sg == 0.0 ? 0.0f : (float)(sg / cnt);
    old[o] = so == 0.0 ? 0.0f : (float)(so / cnt);
}

}  // namespace

void fs_epilogue(const FsEpilogue &e, hipStream_t s) {
    const int64_t ncell = (int64_t)e.ni * e.nj * e.nk;
    const int64_t nface = (int64_t)(e.ni + 1) * e.nj * e.nk + (int64_t)e.ni * (e.nj + 1) * e.nk +
                          (int64_t)e.ni * e.nj * (e.nk + 1);
    hipLaunchKernelGGL(k_fs_layers_init, dim3(blocks_for(ncell)), dim3(kBlock), 0, s, e);
    GFSX_LAUNCH_CHECK();
    for (int L = 1; L <= e.layers; L++) {
        hipLaunchKernelGGL(k_fs_dilate, dim3(blocks_for(ncell)), dim3(kBlock), 0, s, e, L);
        GFSX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_fs_fill, dim3(blocks_for(nface)), dim3(kBlock), 0, s, e, nface);
    GFSX_LAUNCH_CHECK();
    for (int L = 1; L <= e.layers; L++) {
        hipLaunchKernelGGL(k_fs_extrapolate, dim3(blocks_for(nface)), dim3(kBlock), 0, s, e,
                           nface, L);
        GFSX_LAUNCH_CHECK();
    }
}

}  // namespace gfsx
