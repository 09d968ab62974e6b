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
//   3. k_fs_classify: each face's extrapolation layer (at most one) and its
//      neighbours bordering the layer before; k_fs_extrapolate, one launch
//      per layer L that changes the faces of layer L only: on both the
//      snapshot and
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

// component and index of flattened face f (u faces, then v, then w);
// every face count is below 2^31, so 32-bit arithmetic is exact
__device__ __forceinline__ void face_of(const FsEpilogue &e, uint32_t f, int &c, int &i, int &j,
                                        int &k, int d[3]) {
    const uint32_t f0 = (uint32_t)(e.ni + 1) * (uint32_t)e.nj * (uint32_t)e.nk;
    const uint32_t f1 = (uint32_t)e.ni * (uint32_t)(e.nj + 1) * (uint32_t)e.nk;
    c = f < f0 ? 0 : (f < f0 + f1 ? 1 : 2);
    const uint32_t l = c == 0 ? f : (c == 1 ? f - f0 : f - f0 - f1);
    d[0] = e.ni + (c == 0);
    d[1] = e.nj + (c == 1);
    d[2] = e.nk + (c == 2);
    const uint32_t row = l / (uint32_t)d[0];
    i = (int)(l - row * (uint32_t)d[0]);
    j = (int)(row % (uint32_t)d[1]);
    k = (int)(row / (uint32_t)d[1]);
}

__device__ __forceinline__ void cell_of_index(const FsEpilogue &e, uint32_t t, int &i, int &j,
                                              int &k) {
    const uint32_t row = t / (uint32_t)e.ni;
    i = (int)(t - row * (uint32_t)e.ni);
    j = (int)(row % (uint32_t)e.nj);
    k = (int)(row / (uint32_t)e.nj);
}

__global__ __launch_bounds__(kBlock) void k_fs_layers_init(FsEpilogue e) {
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t ncell = (int64_t)e.ni * e.nj * e.nk;
    if (t >= ncell) return;
    int i, j, k;
    cell_of_index(e, (uint32_t)t, i, j, k);
    int8_t v;
    if (e.solid && e.solid[t]) {
        v = kSolid;
    } else if (e.fluid) {
        // the cell's bit in its P2G brick's mask (brick bins: the 8 P2G
        // bricks of each G2P brick consecutive)
        const int bx = i >> 2, by = j >> 2, bz = k >> 2;
        const uint32_t g = (uint32_t)(bx >> 1) +
                           (uint32_t)e.gbx * ((uint32_t)(by >> 1) + (uint32_t)e.gby * (uint32_t)(bz >> 1));
        const uint32_t bin = e.kb0 + ((g << 3) | (uint32_t)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)));
        const int bit = (i & 3) + 4 * (j & 3) + 16 * (k & 3);
        v = ((e.fluid[bin] >> bit) & 1ull) ? 0 : -1;
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
    int i, j, k;
    cell_of_index(e, (uint32_t)t, i, j, k);
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
    face_of(e, (uint32_t)f, c, i, j, k, d);
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

// the layer L at which face (c,i,j,k) is extrapolated: L borders it, L-1
// does not, no solid does, 1 <= L <= layers (macvelocityfield.cpp:694-706);
// 0 when none.  A face borders at most two cells, so at most one L fits.
__device__ __forceinline__ int face_layer(const FsEpilogue &e, int c, int i, int j, int k) {
    const int a = c == 0 ? i : (c == 1 ? j : k);
    const int na = c == 0 ? e.ni : (c == 1 ? e.nj : e.nk);
    const int li = i - (c == 0), lj = j - (c == 1), lk = k - (c == 2);
    const int hi = a < na ? layer_at(e, i, j, k) : -3;   // -3: no cell
    const int lo = a > 0 ? layer_at(e, li, lj, lk) : -3;
    if (hi == kSolid || lo == kSolid) return 0;
    int L = 0;
    if (hi >= 1 && lo != hi - 1) L = hi;
    if (lo >= 1 && hi != lo - 1 && (L == 0 || lo < L)) L = lo;
    return L <= e.layers ? L : 0;
}

// each face's extrapolation layer L (0: none) and, for a face that has one,
// which of its 6 same-component neighbours (in range) border layer L-1:
// one 16-bit word per face, L in the low byte, the mask in the high byte
__global__ __launch_bounds__(kBlock) void k_fs_classify(FsEpilogue e, uint32_t nface) {
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= nface) return;
    int c, i, j, k, d[3];
    face_of(e, f, c, i, j, k, d);
    const int L = face_layer(e, c, i, j, k);
    uint32_t mask = 0;
    if (L > 0) {
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                              {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (int q = 0; q < 6; q++) {
            const int a = nb[q][0], b = nb[q][1], cc = nb[q][2];
            if ((unsigned)a >= (unsigned)d[0] || (unsigned)b >= (unsigned)d[1] ||
                (unsigned)cc >= (unsigned)d[2])
                continue;
            if (face_borders(e, c, a, b, cc, L - 1)) mask |= 1u << q;
        }
    }
    e.fclass[f] = (uint16_t)((uint32_t)L | (mask << 8));
}

// one layer: every face of layer L takes the FP64 average of its masked
// neighbours (in the order -x, +x, -y, +y, -z, +z), 0 when that sum is 0,
// on both grids
__global__ __launch_bounds__(kBlock) void k_fs_extrapolate(FsEpilogue e, uint32_t nface, int L) {
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= nface) return;
    const uint32_t cl = e.fclass[f];
    if ((cl & 0xFFu) != (uint32_t)L) return;
    const uint32_t mask = cl >> 8;
    const uint32_t f0 = (uint32_t)(e.ni + 1) * (uint32_t)e.nj * (uint32_t)e.nk;
    const uint32_t f1 = (uint32_t)e.ni * (uint32_t)(e.nj + 1) * (uint32_t)e.nk;
    const int c = f < f0 ? 0 : (f < f0 + f1 ? 1 : 2);
    const uint32_t o = c == 0 ? f : (c == 1 ? f - f0 : f - f0 - f1);
    const uint32_t d0 = (uint32_t)(e.ni + (c == 0)), d1 = (uint32_t)(e.nj + (c == 1));
    const uint32_t step[3] = {1u, d0, d0 * d1};
    float *grid = c == 0 ? e.grid[0] : (c == 1 ? e.grid[1] : e.grid[2]);
    float *old = c == 0 ? e.old[0] : (c == 1 ? e.old[1] : e.old[2]);
    double sg = 0.0, so = 0.0, cnt = 0.0;
    for (int q = 0; q < 6; q++) {
        if (!(mask & (1u << q))) continue;
        const uint32_t g = (q & 1) ? o + step[q >> 1] : o - step[q >> 1];
        sg += (double)grid[g];
        so += (double)old[g];
        cnt += 1.0;
    }
    grid[o] = sg == 0.0 ? 0.0f : (float)(sg / cnt);
    old[o] = so == 0.0 ? 0.0f : (float)(so / cnt);
}

// ---- collision fix (fluidsimulation.cpp:3145-3207, collision.cpp:303-449)
// FP32 vec3 and FP64 scalar maths as the reference evaluates them; products
// and sums are rounded one by one (no fused multiply-add), as on the CPU.

__device__ __forceinline__ bool solid_at(const FsEpilogue &e, int i, int j, int k) {
    return (unsigned)i < (unsigned)e.ni && (unsigned)j < (unsigned)e.nj &&
           (unsigned)k < (unsigned)e.nk &&
           e.solid[(uint32_t)i + (uint32_t)e.ni * ((uint32_t)j + (uint32_t)e.nj * (uint32_t)k)] != 0;
}

// Grid3d::positionToGridIndex then the solid test; out of the grid is not solid
__device__ __forceinline__ bool point_in_solid(const FsEpilogue &e, float x, float y, float z) {
    const double inv = 1.0 / e.dx;
    const double fi = floor(__dmul_rn((double)x, inv));
    const double fj = floor(__dmul_rn((double)y, inv));
    const double fk = floor(__dmul_rn((double)z, inv));
    if (!(fi >= 0.0 && fj >= 0.0 && fk >= 0.0 && fi < e.ni && fj < e.nj && fk < e.nk)) return false;
    return solid_at(e, (int)fi, (int)fj, (int)fk);
}

// Collision::getLineSegmentVoxelIntersection: the first solid voxel on p0 -> p1
__device__ bool segment_voxel(const FsEpilogue &e, const float a[3], const float b[3], int vox[3]) {
    const float invdx = (float)(1.0 / e.dx);
    float p0[3], p1[3];
    for (int c = 0; c < 3; c++) {
        p0[c] = __fmul_rn(a[c], invdx);
        p1[c] = __fmul_rn(b[c], invdx);
    }
    int g0[3], g1[3], sgn[3];
    for (int c = 0; c < 3; c++) {
        g0[c] = (int)floor((double)p0[c]);
        g1[c] = (int)floor((double)p1[c]);
        sgn[c] = g1[c] > g0[c] ? 1 : g1[c] < g0[c] ? -1 : 0;
    }
    double v[3];
    for (int c = 0; c < 3; c++)
        v[c] = p1[c] == p0[c] ? 1.0 : (double)__fsub_rn(p1[c], p0[c]);
    const double vxvy = __dmul_rn(v[0], v[1]), vxvz = __dmul_rn(v[0], v[2]);
    const double vyvz = __dmul_rn(v[1], v[2]);
    const double gp0 = g0[0] + (g1[0] > g0[0] ? 1 : 0);
    const double gp1 = g0[1] + (g1[1] > g0[1] ? 1 : 0);
    const double gp2 = g0[2] + (g1[2] > g0[2] ? 1 : 0);
    double errx = __dmul_rn(__dsub_rn(gp0, (double)p0[0]), vyvz);
    double erry = __dmul_rn(__dsub_rn(gp1, (double)p0[1]), vxvz);
    double errz = __dmul_rn(__dsub_rn(gp2, (double)p0[2]), vxvy);
    const double derrx = __dmul_rn((double)sgn[0], vyvz);
    const double derry = __dmul_rn((double)sgn[1], vxvz);
    const double derrz = __dmul_rn((double)sgn[2], vxvy);
    int gx = g0[0], gy = g0[1], gz = g0[2];
    // the walk reaches the end voxel in |dgx| + |dgy| + |dgz| steps
    const long long steps = llabs((long long)g1[0] - g0[0]) + llabs((long long)g1[1] - g0[1]) +
                            llabs((long long)g1[2] - g0[2]) + 1;
    for (long long it = 0; it <= steps; it++) {
        if (solid_at(e, gx, gy, gz)) {
            vox[0] = gx;
            vox[1] = gy;
            vox[2] = gz;
            return true;
        }
        if (gx == g1[0] && gy == g1[1] && gz == g1[2]) break;
        const double xr = fabs(errx), yr = fabs(erry), zr = fabs(errz);
        if (sgn[0] != 0 && (sgn[1] == 0 || xr < yr) && (sgn[2] == 0 || xr < zr)) {
            gx += sgn[0];
            errx = __dadd_rn(errx, derrx);
        } else if (sgn[1] != 0 && (sgn[2] == 0 || yr < zr)) {
            gy += sgn[1];
            erry = __dadd_rn(erry, derry);
        } else if (sgn[2] != 0) {
            gz += sgn[2];
            errz = __dadd_rn(errz, derrz);
        }
    }
    return false;
}

// Collision::rayIntersectsAABB, with the reference's test of dir.x where
// dir.z is clamped
__device__ bool ray_aabb(const float p0[3], const float d[3], const float bmin[3],
                         const float bmax[3], float hit[3]) {
    const double eps = 1e-10;
    float dir[3] = {d[0], d[1], d[2]};
    if (fabs((double)dir[0]) < eps) dir[0] = dir[0] < 0 ? (float)-eps : (float)eps;
    if (fabs((double)dir[1]) < eps) dir[1] = dir[1] < 0 ? (float)-eps : (float)eps;
    if (fabs((double)dir[0]) < eps) dir[2] = dir[2] < 0 ? (float)-eps : (float)eps;
    float inv[3];
    for (int c = 0; c < 3; c++) inv[c] = (float)(1.0 / (double)dir[c]);
    double tmin = 0.0, tmax = 0.0;
    for (int c = 0; c < 3; c++) {
        const double t1 = (double)__fmul_rn(__fsub_rn(bmin[c], p0[c]), inv[c]);
        const double t2 = (double)__fmul_rn(__fsub_rn(bmax[c], p0[c]), inv[c]);
        if (c == 0) {
            tmin = fmin(t1, t2);
            tmax = fmax(t1, t2);
        } else {
            tmin = fmax(tmin, fmin(t1, t2));
            tmax = fmin(tmax, fmax(t1, t2));
        }
    }
    if (tmax > fmax(tmin, 0.0)) {
        const float tf = (float)tmin;
        for (int c = 0; c < 3; c++) hit[c] = __fadd_rn(p0[c], __fmul_rn(dir[c], tf));
        return true;
    }
    return false;
}

__global__ __launch_bounds__(kBlock) void k_fs_collide(FsEpilogue e) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= e.np) return;
    float b[3] = {e.x1[0][p], e.x1[1][p], e.x1[2][p]};
    if (!point_in_solid(e, b[0], b[1], b[2])) return;
    const int64_t p0 = e.x0_perm ? (int64_t)e.x0_perm[p] : p;
    const float a[3] = {e.x0[0][p0], e.x0[1][p0], e.x0[2][p0]};
    float r[3] = {a[0], a[1], a[2]};
    int vox[3];
    if (segment_voxel(e, a, b, vox)) {
        float d[3];
        for (int c = 0; c < 3; c++) d[c] = __fsub_rn(b[c], a[c]);
        const float lsq = __fadd_rn(__fadd_rn(__fmul_rn(d[0], d[0]), __fmul_rn(d[1], d[1])),
                                    __fmul_rn(d[2], d[2]));
        const float len = (float)sqrt((double)lsq);
        const float inv = (float)(1.0 / (double)len);
        float rn[3], bmin[3], bmax[3], hit[3];
        for (int c = 0; c < 3; c++) {
            rn[c] = __fmul_rn(d[c], inv);
            bmin[c] = (float)__dmul_rn((double)(float)vox[c], e.dx);
            bmax[c] = __fadd_rn(bmin[c], (float)e.dx);
        }
        if (ray_aabb(a, rn, bmin, bmax, hit)) {
            const float back = (float)(0.05 * e.dx);
            float q[3];
            for (int c = 0; c < 3; c++) q[c] = __fsub_rn(hit[c], __fmul_rn(rn[c], back));
            if (!point_in_solid(e, q[0], q[1], q[2]))
                for (int c = 0; c < 3; c++) r[c] = q[c];
        }
    }
    for (int c = 0; c < 3; c++) e.x1[c][p] = r[c];
}

}  // namespace

void fs_collide(const FsEpilogue &e, hipStream_t s) {
    if (e.np <= 0 || !e.solid) return;
    hipLaunchKernelGGL(k_fs_collide, dim3(blocks_for(e.np)), dim3(kBlock), 0, s, e);
    GFSX_LAUNCH_CHECK();
}

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
    // each face's layer once; a layer's launch changes only faces of that
    // layer and reads only faces bordering the layer before
    hipLaunchKernelGGL(k_fs_classify, dim3(blocks_for(nface)), dim3(kBlock), 0, s, e,
                       (uint32_t)nface);
    GFSX_LAUNCH_CHECK();
    for (int L = 1; L <= e.layers; L++) {
        hipLaunchKernelGGL(k_fs_extrapolate, dim3(blocks_for(nface)), dim3(kBlock), 0, s, e,
                           (uint32_t)nface, L);
        GFSX_LAUNCH_CHECK();
    }
}

}  // namespace gfsx
