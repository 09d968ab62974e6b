// gfsx_internal.h -- device-side types and launch helpers shared by the
// gfx950 kernels of the FLIP transfer + advection library.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace gfsx {

constexpr int kBlock = 256;          // threads per work-group (4 wave64s)
constexpr int kScanItems = 16;       // items per thread in the scan tiles
constexpr int kScanTile = kBlock * kScanItems;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GFSX_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::gfsx::Error(std::string(#call) + ": " +                   \
                                hipGetErrorString(e_));                       \
    } while (0)

#define GFSX_LAUNCH_CHECK()                                                   \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess)                                                 \
            throw ::gfsx::Error(std::string("kernel launch: ") +              \
                                hipGetErrorString(e_));                       \
    } while (0)

// Binning frame.  A particle at p goes to the cell floor((p - off) / dx) of
// the shifted frame (FP64 floor, as Grid3d::positionToGridIndex does,
// grid3d.h:58-63), offset by `pad` so that cells just outside the node
// range still get a bin.  Anything outside [0, b) on an axis goes to the
// single overflow bin `nbin`.
struct BinGrid {
    int b[3];       // bins per axis
    int pad[3];     // bin index of shifted-frame cell 0
    float off[3];   // FP32 shift applied to positions before binning
    double inv_dx;  // 1.0 / dx, rounded on the host like the reference
    uint32_t nbin;  // b[0]*b[1]*b[2]; also the index of the overflow bin
};

// Node grid of one splatted component.  Node n sits at off + n*dx; its
// candidate bins along axis a are [n + dlo[a], n + dhi[a]] (clipped).
struct NodeGrid {
    int n[3];
    float off[3];
    int dlo[3], dhi[3];
    float dx_f;      // unused by the maths; kept for diagnostics
};

// A MAC grid view for G2P (macvelocityfield.cpp:37-45 layout).
struct MacGrid {
    const float *f[3];
    int ni, nj, nk;
    double dx, inv_dx;
    double lim[3];   // dx*ni, dx*nj, dx*nk (Grid3d::isPositionInGrid)
    float limf[3];   // for float x, x < limf[a] exactly when (double)x < lim[a]
};

// Fill lim/limf of g from dx and the cell counts.  limf is the smallest
// float not below lim: every float under it is under lim, every float at
// or above it is at or above lim.
inline void set_limits(MacGrid &g) {
    const int n[3] = {g.ni, g.nj, g.nk};
    for (int a = 0; a < 3; a++) {
        g.lim[a] = g.dx * n[a];
        float f = (float)g.lim[a];
        if ((double)f < g.lim[a]) f = std::nextafter(f, INFINITY);
        g.limf[a] = f;
    }
}

__device__ __forceinline__ uint32_t bin_of(float x, float y, float z,
                                           const BinGrid &g) {
    float px = x - g.off[0], py = y - g.off[1], pz = z - g.off[2];
    double fi = floor((double)px * g.inv_dx) + g.pad[0];
    double fj = floor((double)py * g.inv_dx) + g.pad[1];
    double fk = floor((double)pz * g.inv_dx) + g.pad[2];
    // NaN compares false, so it lands in the overflow bin
    if (!(fi >= 0.0 && fj >= 0.0 && fk >= 0.0 && fi < (double)g.b[0] &&
          fj < (double)g.b[1] && fk < (double)g.b[2]))
        return g.nbin;
    return (uint32_t)fi + (uint32_t)g.b[0] * ((uint32_t)fj + (uint32_t)g.b[1] * (uint32_t)fk);
}

// Wyvill kernel for t = r^2/R^2 (scalarfield.cpp:569-571 rewritten as the
// factorisation K = (1-t)^2 (1 - 4t/9) = a^2 (5/9 + 4a/9), a = 1-t):
// exact zero for t >= 1 without a compare, and no cancellation near t = 1.
__device__ __forceinline__ float wyvill(float t) {
    float a = fmaxf(1.0f - t, 0.0f);
    return a * a * fmaf(4.0f / 9.0f, a, 5.0f / 9.0f);
}

// ---- launches (implemented in gfsx_kernels.hip) ------------------------

void fill_u32(uint32_t *p, uint32_t value, int64_t count, hipStream_t s);
void fill_f32(float *p, float value, int64_t count, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t value, int64_t count, hipStream_t s);
int64_t scan_scratch_words(int64_t m);
void exclusive_scan(const uint32_t *in, uint32_t *out, int64_t m,
                    uint32_t *scratch, hipStream_t s);

struct SortBuffers {
    uint32_t *count;      // nbin+1
    uint32_t *start;      // nbin+2
    uint32_t *key;        // n
    uint32_t *slot;       // n
    uint32_t *perm;       // n: the result
    uint32_t *perm_tmp;   // n: bins filled, not yet ordered inside
    uint32_t *hrank;      // n (octant order only): rank inside the half-cell
    uint8_t *soct;        // n (octant order only): octant, in perm order
    uint8_t *brick_flag;  // (octant order only, nullable): P2G bricks with particles in reach
    uint32_t *big_list;   // big_cap entries: bins too large for one thread
    uint32_t big_cap;
    uint32_t *big_count;  // 2
    uint32_t *block_out;  // blocks(n)
    uint32_t *block_out_scan;  // blocks(n)+1
    uint32_t *scan_scratch;
    uint8_t *oct;         // n: half of its cell each particle lies in
    uint8_t *poct;        // n: the same, in perm order
};
int64_t sort_blocks(int64_t n);

// Stable counting sort of particle indices by bin: perm lists bin 0's
// particles in index order, then bin 1's, ..., then the overflow bin's.
// start[b] is the first perm slot of bin b; start[nbin+1] == n.
// With octant_order, particles inside a bin are ordered by (octant, index)
// instead, the octant being the 3-bit index of the half of the cell the
// particle lies in along x, y, z (so each half-cell is a contiguous run),
// hrank[q] receives the rank of perm[q] inside its half-cell run and
// soct[q] its octant.
// aos != nullptr selects AoS input (x,y,z interleaved); otherwise SoA.
void cell_sort(const float *aos, const float *x, const float *y,
               const float *z, int64_t n, const BinGrid &g,
               const SortBuffers &sb, hipStream_t s, bool octant_order = false);

// Generic P2G gather for one component (ScalarField semantics).
void p2g_gather(const float *aos, const float *val, int64_t n,
                const BinGrid &bins, const NodeGrid &nodes, double radius,
                double dx, const SortBuffers &sb, float *S, float *W,
                hipStream_t s);

void apply_weight(float *S, const float *W, int64_t count, hipStream_t s);

// G2P / RK on AoS particle arrays in input order.
void g2p(const float *pos, int64_t n, const MacGrid &g, int stencil,
         float *out, hipStream_t s);
void advect(const float *pos, int64_t n, const MacGrid &g, int order,
            double dt, int stencil, float *out, hipStream_t s);

struct StepArgs {
    int64_t n;
    int ni, nj, nk;
    double dx;
    const float *xin[3], *vin[3];
    float *xout[3], *vout[3];
    float *xs[3], *vs[3];            // workspace: the input in sorted order
    uint32_t *dense, *dense_count;   // workspace: P2G bricks over capacity
    const uint32_t *id_in;
    uint32_t *id_out;
    float *grid[3], *grid_old[3], *wgt[3];
    float grav_delta[3];
    int grav_on[3];
    float h_half, h_full, h_sixth;   // (float)(0.5*dt), (float)dt, (float)(dt/6.0f)
    float r_pic, r_flip;             // (float)ratio, (float)(1-ratio)
    int stencil;
    int keep_order;
    int p2g_variant;   // 0: brick-staged gather, 1: direct global gather
    int g2p_variant;   // 0: brick-staged LDS boxes, 1: direct global sampling
    int num_cu;        // compute units of the device (persistent launches)
};

// The fused plain step.  Needs SortBuffers sized for n particles and
// (ni+2)*(nj+2)*(nk+2) bins.  When ev is not null, ev[0..3] are recorded on
// the stream before the sort, after the sort, after P2G and after G2P.
void flip_step(const StepArgs &a, const SortBuffers &sb, hipStream_t s,
               hipEvent_t *ev = nullptr);

}  // namespace gfsx
