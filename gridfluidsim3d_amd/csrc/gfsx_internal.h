// gfsx_internal.h -- device-side types and launch helpers shared by the
// gfx950 kernels of the FLIP transfer + advection library.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

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

// What the P2G kernels read of a scene: the particles in sort order, the
// scene's output grids, node grids and body force.
struct StepP2G {
    const float *x, *y, *z;
    const float *vel[3];
    float *grid[3], *grid_old[3], *wgt[3];
    NodeGrid ng[3];
    int64_t nface[3];
    float grav_delta[3];
    int grav_on[3];
    int ni, nj, nk;
};

// The particles of every scene in sort order (slot q of the sort): sorted
// copies x, v indexed by slot, or, when perm is set, each scene's own input
// arrays read at perm[q] - p0 (the particle the slot holds; x, v unused).
struct Sorted {
    const float *x[3], *v[3];
    const uint32_t *perm = nullptr;
};

// One scene of a (batched) fused step or of a cell sort, in a device table
// that the kernels index by particle, bin, binning block or brick: every
// scene owns a contiguous range of each (the offsets p0, b0, blk0, k0, g0
// grow with the scene index).  Bins of a scene: its bin grid bg (nbin
// in-grid bins), then its overflow bin (local index nbin).
struct SceneDesc {
    // particles: binning input (AoS when aos is set), step input and output
    const float *aos;
    const float *xin[3], *vin[3];
    float *xout[3], *vout[3];
    const uint32_t *id_in;
    uint32_t *id_out;
    int64_t n;        // particles
    int64_t p0;       // first global particle index (also its first sort slot)
    BinGrid bg;
    uint32_t b0;      // first global bin
    uint32_t blk0;    // first binning block
    // P2G bricks (4^3 cells) and G2P bricks (8^3 cells)
    int nbx, nby, nbz;
    uint32_t k0;
    int gbx, gby, gbz;
    uint32_t g0;
    // brick sort bins: the 8 P2G bricks of every G2P brick (8*gbx*gby*gbz,
    // some beyond the grid when nbx, nby or nbz is odd), then the overflow
    // bin (local index nbb); kb0 is the first global one
    uint32_t kb0, nbb;
    // occupied-brick masks of the scatter P2G: per (by, bz) row of bricks
    // rmw 64-bit words, bit bx set when brick (bx, by, bz) holds particles;
    // rm0 is the scene's first word
    uint32_t rm0;
    int rmw;
    int64_t f0;       // first global MAC face (faces of u, then v, then w, per scene)
    int64_t r0;       // first combine wave (over the x-rows of u, then v, then w)
    int64_t cw[2];    // combine waves of the scene before its v rows and before its w rows
    // grid and step constants
    int ni, nj, nk;
    double dx;
    float limf[3];     // MacGrid::limf of the scene's grid (set_limits)
    double lim[3];
    float *grid[3], *grid_old[3], *wgt[3];
    float grav_delta[3];
    int grav_on[3];
    float h_half, h_full, h_sixth;   // (float)(0.5*dt), (float)dt, (float)(dt/6.0f)
    float r_pic, r_flip;             // (float)ratio, (float)(1-ratio)
    StepP2G p2g;                     // the P2G view (p2g_view), filled before upload
};

// P2G view of a scene: faces of component c sit at off + n*dx with the
// offset half a cell on the two other axes (fluidsimulation.cpp:2526-2570);
// a face's candidate bins along axis a are [n + dlo, n + dhi] (pad 1, R = dx).
__host__ __device__ inline StepP2G p2g_view(const SceneDesc &sd, const Sorted &so) {
    StepP2G a;
    a.x = so.x[0];
    a.y = so.x[1];
    a.z = so.x[2];
    const float half = (float)(0.5 * sd.dx);
    for (int c = 0; c < 3; c++) {
        a.vel[c] = so.v[c];
        a.grid[c] = sd.grid[c];
        a.grid_old[c] = sd.grid_old[c];
        a.wgt[c] = sd.wgt[c];
        a.grav_delta[c] = sd.grav_delta[c];
        a.grav_on[c] = sd.grav_on[c];
        NodeGrid &ng = a.ng[c];
        ng.n[0] = sd.ni + (c == 0);
        ng.n[1] = sd.nj + (c == 1);
        ng.n[2] = sd.nk + (c == 2);
        ng.dx_f = (float)sd.dx;
        for (int ax = 0; ax < 3; ax++) {
            ng.off[ax] = (ax == c) ? 0.0f : half;
            // node n sits at bin coordinate n + off/dx + pad; its candidate
            // bins cover (n + base - R/dx, n + base + R/dx) with R = dx
            const double base = (double)ng.off[ax] / sd.dx + 1.0;
            ng.dlo[ax] = (int)floor(base - 1.0);
            ng.dhi[ax] = (int)ceil(base + 1.0) - 1;
        }
        a.nface[c] = (int64_t)ng.n[0] * ng.n[1] * ng.n[2];
    }
    a.ni = sd.ni;
    a.nj = sd.nj;
    a.nk = sd.nk;
    return a;
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

// Element i of an array in device memory, read as a global-memory load.
// Pointers taken from descriptor structs are generic (flat) to the
// compiler: a flat load counts against the LDS wait counter too, and may
// alias LDS, so the compiler keeps it in order with LDS accesses and waits
// for it before each one.  Naming the global address space lets loads be
// issued in batches ahead of the LDS work that uses them.
template <class T>
__device__ __forceinline__ T gld(const T *p, size_t i) {
    typedef const __attribute__((address_space(1))) T GT;
    return ((GT *)p)[i];
}

// Element i of an array in device memory written as a global-memory store
// (a flat store counts against the LDS wait counter as well).
template <class T>
__device__ __forceinline__ void gst(T *p, size_t i, T v) {
    typedef __attribute__((address_space(1))) T GT;
    ((GT *)p)[i] = v;
}

// The element at byte offset off of an array in device memory: a 32-bit
// offset lets the access take the array's address from scalar registers
// (no 64-bit address arithmetic per lane).
template <class T>
__device__ __forceinline__ T gld_at(const T *p, uint32_t off) {
    typedef const __attribute__((address_space(1))) char GC;
    typedef const __attribute__((address_space(1))) T GT;
    return *(GT *)((GC *)p + off);
}

template <class T>
__device__ __forceinline__ void gst_at(T *p, uint32_t off, T v) {
    typedef __attribute__((address_space(1))) char GC;
    typedef __attribute__((address_space(1))) T GT;
    *(GT *)((GC *)p + off) = v;
}

// Wyvill kernel for t = r^2/R^2 (scalarfield.cpp:569-571 rewritten as the
// factorisation K = (1-t)^2 (1 - 4t/9) = a^2 (5/9 + 4a/9), a = 1-t):
// exact zero for t >= 1 without a compare, and no cancellation near t = 1.
__device__ __forceinline__ float wyvill(float t) {
    float a = fmaxf(1.0f - t, 0.0f);
    return a * a * fmaf(4.0f / 9.0f, a, 5.0f / 9.0f);
}

// Adjoint of one plain step (trilinear stencil, particles in input order):
// the step's inputs, the adjoints of its outputs, the grids of the forward
// step recomputed from those inputs, and face workspace for the grid
// adjoints.
struct StepAdjoint {
    int64_t n;
    const float *x[3], *v[3];      // step inputs
    const float *ax[3], *av[3];    // dL/dx', dL/dv'
    float *gx[3], *gv[3];          // dL/dx, dL/dv (written)
    const float *g[3];             // grid after the body force
    const float *old[3];           // normalised P2G result
    const float *w[3];             // P2G weight sums
    float *lg[3], *ln[3];          // dL/dg, dL/dn (workspace, zeroed here), then
                                   // the P2G adjoint's per-face factors A, B
    MacGrid mac;                   // dims, dx, limits (f unused)
    float r_pic, r_flip;
    float h_half, h_full, h_sixth;
    float half_dx;                 // (float)(0.5*dx): the face offsets
    float inv_r2;                  // (float)(1/dx^2): P2G radius R = dx
};

// The brick sort of the forward step that the adjoint recomputes (one
// scene): bin lb < nbb is the P2G brick of 4^3 cells numbered as the sort
// numbers them (the 8 bins of a G2P brick consecutive, gbx x gby G2P bricks
// per z layer), bin nbb the particles beyond the padded grid; its particles
// are slots [pstart[lb], pstart[lb + 1]) of the sorted positions xs, and
// perm maps a slot to its input index.  fx: fixed-point accumulators for
// the grid adjoints (int64, the face arrays of dL/dg u,v,w then dL/dn u,v,w
// back to back); bound: three words for the fixed-point scale's bound
// (max |dL/dx'|, max |dL/dv'|, max |g|, float bits).
struct AdjointSort {
    const uint32_t *pstart, *perm;
    const float *xs[3];
    uint32_t nbb;
    int gbx, gby;
    unsigned long long *fx;
    uint32_t *bound;
};

// FluidSimulation's face fill and extrapolation around the transfer (one
// scene): the sort's bin starts give the fluid cells, n and w are the P2G
// gather's normalised values and weights, grid and old receive the field
// (with the body force) and the snapshot, both extrapolated.
struct FsEpilogue {
    int ni, nj, nk;
    const uint32_t *start;      // bin starts of the cell sort (bins (n+2)^3, pad 1), or
    int b[3];                   // bins per axis
    const unsigned long long *fluid;  // the brick sort's per-brick fluid-cell masks
    uint32_t kb0;               // the scene's first brick bin (with fluid)
    int gbx, gby;               // G2P bricks per axis (with fluid)
    const uint8_t *solid;       // optional, ni*nj*nk, nonzero = solid
    int8_t *layer;              // workspace, ni*nj*nk
    const float *n[3], *w[3];
    float *grid[3], *old[3];
    float grav_delta[3];
    int grav_on[3];
    int layers;
    double dx;
    // collision fix after RK4 (with solid cells): the step's outputs x1 and
    // the same particles' inputs x0 (by output index, or at x0_perm[index]
    // when x0_perm is set), np particles
    int64_t np;
    const float *x0[3];
    const uint32_t *x0_perm;
    float *x1[3];
    // workspace: each face's extrapolation layer (low byte, 0: none) and
    // neighbour mask (high byte), nface words
    uint16_t *fclass;
};

// Particle removal (gfsx_remove.hip): positions/velocities in, kept ones
// out (input order), the cell sort over the grid's cells (bg, no pad).
struct RemoveArgs {
    int64_t n;
    const float *x[3], *v[3];
    float *xo[3], *vo[3];
    const uint32_t *id_in;
    uint32_t *id_out;
    BinGrid bg;
    const uint32_t *start, *perm;
    const uint8_t *solid;
    int max_per_cell;
    uint64_t seed;
    uint32_t *keep, *dst;
    uint32_t *full, *nfull;   // over-full cells: list (capacity nbin) and count
};

// ---- launches (implemented in gfsx_kernels.hip) ------------------------

void fill_u32(uint32_t *p, uint32_t value, int64_t count, hipStream_t s);
void fill_f32(float *p, float value, int64_t count, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t value, int64_t count, hipStream_t s);
// gfsx_epilogue.hip
void fs_epilogue(const FsEpilogue &e, hipStream_t s);
void fs_collide(const FsEpilogue &e, hipStream_t s);
// gfsx_remove.hip
void remove_mark(const RemoveArgs &r, hipStream_t s);
void remove_compact(const RemoveArgs &r, hipStream_t s);
// gfsx_adjoint.hip
// so == nullptr selects one thread per particle in input order with every
// grid adjoint added straight to global memory; otherwise bricks of the
// sorted particles accumulate in LDS first.
void flip_step_adjoint(const StepAdjoint &a, const AdjointSort *so, hipStream_t s);
void p2g_adjoint(const float *pos, const float *val, int64_t n, double radius,
                 const float off[3], double dx, int ni, int nj, int nk, const float *gS,
                 const float *gW, float *gpos, float *gval, hipStream_t s);
int64_t scan_scratch_words(int64_t m);
void exclusive_scan(const uint32_t *in, uint32_t *out, int64_t m,
                    uint32_t *scratch, hipStream_t s);

struct SortBuffers {
    const SceneDesc *scenes;   // device table
    int nscene;
    uint32_t nbin_total;  // sum over scenes of (in-grid bins + 1 overflow bin)
    uint32_t nblk_total;  // binning blocks, sum over scenes
    uint32_t *count;      // nbin_total
    uint32_t *start;      // nbin_total+1
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
    uint32_t *block_out;  // nblk_total
    uint32_t *block_out_scan;  // nblk_total+1
    uint32_t *scan_scratch;
    uint8_t *oct;         // n: half of its cell each particle lies in
    uint8_t *poct;        // n: the same, in perm order
};
int64_t sort_blocks(int64_t n);

// Stable counting sort of particle indices by bin, for every scene of
// sb.scenes at once: perm lists scene 0's bins in order (each bin's particles
// in index order, the overflow bin last), then scene 1's, ...; start[b] is
// the first perm slot of global bin b, start[nbin_total] == n.  Indices are
// global (scene p0 + local index).
// With octant_order, particles inside an in-grid bin are ordered by
// (octant, index) instead, the octant being the 3-bit index of the half of
// the cell the particle lies in along x, y, z (so each half-cell is a
// contiguous run), hrank[q] receives the rank of perm[q] inside its
// half-cell run and soct[q] its octant.  With mark_bricks, brick_flag
// receives the P2G bricks with a particle within reach.
void cell_sort(const SortBuffers &sb, int64_t n, hipStream_t s, bool octant_order,
               bool mark_bricks = false);

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

// Face rows along x that one wave of the combine takes in turn: enough for
// about 256 faces, so that short rows share a wave's set-up.
__host__ __device__ inline int combine_rows(int nx) { return nx < 256 ? 256 / nx : 1; }

// Sizes of a scene table laid out by layout_scenes.
struct SceneTotals {
    int64_t n = 0;       // particles
    int64_t nbin = 0;    // bins, overflow bins included
    int64_t nblk = 0;    // binning blocks
    int64_t np2g = 0;    // P2G bricks
    int64_t ng2p = 0;    // G2P bricks
    int64_t nface = 0;   // MAC faces
    int64_t nbb = 0;     // brick sort bins, overflow bins included
    int64_t nrow = 0;    // combine waves (face rows along x, short rows packed), all components
    int64_t nrm = 0;     // occupied-brick mask words
};

// Fill the derived fields of every scene (bin grid, offsets p0/b0/blk0/k0/g0,
// brick counts, grid limits) from n, ni/nj/nk and dx, in table order.
void layout_scenes(SceneDesc *scenes, int nscene, SceneTotals &tot);

// Nodes of one P2G brick's scatter tile, all three components: per
// component 5 nodes along its own axis and 6 along the two others.
constexpr int kP2gTileNodes = 3 * 5 * 6 * 6;

// Workspace of the brick sort (gfsx_brick.hip).  A particle's bin is its
// P2G brick; a wave of 64 consecutive particles of the input forms one run
// per distinct bin among its lanes, and a bin's particles are laid out run
// by run in wave order, each run in lane order: the order of a stable
// counting sort by bin.
struct BrickSort {
    unsigned long long *count;   // nbb_total: (runs << 32) | particles per bin
    unsigned long long *start;   // nbb_total + 1: exclusive scan of count
    uint32_t *pstart;            // nbb_total + 1: first sort slot of each bin
    uint32_t *rstart;            // nbb_total + 1: first entry of each bin's run list
    uint2 *runrec;               // TILE per binning work-group: chunk (bin, provisional rank)
    uint32_t *rlen;              // TILE per binning work-group: chunk length
    uint8_t *nruns;              // per wave: its number of runs
    uint32_t *nchunk;            // per binning work-group: chunks, or every run a chunk
    uint32_t *jr;                // per binning position: chunk (bits 10..) and offset in
                                 // it (bits 0..9)
    uint32_t *list;              // chunks (blk*TILE + c), grouped by bin
    uint32_t *list2;             // the same ordered, for bins of many runs
    uint32_t *runbase;           // TILE per binning work-group: each chunk's first slot
    uint32_t *mid, *huge;        // bins of 17..256 runs, of more
    uint32_t *lcount;            // 2: entries of mid and huge
    unsigned long long *scan_scratch;
    uint32_t *perm;              // nullable: sort slot -> global particle index
    unsigned long long *fluid;   // per bin: the cells of the brick that hold a particle
    // G2P work left to the global-memory kernel: single sort slots (at most
    // n) and their count
    uint32_t *slow_slots;
    uint32_t *slow_count;
    int64_t nwave;
    uint32_t nbb_total;
};
constexpr uint32_t kBrickBinBits = 26;   // global bins below 2^26
int64_t scan64_scratch_words(int64_t m);

// A fused step over one or more scenes.
struct StepBatch {
    std::vector<SceneDesc> scenes;   // host copy, laid out
    const SceneDesc *scenes_dev;     // the same table on the device
    const int64_t *r0s = nullptr;    // the scenes' first combine waves, dense (several scenes)
    SceneTotals tot;
    float *xs[3], *vs[3];            // workspace: all particles in sort order (may be
                                     // null when gather is set)
    bool gather = false;             // p2g_variant 4: the sort writes only the permutation;
                                     // P2G reads each slot's particle from the inputs
                                     // through it, and writes xs/vs when they are set, else
                                     // G2P reads through it as well
    uint32_t *dense, *dense_count;   // workspace: P2G bricks over capacity (list; count
                                     // and next unprocessed entry)
    float2 *tiles = nullptr;         // workspace of the scatter P2G: kP2gTileNodes (S, W) per brick
    unsigned long long *tile_rows = nullptr;   // occupied-brick masks (rm0, rmw)
    int num_cu = 1;                  // compute units of the device
    int stencil;
    int keep_order;
    int p2g_variant;   // 0, 2, 3: brick-staged gather, 1: direct global gather (one scene),
                       // 4: scatter into per-brick tiles + combine
    int g2p_variant;   // 0: brick-staged LDS boxes, 1: direct global sampling (one scene)
    int g2p_brick;     // cells per edge of a G2P brick with g2p_variant 0 and p2g_variant 4: 8 or 4
    const FsEpilogue *epi = nullptr;   // FluidSimulation epilogue after P2G (one scene)
    bool grids_only = false;           // stop after P2G (the adjoint's recompute)
    bool p2 = false;                   // every scene's 1/dx is a power of two: the brick
                                       // G2P takes cell coordinates in FP32 (same bits)
    const BrickSort *bs = nullptr;     // p2g_variant 4: the brick sort's workspace
};

// The fused step with p2g_variant 4: brick sort, scatter P2G into per-brick
// tiles, per-brick combine, G2P over brick ranges (gfsx_brick.hip).
void flip_step_brick(const StepBatch &s, const BrickSort &bs, hipStream_t st, hipEvent_t *ev);

// Write n scene table entries to device memory at dst (and their combine
// wave offsets to r0[0..n) when r0 is set) with one small kernel each (the
// entries travel as kernel arguments).
void put_scenes(SceneDesc *dst, const SceneDesc *src, int n, hipStream_t st,
                int64_t *r0 = nullptr);

// The direct G2P kernel (g2p_variant 1, one scene) over the particles in
// s.xs/s.vs; perm maps sort slots to particle indices (needed for
// GFSX_ORDER_KEEP and particle ids).
void flip_g2p_direct(const StepBatch &s, const uint32_t *perm, hipStream_t st);

// Positions and velocities of every scene copied into the order of perm
// (the cell sort), for the adjoint's bricks.
void sorted_copy(const SceneDesc *scenes_dev, int nsc, const uint32_t *perm, int64_t n,
                 float *const xs[3], float *const vs[3], hipStream_t s);

// The fused plain step of every scene of s, in one launch per kernel.
// Needs SortBuffers laid out for s.tot (and sb.scenes == s.scenes_dev).
// When ev is not null, ev[0..3] are recorded on the stream before the sort,
// after the sort, after P2G and after G2P.
void flip_step(const StepBatch &s, const SortBuffers &sb, hipStream_t st,
               hipEvent_t *ev = nullptr);

}  // namespace gfsx
