// gfsx_kernels.hip -- gfx950 kernels of the FLIP transfer + advection path.
//
//  * cell_sort   : stable counting sort of particle indices by cell (integer
//                  work, bit-identical to a sequential stable counting sort).
//  * p2g gather  : every node sums the kernel-weighted values of the
//                  particles in its candidate cells, in a fixed order (no
//                  float atomics), then normalises, snapshots and adds the
//                  body force on faces bordering fluid.
//  * g2p / RK    : gather-only trilinear or tricubic sampling of the MAC
//                  grid, fused with the PIC/FLIP blend and the RK4 stages.
//
// Reference behaviour followed: ScalarField::addPointValue / applyWeightField
// (scalarfield.cpp:90-106,167-201), MACVelocityField::_interpolate{,Linear}
// U/V/W (macvelocityfield.cpp:350-575), ParticleAdvector::_RK1.._RK4
// (particleadvector.cpp:1045-1078), FluidSimulation stages
// (fluidsimulation.cpp:2010-2017,2526-2595,2766-2805,3104-3129,3181-3209).

#include "gfsx_internal.h"
#include "gfsx_sample.h"
#include "gfsx_device.h"

#include <algorithm>
#include <type_traits>

namespace gfsx {

// ===========================================================================
// Utilities
// ===========================================================================


__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f32(float *p, float v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void fill_u32(uint32_t *p, uint32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    unsigned g = std::min<unsigned>(grid_for(count, kBlock), 8192u);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, s, p, value, count);
    GFSX_LAUNCH_CHECK();
}

__global__ void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void fill_u8(uint8_t *p, uint8_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    unsigned g = std::min<unsigned>(grid_for(count, kBlock), 8192u);
    hipLaunchKernelGGL(k_fill_u8, dim3(g), dim3(kBlock), 0, s, p, value, count);
    GFSX_LAUNCH_CHECK();
}

void fill_f32(float *p, float value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    unsigned g = std::min<unsigned>(grid_for(count, kBlock), 8192u);
    hipLaunchKernelGGL(k_fill_f32, dim3(g), dim3(kBlock), 0, s, p, value, count);
    GFSX_LAUNCH_CHECK();
}

// ===========================================================================
// Device-wide exclusive scan of uint32 (tiles of 4096, three phases)
// ===========================================================================



__global__ __launch_bounds__(kBlock) void k_tile_reduce(const uint32_t *in,
                                                       int64_t m,
                                                       uint32_t *sums) {
    __shared__ uint32_t wsum[kBlock / 64];
    int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        int64_t idx = base + it * kBlock + threadIdx.x;
        if (idx < m) s += in[idx];
    }
    uint32_t total;
    block_exclusive_sum(s, wsum, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__device__ __forceinline__ int padded(int e) { return e + (e >> 5); }

// out[i] = sum(in[0..i)), out[m] = sum(in[0..m)).  tile_off (nullable) holds
// the exclusive prefix of the tile sums.
__global__ __launch_bounds__(kBlock) void k_tile_scan(const uint32_t *in,
                                                     int64_t m,
                                                     const uint32_t *tile_off,
                                                     uint32_t *out) {
    __shared__ uint32_t tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t wsum[kBlock / 64];
    int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        int e = it * kBlock + threadIdx.x;
        int64_t idx = base + e;
        tile[padded(e)] = idx < m ? in[idx] : 0u;
    }
    __syncthreads();
    uint32_t sum = 0;
    const int first = threadIdx.x * kScanItems;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) sum += tile[padded(first + i)];
    uint32_t total;
    uint32_t toff = tile_off ? tile_off[blockIdx.x] : 0u;
    uint32_t run = block_exclusive_sum(sum, wsum, total) + toff;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint32_t v = tile[padded(first + i)];
        tile[padded(first + i)] = run;
        run += v;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        int e = it * kBlock + threadIdx.x;
        int64_t idx = base + e;
        if (idx < m) out[idx] = tile[padded(e)];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[m] = toff + total;
}

int64_t scan_scratch_words(int64_t m) {
    int64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb <= 1) return 0;
    return 2 * nb + 1 + scan_scratch_words(nb);
}

void exclusive_scan(const uint32_t *in, uint32_t *out, int64_t m,
                    uint32_t *scratch, hipStream_t s) {
    if (m <= 0) {
        fill_u32(out, 0u, 1, s);
        return;
    }
    int64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kBlock), 0, s, in, m,
                           (const uint32_t *)nullptr, out);
        GFSX_LAUNCH_CHECK();
        return;
    }
    uint32_t *sums = scratch;
    uint32_t *sums_scan = scratch + nb;
    hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)nb), dim3(kBlock), 0, s,
                       in, m, sums);
    GFSX_LAUNCH_CHECK();
    exclusive_scan(sums, sums_scan, nb, scratch + 2 * nb + 1, s);
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nb), dim3(kBlock), 0, s,
                       in, m, (const uint32_t *)sums_scan, out);
    GFSX_LAUNCH_CHECK();
}

// ===========================================================================
// Stable counting sort by bin
// ===========================================================================

// The binning and scatter kernels take kBinItems particles per thread
// (item m of a block is the consecutive range [m*kBlock, (m+1)*kBlock) of
// the block's kBinItems*kBlock particles), so that every wave has that many
// independent loads and atomics in flight.
constexpr int kBinItems = 4;
constexpr int64_t kBinTile = (int64_t)kBinItems * kBlock;

int64_t sort_blocks(int64_t n) { return (n + kBinTile - 1) / kBinTile; }


// Bin every particle.  A block takes kBinTile consecutive particles of one
// scene.  In-grid particles take a slot in their bin with an integer
// atomic; lanes of a wave that hold a run of equal bins (the input is
// usually close to cell order) take one atomic per run and number themselves
// within it in lane order.  The order inside a bin is then fixed by
// k_order_bins / k_order_big.  Overflow particles get their rank among the
// block's overflow particles (in index order) from wave ballots, so a
// scene's overflow bin is filled in index order directly.  With an octant
// array, each particle also records which half of its cell it occupies
// along each axis.
__global__ __launch_bounds__(kBlock) void k_bin(const SceneDesc *__restrict__ scenes, int nsc,
                                               uint32_t *count, uint32_t *key,
                                               uint32_t *slot, uint32_t *block_out,
                                               uint8_t *oct) {
    __shared__ uint32_t wcnt[kBinItems][kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t blk = blockIdx.x;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, blk, [](const SceneDesc &d) { return d.blk0; })];
    const BinGrid g = sd.bg;
    const int64_t n = sd.n;
    const int64_t base = (int64_t)(blk - sd.blk0) * kBinTile + threadIdx.x;
    float px[kBinItems], py[kBinItems], pz[kBinItems];
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        const int64_t ii = i < n ? i : 0;
        if (sd.aos) {
            px[m] = sd.aos[3 * ii];
            py[m] = sd.aos[3 * ii + 1];
            pz[m] = sd.aos[3 * ii + 2];
        } else {
            px[m] = sd.xin[0][ii];
            py[m] = sd.xin[1][ii];
            pz[m] = sd.xin[2][ii];
        }
    }
    uint32_t b[kBinItems], sl[kBinItems];
    uint8_t o[kBinItems];
    bool over[kBinItems];
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        const bool valid = base + (int64_t)m * kBlock < n;
        b[m] = valid ? bin_of(px[m], py[m], pz[m], g) : g.nbin;
        o[m] = 0;
        if (valid && oct) {
            const double gx = (double)(px[m] - g.off[0]) * g.inv_dx;
            const double gy = (double)(py[m] - g.off[1]) * g.inv_dx;
            const double gz = (double)(pz[m] - g.off[2]) * g.inv_dx;
            o[m] = (uint8_t)(((gx - floor(gx)) >= 0.5) | (((gy - floor(gy)) >= 0.5) << 1) |
                             (((gz - floor(gz)) >= 0.5) << 2));
        }
        over[m] = valid && b[m] == g.nbin;
        // runs of equal in-grid bins among consecutive lanes
        const bool in = valid && !over[m];
        const uint32_t kk = in ? b[m] : 0xFFFFFFFFu;
        const uint32_t kprev = __shfl_up(kk, 1, 64);
        const uint32_t knext = __shfl_down(kk, 1, 64);
        const bool rstart = in && (lane == 0 || kprev != kk);
        const bool rend = in && (lane == 63 || knext != kk);
        const uint64_t S = __ballot(rstart), E = __ballot(rend);
        const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int s0 = in ? 63 - __clzll(S & le) : 0;
        uint32_t rbase = 0;
        if (rstart) {
            const int e0 = __ffsll((unsigned long long)(E & ~((1ull << lane) - 1ull))) - 1;
            rbase = atomicAdd(&count[sd.b0 + b[m]], (uint32_t)(e0 - lane + 1));
        }
        rbase = __shfl(rbase, s0, 64);
        sl[m] = rbase + (uint32_t)(lane - s0);
        const uint64_t mask = __ballot(over[m]);
        if (over[m]) sl[m] = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[m][wid] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    // overflow ranks: all earlier items, then the earlier waves of this item
    uint32_t tot = 0;
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        uint32_t woff = tot;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) {
            woff += (w < wid) ? wcnt[m][w] : 0u;
            tot += wcnt[m][w];
        }
        if (over[m]) sl[m] += woff;
    }
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        if (i < n) {
            key[sd.p0 + i] = sd.b0 + b[m];
            slot[sd.p0 + i] = sl[m];
            if (oct) oct[sd.p0 + i] = o[m];
        }
    }
    if (threadIdx.x == 0) block_out[blk] = tot;
}

// Each scene's overflow count: the sum of its blocks' overflow counts.
__global__ void k_overflow_counts(const SceneDesc *__restrict__ scenes, int nsc,
                                  const uint32_t *__restrict__ block_out_scan,
                                  uint32_t nblk_total, uint32_t *count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsc) return;
    const uint32_t e = s + 1 < nsc ? scenes[s + 1].blk0 : nblk_total;
    count[scenes[s].b0 + scenes[s].bg.nbin] = block_out_scan[e] - block_out_scan[scenes[s].blk0];
}

__global__ void k_copy_one(uint32_t *dst, const uint32_t *src) { *dst = *src; }

__global__ __launch_bounds__(kBlock) void k_scatter(
    const SceneDesc *__restrict__ scenes, int nsc, const uint32_t *key, const uint32_t *slot,
    const uint32_t *start, const uint32_t *block_out_scan, uint32_t *perm,
    const uint8_t *oct, uint8_t *poct) {
    const uint32_t blk = blockIdx.x;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, blk, [](const SceneDesc &d) { return d.blk0; })];
    const int64_t n = sd.n;
    const int64_t base = (int64_t)(blk - sd.blk0) * kBinTile + threadIdx.x;
    const uint32_t ovf = sd.b0 + sd.bg.nbin;
    uint32_t b[kBinItems], sl[kBinItems], st[kBinItems];
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        b[m] = i < n ? key[sd.p0 + i] : ovf;
        sl[m] = i < n ? slot[sd.p0 + i] : 0u;
    }
    // the scene's overflow bin: this block's particles follow the overflow
    // particles of the scene's earlier blocks
    const uint32_t over0 = start[ovf] + block_out_scan[blk] - block_out_scan[sd.blk0];
#pragma unroll
    for (int m = 0; m < kBinItems; m++) st[m] = b[m] == ovf ? over0 : start[b[m]];
#pragma unroll
    for (int m = 0; m < kBinItems; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        if (i < n) {
            const uint32_t dst = st[m] + sl[m];
            perm[dst] = (uint32_t)(sd.p0 + i);
            if (poct) poct[dst] = oct[sd.p0 + i];
        }
    }
}


// The sort key of slot q: the particle index, or (octant, index) when
// octants are given.  Keys of one bin are distinct.
__device__ __forceinline__ uint64_t order_key(const uint32_t *perm,
                                              const uint8_t *poct, uint32_t q) {
    const uint64_t o = poct ? (uint64_t)poct[q] : 0ull;
    return (o << 32) | (uint64_t)perm[q];
}

// Where ordered bins go: the indices (perm), and with octant order each
// particle's rank inside its half-cell run (hrank) and its octant (soct).
struct OrderOut {
    uint32_t *perm;
    uint32_t *hrank;   // nullable
    uint8_t *soct;     // nullable
};

// Write entry e of a bin starting at slot s, given its key and the first
// entry of its octant run.
__device__ __forceinline__ void put_sorted(const OrderOut &out, uint32_t s, uint32_t e,
                                           uint64_t key, uint32_t run_start) {
    out.perm[s + e] = (uint32_t)key;
    if (out.hrank) out.hrank[s + e] = e - run_start;
    if (out.soct) out.soct[s + e] = (uint8_t)(key >> 32);
}

// Keys in 32 bits when every index fits in kNarrowBits: (octant, index)
// as octant << kNarrowBits | index.
constexpr int kNarrowBits = 29;

__device__ __forceinline__ uint32_t order_key32(const uint32_t *perm, const uint8_t *poct,
                                                uint32_t q) {
    const uint32_t o = poct ? (uint32_t)poct[q] : 0u;
    return (o << kNarrowBits) | perm[q];
}

__device__ __forceinline__ uint64_t widen_key(uint64_t k) { return k; }
__device__ __forceinline__ uint64_t widen_key(uint32_t k) {
    return ((uint64_t)(k >> kNarrowBits) << 32) | (k & ((1u << kNarrowBits) - 1u));
}

// Write a bin's sorted keys.
template <int N, class T>
__device__ __forceinline__ void write_sorted(const T (&v)[N], uint32_t s,
                                             uint32_t len, const OrderOut &out) {
    uint32_t rs = 0;
#pragma unroll
    for (int m = 0; m < N; m++) {
        const uint64_t w = widen_key(v[m]);
        if (m > 0 && (w >> 32) != (widen_key(v[m - 1]) >> 32)) rs = (uint32_t)m;
        if ((uint32_t)m < len) put_sorted(out, s, (uint32_t)m, w, rs);
    }
}

// P2G bricks (edge kMarkBrick cells) whose staged cells (the brick plus one
// cell around it) hold a particle: flag[k0 + brick] = 1.
constexpr int kMarkBrick = 4;

// Local bin lb of scene sd (cell = bin - 1 per axis) is staged by the bricks
// bx with 4bx - 1 <= cell <= 4bx + 4, i.e. 4bx <= bin <= 4bx + 5: from
// ceil((bin - 5)/4) = floor((bin - 2)/4) to floor(bin/4); every one of them
// is marked (plain stores of 1).
__device__ __forceinline__ void mark_bricks(const SceneDesc &sd, uint8_t *flag, uint32_t lb) {
    const uint32_t b0 = (uint32_t)sd.bg.b[0], b1 = (uint32_t)sd.bg.b[1];
    const int idx[3] = {(int)(lb % b0), (int)((lb / b0) % b1), (int)(lb / (b0 * b1))};
    const int nb[3] = {sd.nbx, sd.nby, sd.nbz};
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = idx[a] >= 2 ? (idx[a] - 2) / kMarkBrick : 0;
        hi[a] = min(idx[a] / kMarkBrick, nb[a] - 1);
    }
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++)
                flag[sd.k0 + x + nb[0] * (y + nb[1] * z)] = 1;
}

constexpr uint32_t kSmallBin = 16;   // bins up to this size: a thread, in registers
constexpr uint32_t kMidBin = 256;    // bins up to this size: a wave, in registers

// Order every in-grid bin's entries by key (see order_key), out of place
// from perm_in to perm_out.  A thread owns a bin: bins of up to kSmallBin
// entries are sorted in registers by a bitonic network (the network size is
// the largest bin of the wave, rounded up to 8 or 16); larger bins are
// listed for k_order_mid (up to kMidBin entries, from the front of
// big_list, count in big_count[0]) and k_order_big (from the back, count in
// big_count[1]).
template <bool NARROW>
__global__ __launch_bounds__(kBlock) void k_order_bins(
    const SceneDesc *__restrict__ scenes, int nsc, const uint32_t *__restrict__ start,
    uint32_t nbin, const uint32_t *__restrict__ perm_in, const uint8_t *__restrict__ poct,
    OrderOut out, uint32_t *__restrict__ big_list, uint32_t big_cap,
    uint32_t *__restrict__ big_count, uint8_t *__restrict__ brick_flag) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    uint32_t s = 0, len = 0;
    if (b < nbin) {
        s = start[b];
        len = start[b + 1] - s;
    }
    if (len > 0) {
        const SceneDesc &sd = scenes[scene_of(scenes, nsc, b, [](const SceneDesc &d) { return d.b0; })];
        const uint32_t lb = b - sd.b0;
        if (lb == sd.bg.nbin) len = 0;   // overflow bin: already in index order (k_order_big copies it)
        else if (brick_flag) mark_bricks(sd, brick_flag, lb);
    }
    if (len > kSmallBin) {
        // the wave-sorted sizes go to the front of the list, the rest to
        // the back
        if (len <= kMidBin) big_list[atomicAdd(&big_count[0], 1u)] = b;
        else big_list[big_cap - 1u - atomicAdd(&big_count[1], 1u)] = b;
        len = 0;
    }
    if (len == 1) {
        put_sorted(out, s, 0u, order_key(perm_in, poct, s), 0u);
        return;
    }
    // wave-uniform network size
    using KT = typename std::conditional<NARROW, uint32_t, uint64_t>::type;
    auto key = [&](uint32_t q) -> KT {
        if (NARROW) return (KT)order_key32(perm_in, poct, q);
        return (KT)order_key(perm_in, poct, q);
    };
    const bool need16 = __any(len > 8u), need8 = __any(len > 1u);
    if (need16) {
        KT v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = (uint32_t)m < len ? key(s + m) : (KT)~(KT)0;
        bitonic_sort<16>(v);
        write_sorted<16>(v, s, len, out);
    } else if (need8) {
        KT v[8];
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = (uint32_t)m < len ? key(s + m) : (KT)~(KT)0;
        bitonic_sort<8>(v);
        write_sorted<8>(v, s, len, out);
    }
}


// Sort one bin of 16 < len <= 64*K entries with the calling wave and write
// it out.
template <int K>
__device__ __forceinline__ void wave_order_bin(uint32_t s, uint32_t len,
                                               const uint32_t *__restrict__ perm_in,
                                               const uint8_t *__restrict__ poct,
                                               const OrderOut &out) {
    const int lane = threadIdx.x & 63;
    uint64_t v[K];
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        v[r] = e < len ? order_key(perm_in, poct, s + e) : ~0ull;
    }
    wave_bitonic<K>(v);
    // first element of each octant run: the running maximum of the run
    // starts seen so far, over registers then lanes
    const uint32_t prev_hi = __shfl_up((uint32_t)(v[K - 1] >> 32), 1, 64);
    uint32_t rs[K];
    uint32_t m = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        const uint32_t before = r > 0 ? (uint32_t)(v[r - 1] >> 32) : prev_hi;
        const bool first = e == 0 || before != (uint32_t)(v[r] >> 32);
        m = first ? e : m;
        rs[r] = m;
    }
    // carry the last run start of lower lanes into this lane's registers
    uint32_t carry = m;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(carry, d, 64);
        if (lane >= d) carry = max(carry, t);
    }
    // (the shuffle runs on every lane: evaluated under lane > 0, lane 0
    // would be inactive and lane 1 would read 0 instead of its value)
    const uint32_t carry_up = __shfl_up(carry, 1, 64);
    const uint32_t from_below = lane > 0 ? carry_up : 0u;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        if (e < len) put_sorted(out, s, e, v[r], max(rs[r], from_below));
    }
}

// Bins of kSmallBin < len <= kMidBin entries: one wave per bin, sorted in
// registers (64 or 256 keys per wave).
__global__ __launch_bounds__(kBlock) void k_order_mid(
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm_in,
    const uint8_t *__restrict__ poct, OrderOut out, const uint32_t *__restrict__ big_list,
    const uint32_t *__restrict__ big_count) {
    const uint32_t nmid = big_count[0];
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t c = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); c < nmid; c += waves) {
        const uint32_t b = big_list[c];
        const uint32_t s = start[b], len = start[b + 1] - s;
        if (len <= 64u) wave_order_bin<1>(s, len, perm_in, poct, out);
        else wave_order_bin<4>(s, len, perm_in, poct, out);
    }
}

constexpr int kBigTile = 2048;   // keys per LDS tile in k_order_big
constexpr int kOrderBigBlocks = 1024;
constexpr int kOrderMidBlocks = 4096;

// Bins of more than kMidBin entries: one wave per bin ranks every entry by
// counting the smaller keys of its bin (tiles of the keys pass through LDS,
// so any size works).  Also copies the scenes' overflow bins, which k_bin
// already filled in index order.
__global__ __launch_bounds__(64) void k_order_big(
    const SceneDesc *__restrict__ scenes, int nsc, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ perm_in, const uint8_t *__restrict__ poct, OrderOut out,
    const uint32_t *__restrict__ big_list, uint32_t big_cap,
    const uint32_t *__restrict__ big_count) {
    __shared__ uint64_t tile[kBigTile];
    const uint32_t lane = threadIdx.x;
    for (int sc = 0; sc < nsc; sc++) {
        const uint32_t ovf = scenes[sc].b0 + scenes[sc].bg.nbin;
        const uint32_t o0 = start[ovf], o1 = start[ovf + 1];
        for (uint32_t q = o0 + blockIdx.x * 64u + lane; q < o1; q += gridDim.x * 64u)
            out.perm[q] = perm_in[q];
    }
    const uint32_t nbig = big_count[1];
    for (uint32_t c = blockIdx.x; c < nbig; c += gridDim.x) {
        const uint32_t b = big_list[big_cap - 1u - c];
        const uint32_t s = start[b], len = start[b + 1] - s;
        for (uint32_t e0 = 0; e0 < len; e0 += 64u) {
            const uint32_t e = e0 + lane;
            const uint64_t mine = e < len ? order_key(perm_in, poct, s + e) : ~0ull;
            uint32_t rank = 0, hr = 0;
            for (uint32_t t0 = 0; t0 < len; t0 += kBigTile) {
                const uint32_t tl = min(len - t0, (uint32_t)kBigTile);
                __syncthreads();
                for (uint32_t k = lane; k < tl; k += 64u) tile[k] = order_key(perm_in, poct, s + t0 + k);
                __syncthreads();
                for (uint32_t k = 0; k < tl; k++) {
                    const uint64_t kk = tile[k];
                    const bool lt = kk < mine;
                    rank += lt;
                    hr += lt && (kk >> 32) == (mine >> 32);
                }
            }
            if (e < len) put_sorted(out, s, rank, mine, rank - hr);
        }
    }
}

void cell_sort(const SortBuffers &sb, int64_t n, hipStream_t s, bool octant_order,
               bool mark_bricks) {
    fill_u32(sb.count, 0u, (int64_t)sb.nbin_total, s);
    if (n <= 0) {
        fill_u32(sb.start, 0u, (int64_t)sb.nbin_total + 1, s);
        return;
    }
    uint8_t *oct = octant_order ? sb.oct : nullptr;
    uint8_t *poct = octant_order ? sb.poct : nullptr;
    const unsigned nb = sb.nblk_total;
    hipLaunchKernelGGL(k_bin, dim3(nb), dim3(kBlock), 0, s, sb.scenes, sb.nscene, sb.count,
                       sb.key, sb.slot, sb.block_out, oct);
    GFSX_LAUNCH_CHECK();
    exclusive_scan(sb.block_out, sb.block_out_scan, nb, sb.scan_scratch, s);
    hipLaunchKernelGGL(k_overflow_counts, dim3(grid_for(sb.nscene, 64)), dim3(64), 0, s,
                       sb.scenes, sb.nscene, (const uint32_t *)sb.block_out_scan, nb, sb.count);
    GFSX_LAUNCH_CHECK();
    exclusive_scan(sb.count, sb.start, (int64_t)sb.nbin_total, sb.scan_scratch, s);
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, s, sb.scenes, sb.nscene,
                       (const uint32_t *)sb.key, (const uint32_t *)sb.slot,
                       (const uint32_t *)sb.start, (const uint32_t *)sb.block_out_scan,
                       sb.perm_tmp, (const uint8_t *)oct, poct);
    GFSX_LAUNCH_CHECK();
    fill_u32(sb.big_count, 0u, 2, s);
    OrderOut out;
    out.perm = sb.perm;
    out.hrank = octant_order ? sb.hrank : nullptr;
    out.soct = octant_order ? sb.soct : nullptr;
    uint8_t *flag = mark_bricks ? sb.brick_flag : nullptr;
    auto *order = n < ((int64_t)1 << kNarrowBits) ? k_order_bins<true> : k_order_bins<false>;
    hipLaunchKernelGGL(order, dim3(grid_for(sb.nbin_total, kBlock)), dim3(kBlock), 0, s,
                       sb.scenes, sb.nscene, (const uint32_t *)sb.start, sb.nbin_total,
                       (const uint32_t *)sb.perm_tmp, (const uint8_t *)poct, out, sb.big_list,
                       sb.big_cap, sb.big_count, flag);
    GFSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_order_mid, dim3(kOrderMidBlocks), dim3(kBlock), 0, s,
                       (const uint32_t *)sb.start, (const uint32_t *)sb.perm_tmp,
                       (const uint8_t *)poct, out, (const uint32_t *)sb.big_list,
                       (const uint32_t *)sb.big_count);
    GFSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_order_big, dim3(kOrderBigBlocks), dim3(64), 0, s, sb.scenes,
                       sb.nscene, (const uint32_t *)sb.start, (const uint32_t *)sb.perm_tmp,
                       (const uint8_t *)poct, out, (const uint32_t *)sb.big_list, sb.big_cap,
                       (const uint32_t *)sb.big_count);
    GFSX_LAUNCH_CHECK();
}

// ===========================================================================
// P2G gather
// ===========================================================================

// Sum K and K*val over the candidate bins of node (i,j,k).  Bins along x in
// one (y,z) row are contiguous in perm, so each row is one index range; the
// order of the sum is fixed: rows in (k,j) order, then perm order.
template <bool HAS_VAL, class Load>
__device__ __forceinline__ void gather_node(int i, int j, int k,
                                            const NodeGrid &ng,
                                            const BinGrid &bg, double dx,
                                            float inv_r2,
                                            const uint32_t *__restrict__ start,
                                            const uint32_t *__restrict__ perm,
                                            Load load, float &S, float &W) {
    // node position in the shifted frame, rounded as
    // Grid3d::GridIndexToPosition does (grid3d.h:79-81)
    const float gx = (float)((float)i * dx);
    const float gy = (float)((float)j * dx);
    const float gz = (float)((float)k * dx);
    const int ilo = max(i + ng.dlo[0], 0), ihi = min(i + ng.dhi[0], bg.b[0] - 1);
    const int jlo = max(j + ng.dlo[1], 0), jhi = min(j + ng.dhi[1], bg.b[1] - 1);
    const int klo = max(k + ng.dlo[2], 0), khi = min(k + ng.dhi[2], bg.b[2] - 1);
    if (ilo > ihi) return;
    for (int bk = klo; bk <= khi; bk++) {
        for (int bj = jlo; bj <= jhi; bj++) {
            const uint32_t row = (uint32_t)bg.b[0] * ((uint32_t)bj + (uint32_t)bg.b[1] * (uint32_t)bk);
            const uint32_t q0 = start[row + ilo], q1 = start[row + ihi + 1];
            for (uint32_t q = q0; q < q1; q++) {
                const uint32_t p = perm ? perm[q] : q;
                float px, py, pz, val;
                load(p, px, py, pz, val);
                // p -= offset, then v = gpos - p (scalarfield.cpp:168,186-188)
                px -= ng.off[0];
                py -= ng.off[1];
                pz -= ng.off[2];
                const float vx = gx - px, vy = gy - py, vz = gz - pz;
                const float d2 = vx * vx + vy * vy + vz * vz;
                const float t = d2 * inv_r2;
                const float K = wyvill(t);
                // only inside the radius (scalarfield.cpp:191-196): a value
                // that is not finite must not reach the nodes beyond it
                if (HAS_VAL) S = fmaf(K, t < 1.0f ? val : 0.0f, S);
                else S += K;
                W += K;
            }
        }
    }
}

struct AosLoad {
    const float *__restrict__ pos;
    const float *__restrict__ val;
    __device__ __forceinline__ void operator()(uint32_t p, float &x, float &y,
                                               float &z, float &v) const {
        x = pos[3 * (size_t)p];
        y = pos[3 * (size_t)p + 1];
        z = pos[3 * (size_t)p + 2];
        v = val ? val[p] : 1.0f;
    }
};

template <bool HAS_VAL>
__global__ __launch_bounds__(kBlock) void k_p2g_gather(
    AosLoad ld, BinGrid bg, NodeGrid ng, double dx, float inv_r2,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm,
    float *__restrict__ S, float *__restrict__ W) {
    const int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t nnode = (int64_t)ng.n[0] * ng.n[1] * ng.n[2];
    if (f >= nnode) return;
    const int i = (int)(f % ng.n[0]);
    const int j = (int)((f / ng.n[0]) % ng.n[1]);
    const int k = (int)(f / ((int64_t)ng.n[0] * ng.n[1]));
    float s = 0.0f, w = 0.0f;
    gather_node<HAS_VAL>(i, j, k, ng, bg, dx, inv_r2, start, perm, ld, s, w);
    // accumulate into the caller's arrays (clscalarfield.cpp:1448-1449)
    S[f] += s;
    if (W) W[f] += w;
}

void p2g_gather(const float *aos, const float *val, int64_t n,
                const BinGrid &bins, const NodeGrid &nodes, double radius,
                double dx, const SortBuffers &sb, float *S, float *W,
                hipStream_t s) {
    (void)n;
    int64_t nnode = (int64_t)nodes.n[0] * nodes.n[1] * nodes.n[2];
    if (nnode <= 0) return;
    float inv_r2 = (float)(1.0 / (radius * radius));
    AosLoad ld{aos, val};
    if (val)
        hipLaunchKernelGGL(k_p2g_gather<true>, dim3(grid_for(nnode, kBlock)),
                           dim3(kBlock), 0, s, ld, bins, nodes, dx, inv_r2,
                           (const uint32_t *)sb.start, (const uint32_t *)sb.perm,
                           S, W);
    else
        hipLaunchKernelGGL(k_p2g_gather<false>, dim3(grid_for(nnode, kBlock)),
                           dim3(kBlock), 0, s, ld, bins, nodes, dx, inv_r2,
                           (const uint32_t *)sb.start, (const uint32_t *)sb.perm,
                           S, W);
    GFSX_LAUNCH_CHECK();
}

__global__ void k_apply_weight(float *S, const float *W, int64_t count) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    float w = W[i];
    if (w > 0.0f) S[i] = S[i] / w;
}

void apply_weight(float *S, const float *W, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_apply_weight, dim3(grid_for(count, kBlock)),
                       dim3(kBlock), 0, s, S, W, count);
    GFSX_LAUNCH_CHECK();
}

// ===========================================================================
// G2P and RK kernels (samplers in gfsx_sample.h)
// ===========================================================================

template <int ST>
__global__ __launch_bounds__(kBlock) void k_g2p(const float *pos, int64_t n,
                                               MacGrid g, float *out) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= n) return;
    const float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
    float v[3];
    sample_mac<ST, false>(g, nullptr, nullptr, nullptr, x, y, z, v, nullptr);
    validate3(v);
    out[3 * p] = v[0];
    out[3 * p + 1] = v[1];
    out[3 * p + 2] = v[2];
}

void g2p(const float *pos, int64_t n, const MacGrid &g, int stencil,
         float *out, hipStream_t s) {
    if (n <= 0) return;
    if (stencil == 0)
        hipLaunchKernelGGL(k_g2p<0>, dim3(grid_for(n, kBlock)), dim3(kBlock),
                           0, s, pos, n, g, out);
    else
        hipLaunchKernelGGL(k_g2p<1>, dim3(grid_for(n, kBlock)), dim3(kBlock),
                           0, s, pos, n, g, out);
    GFSX_LAUNCH_CHECK();
}



template <int ST>
__global__ __launch_bounds__(kBlock) void k_advect(const float *pos, int64_t n,
                                                  MacGrid g, int order,
                                                  RkConsts c, float *out) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= n) return;
    const float x0[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    float k1[3], unused[3], x1[3];
    sample_mac<ST, false>(g, nullptr, nullptr, nullptr, x0[0], x0[1], x0[2], k1, unused);
    rk_step<ST>(g, order, c, x0, k1, x1);
    out[3 * p] = x1[0];
    out[3 * p + 1] = x1[1];
    out[3 * p + 2] = x1[2];
}

static RkConsts rk_consts(double dt) {
    RkConsts c;
    c.half = (float)(0.5 * dt);
    c.three_q = (float)(0.75 * dt);
    c.full = (float)dt;
    c.sixth = (float)(dt / 6.0f);
    c.ninth = (float)(dt / 9.0f);
    return c;
}

void advect(const float *pos, int64_t n, const MacGrid &g, int order,
            double dt, int stencil, float *out, hipStream_t s) {
    if (n <= 0) return;
    RkConsts c = rk_consts(dt);
    if (stencil == 0)
        hipLaunchKernelGGL(k_advect<0>, dim3(grid_for(n, kBlock)), dim3(kBlock),
                           0, s, pos, n, g, order, c, out);
    else
        hipLaunchKernelGGL(k_advect<1>, dim3(grid_for(n, kBlock)), dim3(kBlock),
                           0, s, pos, n, g, order, c, out);
    GFSX_LAUNCH_CHECK();
}

// ===========================================================================
// The fused plain step
// ===========================================================================

struct SoaLoad {
    const float *__restrict__ x;
    const float *__restrict__ y;
    const float *__restrict__ z;
    const float *__restrict__ v;
    __device__ __forceinline__ void operator()(uint32_t p, float &px, float &py,
                                               float &pz, float &pv) const {
        px = x[p];
        py = y[p];
        pz = z[p];
        pv = v[p];
    }
};

// One thread per face of u, v or w: gather, normalise
// (scalarfield.cpp:90-106), snapshot (fluidsimulation.cpp:3306), then add
// the body force where the face borders a cell that holds a particle
// (fluidsimulation.cpp:2766-2805, fluidmaterialgrid.cpp:119-143).
__global__ __launch_bounds__(kBlock) void k_flip_p2g(
    StepP2G a, BinGrid bg, double dx, float inv_r2,
    const uint32_t *__restrict__ count, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ perm) {
    int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    int c = 0;
    if (f >= a.nface[0]) {
        f -= a.nface[0];
        c = 1;
        if (f >= a.nface[1]) {
            f -= a.nface[1];
            c = 2;
            if (f >= a.nface[2]) return;
        }
    }
    const NodeGrid &ng = a.ng[c];
    const int i = (int)(f % ng.n[0]);
    const int j = (int)((f / ng.n[0]) % ng.n[1]);
    const int k = (int)(f / ((int64_t)ng.n[0] * ng.n[1]));
    float S = 0.0f, W = 0.0f;
    SoaLoad ld{a.x, a.y, a.z, a.vel[c]};
    gather_node<true>(i, j, k, ng, bg, dx, inv_r2, start, perm, ld, S, W);
    float val = W > 0.0f ? S / W : S;
    if (a.grid_old[c]) a.grid_old[c][f] = val;
    if (a.wgt[c]) a.wgt[c][f] = W;
    if (a.grav_on[c]) {
        // the two cells on either side of the face along axis c
        const int ax = c == 0 ? i : (c == 1 ? j : k);
        const int nax = c == 0 ? a.ni : (c == 1 ? a.nj : a.nk);
        bool fluid = false;
        // bin of in-domain cell (ci,cj,ck) is (ci+1, cj+1, ck+1)
        const uint32_t bhi = (uint32_t)(i + 1) + (uint32_t)bg.b[0] *
                             ((uint32_t)(j + 1) + (uint32_t)bg.b[1] * (uint32_t)(k + 1));
        const uint32_t step = c == 0 ? 1u : (c == 1 ? (uint32_t)bg.b[0]
                                                    : (uint32_t)bg.b[0] * (uint32_t)bg.b[1]);
        if (ax < nax) fluid = fluid || count[bhi] > 0u;
        if (ax > 0) fluid = fluid || count[bhi - step] > 0u;
        if (fluid) val += a.grav_delta[c];
    }
    a.grid[c][f] = val;
}

// ---------------------------------------------------------------------------
// Brick-staged P2G gather for the fused step.
//
// One work-group per 4x4x4-cell brick.  It stages every particle of the
// brick plus a one-cell halo (6x6x6 cells) into LDS, re-laid out in
// half-cell-major order: half-cell (2c+o) per axis, where o says in which
// half of its cell the particle lies.  A face then has all its candidate
// particles in 16 contiguous LDS ranges -- the 4x4x4 half-cells whose
// union is the 2x2x2 "shifted cells" around it -- instead of 18 whole
// cells: 64 candidates instead of 144 at 8 particles per cell.  The order
// inside a half-cell is the order of the sorted particle list, so the sum
// for every face is taken in a fixed order.
// ---------------------------------------------------------------------------
namespace brick {
constexpr int BX = 4, BY = 4, BZ = 4;
constexpr int SX = BX + 2, SY = BY + 2, SZ = BZ + 2;  // staged cells
constexpr int NROW = SY * SZ;
constexpr int HX = 2 * SX, HY = 2 * SY, HZ = 2 * SZ;  // half-cells
constexpr int NH = HX * HY * HZ;
constexpr int NCELL = BX * BY * BZ;       // 64 cells: one wave per component
constexpr int NT = 3 * NCELL;             // 192 threads (dense pass)
constexpr int NTC = 4 * NCELL;            // 256 threads (common pass)
// Particles staged per chunk (6 floats each); a brick at 8 particles per
// cell stages 6x6x6 cells * 8 = 1728.
constexpr int CAP = 2016;
constexpr int DENSE_BLOCKS = 2048;   // work-groups of the pass over dense bricks
constexpr int DENSE_CAP = 1344;      // its chunk size (smaller: more work-groups per CU)
static_assert(NCELL == 64, "one wave per component");
static_assert(BX == kMarkBrick && BY == kMarkBrick && BZ == kMarkBrick, "the sort marks these bricks");
}  // namespace brick

__device__ __forceinline__ void face_epilogue(const StepP2G &a,
                                              const BinGrid &bg,
                                              const uint32_t *__restrict__ count,
                                              int c, int i, int j, int k,
                                              float S, float W) {
    const NodeGrid &ng = a.ng[c];
    const int64_t f = (int64_t)i + (int64_t)ng.n[0] * ((int64_t)j + (int64_t)ng.n[1] * k);
    float val = W > 0.0f ? S / W : S;
    if (a.grid_old[c]) a.grid_old[c][f] = val;
    if (a.wgt[c]) a.wgt[c][f] = W;
    if (a.grav_on[c]) {
        const int ax = c == 0 ? i : (c == 1 ? j : k);
        const int nax = c == 0 ? a.ni : (c == 1 ? a.nj : a.nk);
        bool fluid = false;
        const uint32_t bhi = (uint32_t)(i + 1) + (uint32_t)bg.b[0] *
                             ((uint32_t)(j + 1) + (uint32_t)bg.b[1] * (uint32_t)(k + 1));
        const uint32_t step = c == 0 ? 1u : (c == 1 ? (uint32_t)bg.b[0]
                                                    : (uint32_t)bg.b[0] * (uint32_t)bg.b[1]);
        if (ax < nax) fluid = fluid || count[bhi] > 0u;
        if (ax > 0) fluid = fluid || count[bhi - step] > 0u;
        if (fluid) val += a.grav_delta[c];
    }
    a.grid[c][f] = val;
}


// face_epilogue for a brick face: whether the cells on either side hold a
// particle comes from the brick's occupancy table (staged cells, x fastest)
// instead of the global counts.
__device__ __forceinline__ void brick_epilogue(const FaceOut &o, const uint8_t *occ,
                                               int c, int i, int j, int k, int i0, int j0,
                                               int k0, float S, float W) {
    using namespace brick;
    const int64_t f = (int64_t)i + (int64_t)o.n0 * ((int64_t)j + (int64_t)o.n1 * k);
    float val = W > 0.0f ? S / W : S;
    if (o.old) o.old[f] = val;
    if (o.wgt) o.wgt[f] = W;
    if (o.grav_on) {
        const int ax = c == 0 ? i : (c == 1 ? j : k);
        // staged cell of (i,j,k); the cell below along c is one step back
        const int hi = (i - i0 + 1) + SX * ((j - j0 + 1) + SY * (k - k0 + 1));
        const int step = c == 0 ? 1 : (c == 1 ? SX : SX * SY);
        bool fluid = false;
        if (ax < o.nax) fluid = fluid || occ[hi] != 0;
        if (ax > 0) fluid = fluid || occ[hi - step] != 0;
        if (fluid) val += o.grav_delta;
    }
    o.grid[f] = val;
}

// Range r (0..15) of the 4x4 (y,z) rows of half-cells around a face; each
// row covers 4 consecutive half-cells along x, contiguous in LDS.
__device__ __forceinline__ void brick_range(const uint16_t *__restrict__ hs,
                                            int hx0, int hy0, int hz0, int r,
                                            uint32_t &q, uint32_t &e) {
    using namespace brick;
    const int base = hx0 + HX * ((hy0 + (r & 3)) + HY * (hz0 + (r >> 2)));
    q = hs[base];
    e = hs[base + 4];
}

// Gather over the 4x4x4 half-cells around face (li,lj,lk) (staged cell
// coordinates) of component C, as one stream of candidates per lane: the
// lane steps to its next non-empty range when it exhausts one, so a wave
// runs for the longest lane's total, not for the sum of per-range maxima.
template <int C>
__device__ __forceinline__ void brick_gather(const float *__restrict__ sx,
                                             const float *__restrict__ sy,
                                             const float *__restrict__ sz,
                                             const float *__restrict__ sv,
                                             const uint16_t *__restrict__ hs,
                                             int li, int lj, int lk, float gx,
                                             float gy, float gz, float offh,
                                             float inv_r2, float &S, float &W) {
    const int hx0 = C == 0 ? 2 * li - 2 : 2 * li - 1;
    const int hy0 = C == 1 ? 2 * lj - 2 : 2 * lj - 1;
    const int hz0 = C == 2 ? 2 * lk - 2 : 2 * lk - 1;
    int r = 0;
    uint32_t q, e;
    brick_range(hs, hx0, hy0, hz0, 0, q, e);
    while (true) {
        while (q >= e && r < 15) {
            r++;
            brick_range(hs, hx0, hy0, hz0, r, q, e);
        }
        if (q >= e) break;
        float px = sx[q], py = sy[q], pz = sz[q];
        const float pv = sv[q];
        // p -= offset (scalarfield.cpp:168); the offset is zero on the
        // component's own axis
        if (C != 0) px -= offh;
        if (C != 1) py -= offh;
        if (C != 2) pz -= offh;
        const float vx = gx - px, vy = gy - py, vz = gz - pz;
        const float d2 = vx * vx + vy * vy + vz * vz;
        const float t = d2 * inv_r2;
        const float K = wyvill(t);
        S = fmaf(K, t < 1.0f ? pv : 0.0f, S);   // only inside the radius
        W += K;
        q++;
    }
}

// The same gather as 16 nested range loops.
template <int C>
__device__ __forceinline__ void brick_gather_nested(const float *__restrict__ sx,
                                                    const float *__restrict__ sy,
                                                    const float *__restrict__ sz,
                                                    const float *__restrict__ sv,
                                                    const uint16_t *__restrict__ hs,
                                                    int li, int lj, int lk, float gx,
                                                    float gy, float gz, float offh,
                                                    float inv_r2, float &S, float &W) {
    const int hx0 = C == 0 ? 2 * li - 2 : 2 * li - 1;
    const int hy0 = C == 1 ? 2 * lj - 2 : 2 * lj - 1;
    const int hz0 = C == 2 ? 2 * lk - 2 : 2 * lk - 1;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint32_t q, e;
        brick_range(hs, hx0, hy0, hz0, r, q, e);
        for (; q < e; q++) {
            float px = sx[q], py = sy[q], pz = sz[q];
            const float pv = sv[q];
            if (C != 0) px -= offh;
            if (C != 1) py -= offh;
            if (C != 2) pz -= offh;
            const float vx = gx - px, vy = gy - py, vz = gz - pz;
            const float d2 = vx * vx + vy * vy + vz * vz;
            const float t = d2 * inv_r2;
            const float K = wyvill(t);
            S = fmaf(K, t < 1.0f ? pv : 0.0f, S);   // only inside the radius
            W += K;
        }
    }
}

// The same gather range by range, candidates four at a time.  The four
// reads of a chunk are consecutive (merged into paired LDS reads); reads past
// the range's end land on other staged data (or the next LDS array) and are
// masked: their weight and value are selected to zero, so nothing of them
// reaches the sums.  The order of the sum per face is fixed (ranges in
// order, then candidates in order).
template <int C>
__device__ __forceinline__ void brick_gather_chunked(const float *__restrict__ sx,
                                                     const float *__restrict__ sy,
                                                     const float *__restrict__ sz,
                                                     const float *__restrict__ sv,
                                                     const uint16_t *__restrict__ hs,
                                                     int li, int lj, int lk, float gx,
                                                     float gy, float gz, float offh,
                                                     float inv_r2, float &S, float &W) {
    const int hx0 = C == 0 ? 2 * li - 2 : 2 * li - 1;
    const int hy0 = C == 1 ? 2 * lj - 2 : 2 * lj - 1;
    const int hz0 = C == 2 ? 2 * lk - 2 : 2 * lk - 1;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint32_t q, e;
        brick_range(hs, hx0, hy0, hz0, r, q, e);
        for (; q < e; q += 4u) {
            float px[4], py[4], pz[4], pv[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                px[c] = sx[q + c];
                py[c] = sy[q + c];
                pz[c] = sz[q + c];
                pv[c] = sv[q + c];
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const bool valid = q + (uint32_t)c < e;
                // p -= offset (scalarfield.cpp:168); the offset is zero on
                // the component's own axis
                if (C != 0) px[c] -= offh;
                if (C != 1) py[c] -= offh;
                if (C != 2) pz[c] -= offh;
                const float vx = gx - px[c], vy = gy - py[c], vz = gz - pz[c];
                const float d2 = vx * vx + vy * vy + vz * vz;
                const float t = d2 * inv_r2;
                float K = wyvill(t);
                K = valid ? K : 0.0f;
                // values only inside the radius (scalarfield.cpp:191-196)
                const float v = valid && t < 1.0f ? pv[c] : 0.0f;
                S = fmaf(K, v, S);
                W += K;
            }
        }
    }
}

// Row of staged slot s: the number of row starts (after the first) that
// are <= s.  All lanes read the same LDS words, so the reads broadcast and
// do not depend on one another.
__device__ __forceinline__ int brick_row_of(const uint32_t *row_s0, uint32_t s) {
    int r = 0;
#pragma unroll
    for (int q = 1; q < brick::NROW; q++) r += row_s0[q] <= s;
    return r;
}

// Add the staged candidates of face (li,lj,lk) (brick-local) of component C
// to S and W.
template <int C, int V>
__device__ __forceinline__ void brick_accum(const StepP2G &a, const float *sx,
                                            const float *sy, const float *sz,
                                            const float *sv, const uint16_t *hs,
                                            int i0, int j0, int k0, int li, int lj,
                                            int lk, double dx, float inv_r2,
                                            float &S, float &W) {
    // face position in the component's shifted frame (grid3d.h:79-81)
    const float gx = (float)((float)(i0 + li) * dx);
    const float gy = (float)((float)(j0 + lj) * dx);
    const float gz = (float)((float)(k0 + lk) * dx);
    const float offh = a.ng[C].off[C == 0 ? 1 : 0];
    if (V == 0)
        brick_gather_chunked<C>(sx, sy, sz, sv, hs, li + 1, lj + 1, lk + 1, gx, gy, gz, offh, inv_r2, S, W);
    else if (V == 2)
        brick_gather_nested<C>(sx, sy, sz, sv, hs, li + 1, lj + 1, lk + 1, gx, gy, gz, offh, inv_r2, S, W);
    else
        brick_gather<C>(sx, sy, sz, sv, hs, li + 1, lj + 1, lk + 1, gx, gy, gz, offh, inv_r2, S, W);
}

__device__ __forceinline__ void brick_zero_face(const StepP2G &a, const FaceOut &o, int c,
                                                int fi, int fj, int fk) {
    if (fi >= a.ni + (c == 0) || fj >= a.nj + (c == 1) || fk >= a.nk + (c == 2)) return;
    const int64_t f = (int64_t)fi + (int64_t)o.n0 * ((int64_t)fj + (int64_t)o.n1 * fk);
    if (o.old) o.old[f] = 0.0f;
    if (o.wgt) o.wgt[f] = 0.0f;
    o.grid[f] = 0.0f;
}

// One work-group per brick of BX*BY*BZ cells; wave w computes the faces of
// component w.  The particles of the staged cells (the brick plus one cell
// around it) are copied into LDS in half-cell-major order, CAP at a time.
// With CHUNKED false a brick holding more than CAP is appended to dense[]
// and left for the CHUNKED pass, which processes it in chunks of
// consecutive sort slots, the faces accumulating over the chunks (a
// separate kernel, so that the common case keeps its register budget).
template <int V, int CAP, bool CHUNKED>
__device__ __forceinline__ void p2g_brick(
    const StepP2G &a, const BinGrid &bg, double dx, float inv_r2, int nbx, int nby,
    int brick_id, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ hrank,
    uint32_t *__restrict__ dense, uint32_t *__restrict__ dense_count,
    float *s_buf, uint16_t *s_hs, uint32_t *s_rowq, uint32_t *s_rows0, uint32_t *s_wsum,
    uint8_t *s_occ) {
    using namespace brick;
    constexpr int PER = (CAP + NT - 1) / NT;
    float *s_x = s_buf, *s_y = s_buf + CAP, *s_z = s_buf + 2 * CAP;
    float *s_v[3] = {s_buf + 3 * CAP, s_buf + 4 * CAP, s_buf + 5 * CAP};
    // half-cell counts (32-bit, for LDS atomics) borrow the particle
    // arrays, which are only written once the layout is known
    uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_buf);

    const int bxi = brick_id % nbx, byi = (brick_id / nbx) % nby, bzi = brick_id / (nbx * nby);
    const int i0 = bxi * BX, j0 = byi * BY, k0 = bzi * BZ;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const uint32_t row_stride = (uint32_t)bg.b[0];
    const int bx_hi = min(i0 + BX + 1, bg.b[0] - 1);

    // rows of staged cells: world cells [i0-1, i0+BX] along x are bins
    // [i0, i0+BX+1] (pad 1), contiguous in sort order; wave 0 scans the
    // lengths
    if (wave == 0) {
        uint32_t q0 = 0, len = 0;
        if (lane < NROW) {
            const int cy = j0 - 1 + lane % SY, cz = k0 - 1 + lane / SY;
            if (cy >= -1 && cy <= a.nj && cz >= -1 && cz <= a.nk) {
                const uint32_t row = row_stride * ((uint32_t)(cy + 1) + (uint32_t)bg.b[1] * (uint32_t)(cz + 1));
                q0 = start[row + i0];
                len = start[row + bx_hi + 1] - q0;
            }
            s_rowq[lane] = q0;
        }
        const uint32_t inc = wave_inclusive_sum(len);
        if (lane < NROW) s_rows0[lane + 1] = inc;
        if (lane == 0) s_rows0[0] = 0u;
    }
    for (int h = t; h <= NH; h += NT) s_cnt[h] = 0u;
    for (int h = t; h < SX * SY * SZ; h += NT) s_occ[h] = 0;
    __syncthreads();
    const uint32_t N = s_rows0[NROW];
    if (!CHUNKED && N > (uint32_t)CAP) {
        if (t == 0) dense[atomicAdd(dense_count, 1u)] = (uint32_t)brick_id;
        return;
    }

    // this thread's face: cell (li,lj,lk) of the brick, component = wave
    const int li = lane % BX, lj = (lane / BX) % BY, lk = lane / (BX * BY);
    const bool own_ok = i0 + li < a.ni + (wave == 0) && j0 + lj < a.nj + (wave == 1) &&
                        k0 + lk < a.nk + (wave == 2);
    // the faces on the domain's upper planes belong to the last brick:
    // lanes 0..15 of the wave take its 4x4 plane
    const int e0 = lane & 3, e1 = (lane >> 2) & 3;
    const FaceOut fo = face_out(a, wave);
    const bool xtra = (wave == 0 && i0 + BX == a.ni) || (wave == 1 && j0 + BY == a.nj) ||
                      (wave == 2 && k0 + BZ == a.nk);
    int eli, elj, elk;
    if (wave == 0) { eli = BX; elj = e0; elk = e1; }
    else if (wave == 1) { eli = e0; elj = BY; elk = e1; }
    else { eli = e0; elj = e1; elk = BZ; }
    const bool ext_ok = xtra && lane < 16 && i0 + eli < a.ni + (wave == 0) &&
                        j0 + elj < a.nj + (wave == 1) && k0 + elk < a.nk + (wave == 2);

    if (N == 0) {
        // nothing within reach of any face of this brick: all values are 0
        // and no face borders fluid
        if (own_ok) brick_zero_face(a, fo, wave, i0 + li, j0 + lj, k0 + lk);
        if (ext_ok) brick_zero_face(a, fo, wave, i0 + eli, j0 + elj, k0 + elk);
        return;
    }

    float S = 0.0f, W = 0.0f, SE = 0.0f, WE = 0.0f;
    for (uint32_t c0 = 0; c0 < N; c0 += (uint32_t)CAP) {
        const uint32_t ce = min(N, c0 + (uint32_t)CAP);
        if (c0 > 0) {
            __syncthreads();   // the last chunk's gather is done with LDS
            for (int h = t; h <= NH; h += NT) s_cnt[h] = 0u;
            __syncthreads();
        }

        // stage 1: positions -> half-cell of every particle of the chunk;
        // count them.  Every slot's loads (position, velocity, rank inside
        // its half-cell) are issued before any is used (slots past the
        // chunk load the brick's first slot and are masked), so the global
        // latency is paid once per batch.  The particle arrays are in sort
        // order: slot q of the sort is index q.
        uint32_t rq[PER], rh[PER], rk[PER];
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const uint32_t s = c0 + (uint32_t)t + (uint32_t)m * NT;
            const int r = brick_row_of(s_rows0, s);
            rq[m] = s < ce ? s_rowq[r] + (s - s_rows0[r]) : s_rowq[0];
        }
        float rx[PER], ry[PER], rz[PER], rvx[PER], rvy[PER], rvz[PER];
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const uint32_t q = rq[m];
            rx[m] = gld(a.x, q);
            ry[m] = gld(a.y, q);
            rz[m] = gld(a.z, q);
            rvx[m] = gld(a.vel[0], q);
            rvy[m] = gld(a.vel[1], q);
            rvz[m] = gld(a.vel[2], q);
            rk[m] = gld(hrank, q);
        }
#pragma unroll
        for (int m = 0; m < PER; m++) {
            // the same FP64 floor as the binning, so the cell agrees
            const double gxd = (double)rx[m] * bg.inv_dx;
            const double gyd = (double)ry[m] * bg.inv_dx;
            const double gzd = (double)rz[m] * bg.inv_dx;
            const double fx = floor(gxd), fy = floor(gyd), fz = floor(gzd);
            const int lx = (int)fx - (i0 - 1), ly = (int)fy - (j0 - 1), lz = (int)fz - (k0 - 1);
            const int ox = (gxd - fx) >= 0.5, oy = (gyd - fy) >= 0.5, oz = (gzd - fz) >= 0.5;
            rh[m] = (uint32_t)((2 * lx + ox) + HX * ((2 * ly + oy) + HY * (2 * lz + oz)));
            const uint32_t s = c0 + (uint32_t)t + (uint32_t)m * NT;
            if (s < ce) {
                atomicAdd(&s_cnt[rh[m]], 1u);
                s_occ[lx + SX * (ly + SY * lz)] = 1;
            }
        }
        __syncthreads();

        // stage 2: exclusive prefix over the half-cells
        {
            constexpr int ITEMS = (NH + NT - 1) / NT;
            uint32_t loc[ITEMS];
            uint32_t sum = 0;
#pragma unroll
            for (int m = 0; m < ITEMS; m++) {
                const int h = t * ITEMS + m;
                loc[m] = h < NH ? s_cnt[h] : 0u;
                sum += loc[m];
            }
            // block-wide exclusive sum over the 3 waves
            const uint32_t inc = wave_inclusive_sum(sum);
            if (lane == 63) s_wsum[wave] = inc;
            __syncthreads();
            uint32_t run = inc - sum;
            for (int w = 0; w < wave; w++) run += s_wsum[w];
#pragma unroll
            for (int m = 0; m < ITEMS; m++) {
                const int h = t * ITEMS + m;
                if (h < NH) s_hs[h] = (uint16_t)run;
                run += loc[m];
            }
            if (t == 0) s_hs[NH] = (uint16_t)(ce - c0);
        }
        __syncthreads();

        // stage 3: place every particle at its half-cell's first LDS slot
        // plus its rank among the half-cell's particles in this chunk (a
        // half-cell's particles are consecutive sort slots, the run
        // starting rank slots before the particle)
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const uint32_t s = c0 + (uint32_t)t + (uint32_t)m * NT;
            if (s < ce) {
                const uint32_t d = (uint32_t)s_hs[rh[m]] + min(rk[m], s - c0);
                s_x[d] = rx[m];
                s_y[d] = ry[m];
                s_z[d] = rz[m];
                s_v[0][d] = rvx[m];
                s_v[1][d] = rvy[m];
                s_v[2][d] = rvz[m];
            }
        }
        __syncthreads();

        // gather: wave w accumulates component w for its 64 faces, then the
        // domain-boundary plane if this brick has one.  (In the chunk loop
        // the face coordinates are hidden from the optimiser, so that it does
        // not keep every range address of both faces live across chunks.)
        int fli = li, flj = lj, flk = lk, feli = eli, felj = elj, felk = elk;
        if (CHUNKED) {
            asm volatile("" : "+v"(fli), "+v"(flj), "+v"(flk));
            asm volatile("" : "+v"(feli), "+v"(felj), "+v"(felk));
        }
        if (wave == 0) {
            if (own_ok) brick_accum<0, V>(a, s_x, s_y, s_z, s_v[0], s_hs, i0, j0, k0, fli, flj, flk, dx, inv_r2, S, W);
            if (ext_ok) brick_accum<0, V>(a, s_x, s_y, s_z, s_v[0], s_hs, i0, j0, k0, feli, felj, felk, dx, inv_r2, SE, WE);
        } else if (wave == 1) {
            if (own_ok) brick_accum<1, V>(a, s_x, s_y, s_z, s_v[1], s_hs, i0, j0, k0, fli, flj, flk, dx, inv_r2, S, W);
            if (ext_ok) brick_accum<1, V>(a, s_x, s_y, s_z, s_v[1], s_hs, i0, j0, k0, feli, felj, felk, dx, inv_r2, SE, WE);
        } else {
            if (own_ok) brick_accum<2, V>(a, s_x, s_y, s_z, s_v[2], s_hs, i0, j0, k0, fli, flj, flk, dx, inv_r2, S, W);
            if (ext_ok) brick_accum<2, V>(a, s_x, s_y, s_z, s_v[2], s_hs, i0, j0, k0, feli, felj, felk, dx, inv_r2, SE, WE);
        }
        if (!CHUNKED) break;
    }
    if (own_ok) brick_epilogue(fo, s_occ, wave, i0 + li, j0 + lj, k0 + lk, i0, j0, k0, S, W);
    if (ext_ok) brick_epilogue(fo, s_occ, wave, i0 + eli, j0 + elj, k0 + elk, i0, j0, k0, SE, WE);
}

// The common P2G kernel: one work-group of four waves per brick.  A brick
// whose staged cells hold no particle (flag 0, set by the sort) writes zeros
// and leaves at once.  Otherwise the staged cells (the brick and one cell
// around it) form 36 rows along x, each one contiguous range of the sort;
// wave w stages rows w, w+4, ..., 64 consecutive sort slots per lane step
// (coalesced loads).  A particle's cell in its row comes from the row's bin
// starts, its octant from the sort, and its LDS slot is its half-cell's run
// start (prefix of the half-cell counts) plus its rank in the run (from the
// sort).  Waves 0..2 then gather the faces of components 0..2, wave 3 the
// faces on the domain's upper planes when the brick has them.  A brick
// holding more than CAP particles is appended to dense[] for
// k_flip_p2g_dense.
// Row bin starts of a P2G brick: row (cy, cz) holds world cells i0-1 ..
// i0+4 at (j0-1+cy, k0-1+cz), bins i0 .. i0+5 (pad 1); thread t < 36*7
// loads start t of row t/7.  A row outside the bin grid, or cells beyond
// its end, are empty.
__device__ __forceinline__ uint32_t brick_row_start(const StepP2G &a, const BinGrid &bg,
                                                    int nbx, int nby, int brick_id,
                                                    const uint32_t *__restrict__ start, int t) {
    using namespace brick;
    constexpr int NB = SX + 1;
    const int bxi = brick_id % nbx, byi = (brick_id / nbx) % nby, bzi = brick_id / (nbx * nby);
    const int i0 = bxi * BX, j0 = byi * BY, k0 = bzi * BZ;
    uint32_t v = 0;
    if (t < NROW * NB) {
        const int r = t / NB, k = t % NB;
        const int wy = j0 - 1 + r % SY, wz = k0 - 1 + r / SY;
        if (wy <= a.nj && wz <= a.nk) {
            const uint32_t row = (uint32_t)bg.b[0] * ((uint32_t)(wy + 1) + (uint32_t)bg.b[1] * (uint32_t)(wz + 1));
            v = gld(start, row + (uint32_t)min(i0 + k, bg.b[0]));
        }
    }
    return v;
}

template <int V, int CAP>
__device__ __forceinline__ void p2g_brick_rows(
    const StepP2G &a, const BinGrid &bg, double dx, float inv_r2, int nbx, int nby,
    int brick_id, uint32_t gid, uint32_t v,
    const uint32_t *__restrict__ hrank, const uint8_t *__restrict__ soct, bool flagged,
    uint32_t *__restrict__ dense, uint32_t *__restrict__ dense_count) {
    using namespace brick;
    constexpr int NCS = SX * SY * SZ;          // staged cells
    constexpr int NB = SX + 1;                 // bin starts per row
    constexpr int RPW = (NROW + 3) / 4;        // rows per wave
    static_assert(NROW * NB <= NTC, "one thread per row bin start");
    static_assert(NH + 1 <= 6 * CAP, "half-cell counts live in the particle arrays");
    static_assert(CAP <= 65535, "LDS slots are 16-bit");
    __shared__ float s_buf[6 * CAP + 4];   // +4: masked reads past a range end
    __shared__ uint16_t s_hs[NH + 1];
    __shared__ uint32_t s_bs[NROW][NB];
    __shared__ uint32_t s_wsum[NTC / 64];
    __shared__ uint8_t s_occ[NCS];
    float *const s_x = s_buf, *const s_y = s_buf + CAP, *const s_z = s_buf + 2 * CAP;
    float *const s_v0 = s_buf + 3 * CAP, *const s_v1 = s_buf + 4 * CAP;
    float *const s_v2 = s_buf + 5 * CAP;
    // half-cell counts borrow the particle arrays, which are only written
    // once the layout is known
    uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_buf);

    const int bxi = brick_id % nbx, byi = (brick_id / nbx) % nby, bzi = brick_id / (nbx * nby);
    const int i0 = bxi * BX, j0 = byi * BY, k0 = bzi * BZ;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;

    // faces of this thread: waves 0..2 own the brick's faces of component
    // `wave`; wave 3's lanes 16c..16c+15 the upper-plane faces of component c
    const int c = wave < 3 ? wave : (lane >> 4);
    int fi, fj, fk;
    bool ok;
    if (wave < 3) {
        fi = lane % BX;
        fj = (lane / BX) % BY;
        fk = lane / (BX * BY);
        ok = true;
    } else {
        const int e0 = lane & 3, e1 = (lane >> 2) & 3;
        if (c == 0) { fi = BX; fj = e0; fk = e1; ok = i0 + BX == a.ni; }
        else if (c == 1) { fi = e0; fj = BY; fk = e1; ok = j0 + BY == a.nj; }
        else { fi = e0; fj = e1; fk = BZ; ok = k0 + BZ == a.nk; }
        ok = ok && lane < 48;
    }
    ok = ok && i0 + fi < a.ni + (c == 0) && j0 + fj < a.nj + (c == 1) && k0 + fk < a.nk + (c == 2);
    const FaceOut fo = face_out(a, c);

    // v: this thread's row bin start (brick_row_start), loaded by the
    // caller ahead of time together with the brick's flag
    if (!flagged) {
        // no particle within reach of any face of this brick: all values
        // are 0 and no face borders fluid
        if (ok) brick_zero_face(a, fo, c, i0 + fi, j0 + fj, k0 + fk);
        return;
    }
    if (t < NROW * NB) s_bs[t / NB][t % NB] = v;
    for (int h = t; h <= NH; h += NTC) s_cnt[h] = 0u;
    __syncthreads();
    uint32_t total = 0;
#pragma unroll 4
    for (int r = 0; r < NROW; r++) total += s_bs[r][NB - 1] - s_bs[r][0];
    if (total > (uint32_t)CAP) {
        if (t == 0) dense[atomicAdd(dense_count, 1u)] = gid;
        return;
    }
    if (total == 0) {
        if (ok) brick_zero_face(a, fo, c, i0 + fi, j0 + fj, k0 + fk);
        return;
    }
    if (t < NCS) {
        const int r = t / SX, k = t % SX;
        s_occ[t] = s_bs[r][k + 1] > s_bs[r][k];
    }

    // the half-cell of sort slot q in row r (staged row coordinates cy, cz)
    auto half_cell = [&](int r, uint32_t q, uint32_t o) {
        int ix = 0;
#pragma unroll
        for (int k = 1; k < SX; k++) ix += q >= s_bs[r][k];
        const int cy = r % SY, cz = r / SY;
        return (uint32_t)((2 * ix + (int)(o & 1u)) +
                          HX * ((2 * cy + (int)((o >> 1) & 1u)) + HY * (2 * cz + (int)(o >> 2))));
    };

    // count: each wave its rows; the first 64 particles of a row stay in
    // registers for placing, the rest are read again then.  Every row's
    // loads are issued before any is used (octants and ranks first, which
    // the counting needs first), so the wave waits for global memory once
    // rather than once per row.
    uint32_t rq[RPW], rlen[RPW];
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        const int r = wave + 4 * m;
        uint32_t base = 0, len = 0;
        if (r < NROW) {
            base = s_bs[r][0];
            len = s_bs[r][NB - 1] - base;
        }
        rlen[m] = len;
        rq[m] = (uint32_t)lane < len ? base + (uint32_t)lane : 0u;
    }
    uint32_t ro[RPW], rk[RPW];
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        ro[m] = gld(soct, rq[m]);
        rk[m] = gld(hrank, rq[m]);
    }
    float rx[RPW], ry[RPW], rz[RPW], rvx[RPW], rvy[RPW], rvz[RPW];
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        rx[m] = gld(a.x, rq[m]);
        ry[m] = gld(a.y, rq[m]);
        rz[m] = gld(a.z, rq[m]);
        rvx[m] = gld(a.vel[0], rq[m]);
        rvy[m] = gld(a.vel[1], rq[m]);
        rvz[m] = gld(a.vel[2], rq[m]);
    }
    uint32_t rd[RPW];   // half-cell | rank << 16
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        rd[m] = 0xFFFFFFFFu;
        if ((uint32_t)lane < rlen[m]) {
            const uint32_t h = half_cell(wave + 4 * m, rq[m], ro[m]);
            atomicAdd(&s_cnt[h], 1u);
            rd[m] = h | (rk[m] << 16);
        }
    }
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        const int r = wave + 4 * m;
        const uint32_t base = r < NROW ? s_bs[r][0] : 0u;
        for (uint32_t off = 64; off < rlen[m]; off += 64) {
            if (off + (uint32_t)lane < rlen[m]) {
                const uint32_t q2 = base + off + (uint32_t)lane;
                atomicAdd(&s_cnt[half_cell(r, q2, gld(soct, q2))], 1u);
            }
        }
    }
    __syncthreads();

    // exclusive prefix over the half-cells
    {
        constexpr int ITEMS = (NH + NTC - 1) / NTC;
        uint32_t loc[ITEMS];
        uint32_t sum = 0;
#pragma unroll
        for (int m = 0; m < ITEMS; m++) {
            const int h = t * ITEMS + m;
            loc[m] = h < NH ? s_cnt[h] : 0u;
            sum += loc[m];
        }
        const uint32_t inc = wave_inclusive_sum(sum);
        if (lane == 63) s_wsum[wave] = inc;
        __syncthreads();
        uint32_t run = inc - sum;
#pragma unroll
        for (int w = 0; w < NTC / 64; w++) run += w < wave ? s_wsum[w] : 0u;
#pragma unroll
        for (int m = 0; m < ITEMS; m++) {
            const int h = t * ITEMS + m;
            if (h < NH) s_hs[h] = (uint16_t)run;
            run += loc[m];
        }
        if (t == 0) s_hs[NH] = (uint16_t)total;
    }
    __syncthreads();

    // place: half-cell run start + rank in the run
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        if (rd[m] != 0xFFFFFFFFu) {
            const uint32_t d = (uint32_t)s_hs[rd[m] & 0xFFFFu] + (rd[m] >> 16);
            s_x[d] = rx[m];
            s_y[d] = ry[m];
            s_z[d] = rz[m];
            s_v0[d] = rvx[m];
            s_v1[d] = rvy[m];
            s_v2[d] = rvz[m];
        }
        const int r = wave + 4 * m;
        if (r < NROW) {
            const uint32_t base = s_bs[r][0], len = s_bs[r][NB - 1] - base;
            for (uint32_t off = 64; off < len; off += 64) {
                if (off + (uint32_t)lane < len) {
                    const uint32_t q2 = base + off + (uint32_t)lane;
                    const float px = gld(a.x, q2), py = gld(a.y, q2), pz = gld(a.z, q2);
                    const float pvx = gld(a.vel[0], q2), pvy = gld(a.vel[1], q2);
                    const float pvz = gld(a.vel[2], q2);
                    const uint32_t d = (uint32_t)s_hs[half_cell(r, q2, gld(soct, q2))] +
                                       gld(hrank, q2);
                    s_x[d] = px;
                    s_y[d] = py;
                    s_z[d] = pz;
                    s_v0[d] = pvx;
                    s_v1[d] = pvy;
                    s_v2[d] = pvz;
                }
            }
        }
    }
    __syncthreads();

    float S = 0.0f, W = 0.0f;
    if (ok) {
        if (c == 0) brick_accum<0, V>(a, s_x, s_y, s_z, s_v0, s_hs, i0, j0, k0, fi, fj, fk, dx, inv_r2, S, W);
        else if (c == 1) brick_accum<1, V>(a, s_x, s_y, s_z, s_v1, s_hs, i0, j0, k0, fi, fj, fk, dx, inv_r2, S, W);
        else brick_accum<2, V>(a, s_x, s_y, s_z, s_v2, s_hs, i0, j0, k0, fi, fj, fk, dx, inv_r2, S, W);
        brick_epilogue(fo, s_occ, c, i0 + fi, j0 + fj, k0 + fk, i0, j0, k0, S, W);
    }
}

// One work-group per P2G brick of any scene (global brick numbering).
template <int V, int CAP>
__global__ __launch_bounds__(brick::NTC) void k_flip_p2g_brick(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, uint32_t nbrick, int per_xcd,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ hrank,
    const uint8_t *__restrict__ soct, const uint8_t *__restrict__ brick_flag,
    uint32_t *__restrict__ dense, uint32_t *__restrict__ dense_count) {
    // consecutive bricks share halo particles: keep them on one XCD
    // (work-groups are dealt round-robin over the 8 XCDs)
    const uint32_t gid = (blockIdx.x & 7u) * (uint32_t)per_xcd + (blockIdx.x >> 3);
    if (gid >= nbrick) return;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &d) { return d.k0; })];
    const bool flagged = gld(brick_flag, gid) != 0;
    const uint32_t v = brick_row_start(sd.p2g, sd.bg, sd.nbx, sd.nby, (int)(gid - sd.k0),
                                       start + sd.b0, threadIdx.x);
    p2g_brick_rows<V, CAP>(sd.p2g, sd.bg, sd.dx, (float)(1.0 / (sd.dx * sd.dx)), sd.nbx, sd.nby,
                           (int)(gid - sd.k0), gid, v, hrank, soct, flagged, dense,
                           dense_count);
}

// The bricks listed by k_flip_p2g_brick as over its capacity, in chunks.
// Three waves per SIMD (a register budget of 168).  A work-group takes the
// next unprocessed brick of the list from a counter (dense_count[1]) each
// time it finishes one, so large and small bricks balance over the
// work-groups; the order does not change any result.
template <int V, int CAP>
__global__ __launch_bounds__(brick::NT, 3) void k_flip_p2g_dense(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ hrank,
    uint32_t *__restrict__ dense, uint32_t *__restrict__ dense_count) {
    using namespace brick;
    static_assert(NH + 1 <= 6 * CAP, "half-cell counts live in the particle arrays");
    static_assert(CAP <= 65535, "LDS slots are 16-bit");
    __shared__ float s_buf[6 * CAP + 4];   // +4: masked reads past a range end
    __shared__ uint16_t s_hs[NH + 1];
    __shared__ uint32_t s_rowq[NROW], s_rows0[NROW + 1];
    __shared__ uint32_t s_wsum[NT / 64];
    __shared__ uint8_t s_occ[SX * SY * SZ];
    __shared__ uint32_t s_next;
    const uint32_t nd = dense_count[0];
    while (true) {
        __syncthreads();   // the previous brick is done with LDS and s_next
        if (threadIdx.x == 0) s_next = atomicAdd(&dense_count[1], 1u);
        __syncthreads();
        const uint32_t d = s_next;
        if (d >= nd) break;
        const uint32_t gid = dense[d];
        const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &e) { return e.k0; })];
        p2g_brick<V, CAP, true>(sd.p2g, sd.bg, sd.dx, (float)(1.0 / (sd.dx * sd.dx)), sd.nbx, sd.nby,
                                (int)(gid - sd.k0), nullptr, start + sd.b0, hrank, dense,
                                dense_count, s_buf, s_hs, s_rowq, s_rows0, s_wsum, s_occ);
    }
}

// ---------------------------------------------------------------------------


// Positions and velocities copied into sorted order, so that the transfer
// kernels read them by sort slot with no indirection.  Slot q belongs to the
// scene whose particle range holds q (the sort keeps scenes in order).
__global__ __launch_bounds__(kBlock) void k_sorted_copy(const SceneDesc *__restrict__ scenes,
                                                       int nsc, const uint32_t *__restrict__ perm,
                                                       int64_t n, float *xs0, float *xs1,
                                                       float *xs2, float *vs0, float *vs1,
                                                       float *vs2) {
    const int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (q >= n) return;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, (uint64_t)q, [](const SceneDesc &d) { return d.p0; })];
    const uint32_t p = perm[q] - (uint32_t)sd.p0;
    const float x = sd.xin[0][p], y = sd.xin[1][p], z = sd.xin[2][p];
    const float vx = sd.vin[0][p], vy = sd.vin[1][p], vz = sd.vin[2][p];
    xs0[q] = x;
    xs1[q] = y;
    xs2[q] = z;
    vs0[q] = vx;
    vs1[q] = vy;
    vs2[q] = vz;
}

// One thread per particle, in cell order: PIC/FLIP blend
// (fluidsimulation.cpp:3115-3127) and RK4 (particleadvector.cpp:1045-1054)
// through the new field.
template <int ST>
__global__ __launch_bounds__(kBlock) void k_flip_g2p(
    StepG2P a, MacGrid g, const uint32_t *__restrict__ perm) {
    const int64_t jdx = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (jdx >= a.n) return;
    const uint32_t q = (uint32_t)jdx;
    const float x0[3] = {a.xin[0][q], a.xin[1][q], a.xin[2][q]};
    const float v0[3] = {a.vin[0][q], a.vin[1][q], a.vin[2][q]};
    float k1[3], vo[3];
    sample_mac<ST, true>(g, a.old[0], a.old[1], a.old[2], x0[0], x0[1], x0[2], k1, vo);
    float vn[3] = {k1[0], k1[1], k1[2]};
    validate3(vn);
    validate3(vo);
    float vnew[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float vflip = (v0[c] + vn[c]) - vo[c];
        vnew[c] = a.r_pic * vn[c] + a.r_flip * vflip;
    }
    float x1[3];
    rk_step<ST>(g, 4, a.rk, x0, k1, x1);
    store_particle(a, perm, q, x1, vnew);
}

// ---------------------------------------------------------------------------
// Brick-staged G2P: one work-group per brick of GB cells per axis.  The
// work-group copies the box of the new field around its brick (the stencil
// footprint of every position inside the brick plus a halo of GH cells for
// the RK stage positions) and the box of the pre-force snapshot (no halo:
// it is only sampled at the start position) into LDS, with 0 wherever the
// box leaves the arrays -- the value the reference reads there
// (macvelocityfield.cpp:372-378,470-478).  A sample whose stencil leaves
// the box reads global memory instead, so every position is handled and
// the result is bit-identical to k_flip_g2p.
// ---------------------------------------------------------------------------

// The brick body of k_flip_g2p_brick for brick brick_id of one scene
// (start: the scene's bin starts, local bin indices).
template <int ST>
__device__ __forceinline__ void g2p_brick(const StepG2P &a, const MacGrid &g, const BinGrid &bg,
                                          int nbx, int nby, int nbz, int brick_id,
                                          const uint32_t *__restrict__ start,
                                          const uint32_t *__restrict__ perm) {
    using namespace gbrick;
    using B = Box<ST>;
    __shared__ float s_cur[3][B::E * B::E * B::E];
    __shared__ float s_old[3][B::EO * B::EO * B::EO];
    __shared__ uint32_t s_rowq[MAXROW], s_rows0[MAXROW + 1];
    __shared__ uint32_t s_wsum;
    const int t = threadIdx.x;
    const int bxi = brick_id % nbx, byi = (brick_id / nbx) % nby, bzi = brick_id / (nbx * nby);
    const int c0x = bxi * GB, c0y = byi * GB, c0z = bzi * GB;
    // bins of the brick (bin = cell + 1); the first and last bricks of an
    // axis also take the pad bin beyond the domain
    const int bx_lo = bxi == 0 ? 0 : c0x + 1, bx_hi = bxi == nbx - 1 ? bg.b[0] : c0x + GB + 1;
    const int by_lo = byi == 0 ? 0 : c0y + 1, by_hi = byi == nby - 1 ? bg.b[1] : c0y + GB + 1;
    const int bz_lo = bzi == 0 ? 0 : c0z + 1, bz_hi = bzi == nbz - 1 ? bg.b[2] : c0z + GB + 1;
    const int ny = by_hi - by_lo;
    const int nrow = ny * (bz_hi - bz_lo);

    // (y,z) rows of bins: each is one contiguous range of perm
    if (t < 128) {
        uint32_t q0 = 0, len = 0;
        if (t < nrow) {
            const uint32_t by = (uint32_t)(by_lo + t % ny), bz = (uint32_t)(bz_lo + t / ny);
            const uint32_t row = (uint32_t)bg.b[0] * (by + (uint32_t)bg.b[1] * bz);
            q0 = start[row + (uint32_t)bx_lo];
            len = start[row + (uint32_t)bx_hi] - q0;
            s_rowq[t] = q0;
        }
        const uint32_t inc = wave_inclusive_sum(len);
        if (t == 63) s_wsum = inc;
        if (t < nrow) s_rows0[t + 1] = inc;   // the second wave adds the first's total below
        if (t == 0) s_rows0[0] = 0u;
    }
    __syncthreads();
    if (t >= 64 && t < nrow) s_rows0[t + 1] += s_wsum;
    __syncthreads();
    const uint32_t N = s_rows0[nrow];
    if (N == 0) return;

    const int lx = c0x - B::LO, ly = c0y - B::LO, lz = c0z - B::LO;
    const int lox = c0x - B::LOO, loy = c0y - B::LOO, loz = c0z - B::LOO;
    stage_box<B::E>(s_cur[0], g.f[0], g.ni + 1, g.nj, g.nk, lx, ly, lz);
    stage_box<B::E>(s_cur[1], g.f[1], g.ni, g.nj + 1, g.nk, lx, ly, lz);
    stage_box<B::E>(s_cur[2], g.f[2], g.ni, g.nj, g.nk + 1, lx, ly, lz);
    stage_box<B::EO>(s_old[0], a.old[0], g.ni + 1, g.nj, g.nk, lox, loy, loz);
    stage_box<B::EO>(s_old[1], a.old[1], g.ni, g.nj + 1, g.nk, lox, loy, loz);
    stage_box<B::EO>(s_old[2], a.old[2], g.ni, g.nj, g.nk + 1, lox, loy, loz);
    __syncthreads();

    // slot s -> sort index: the row is the last r with rows0[r] <= s
    auto slot_q = [&](uint32_t s) {
        int lo = 0, hi = nrow - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_rows0[mid] <= s) lo = mid;
            else hi = mid - 1;
        }
        return s_rowq[lo] + (s - s_rows0[lo]);
    };
    // the next particle's position and velocity are loaded while this one
    // is computed
    uint32_t qn = 0;
    float xn[3] = {0.0f, 0.0f, 0.0f}, vn[3] = {0.0f, 0.0f, 0.0f};
    if ((uint32_t)t < N) {
        qn = slot_q((uint32_t)t);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            xn[c] = gld(a.xin[c], qn);
            vn[c] = gld(a.vin[c], qn);
        }
    }
    for (uint32_t s = (uint32_t)t; s < N; s += NT) {
        const uint32_t q = qn;
        const float x0[3] = {xn[0], xn[1], xn[2]};
        const float v0[3] = {vn[0], vn[1], vn[2]};
        if (s + NT < N) {
            qn = slot_q(s + NT);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                xn[c] = gld(a.xin[c], qn);
                vn[c] = gld(a.vin[c], qn);
            }
        }
        float k1[3], vo[3], vnew[3];
        brick_sample<ST, true>(g, a.old[0], a.old[1], a.old[2], s_cur[0], s_cur[1], s_cur[2],
                               s_old[0], s_old[1], s_old[2], lx, ly, lz, lox, loy, loz,
                               x0[0], x0[1], x0[2], k1, vo);
        flip_blend(a, v0, k1, vo, vnew);
        // RK4 (particleadvector.cpp:1045-1054) through the staged field
        float xq[3], k2[3], k3[3], k4[3], unused[3], x1[3];
#pragma unroll
        for (int c = 0; c < 3; c++) xq[c] = x0[c] + a.rk.half * k1[c];
        brick_sample<ST, false>(g, nullptr, nullptr, nullptr, s_cur[0], s_cur[1], s_cur[2],
                                nullptr, nullptr, nullptr, lx, ly, lz, 0, 0, 0,
                                xq[0], xq[1], xq[2], k2, unused);
#pragma unroll
        for (int c = 0; c < 3; c++) xq[c] = x0[c] + a.rk.half * k2[c];
        brick_sample<ST, false>(g, nullptr, nullptr, nullptr, s_cur[0], s_cur[1], s_cur[2],
                                nullptr, nullptr, nullptr, lx, ly, lz, 0, 0, 0,
                                xq[0], xq[1], xq[2], k3, unused);
#pragma unroll
        for (int c = 0; c < 3; c++) xq[c] = x0[c] + a.rk.full * k3[c];
        brick_sample<ST, false>(g, nullptr, nullptr, nullptr, s_cur[0], s_cur[1], s_cur[2],
                                nullptr, nullptr, nullptr, lx, ly, lz, 0, 0, 0,
                                xq[0], xq[1], xq[2], k4, unused);
#pragma unroll
        for (int c = 0; c < 3; c++)
            x1[c] = x0[c] + a.rk.sixth * (((k1[c] + 2.0f * k2[c]) + 2.0f * k3[c]) + k4[c]);
        store_particle(a, perm, q, x1, vnew);
    }
}

// One work-group per G2P brick of any scene (global brick numbering); the
// last NEXTRA work-groups take the particles beyond each scene's padded bin
// grid: outside the domain, so every sample is 0 and the global path
// returns at once.
template <int ST>
__device__ __forceinline__ void g2p_brick_kernel(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order,
    uint32_t nbrick, int per_xcd, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ perm) {
    using namespace gbrick;
    const int t = threadIdx.x;
    const int nslot = 8 * per_xcd;
    if ((int)blockIdx.x >= nslot) {
        for (int sc = 0; sc < nsc; sc++) {
            const SceneDesc &sd = scenes[sc];
            const StepG2P a = g2p_view(sd, so, keep_order);
            const MacGrid g = mac_view(sd);
            const uint32_t ovf = sd.b0 + sd.bg.nbin;
            const uint32_t q0 = start[ovf], q1 = start[ovf + 1];
            for (uint32_t q = q0 + (uint32_t)((blockIdx.x - nslot) * NT + t); q < q1;
                 q += (uint32_t)(NEXTRA * NT)) {
                const float x0[3] = {a.xin[0][q], a.xin[1][q], a.xin[2][q]};
                const float v0[3] = {a.vin[0][q], a.vin[1][q], a.vin[2][q]};
                float k1[3], vo[3], vnew[3], x1[3];
                sample_mac<ST, true>(g, a.old[0], a.old[1], a.old[2], x0[0], x0[1], x0[2], k1, vo);
                flip_blend(a, v0, k1, vo, vnew);
                rk_step<ST>(g, 4, a.rk, x0, k1, x1);
                store_particle(a, perm, q, x1, vnew);
            }
        }
        return;
    }
    // consecutive bricks share box halos: keep them on one XCD (work-groups
    // are dealt round-robin over the 8 XCDs)
    const uint32_t gid = (blockIdx.x & 7u) * (uint32_t)per_xcd + (blockIdx.x >> 3);
    if (gid >= nbrick) return;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &d) { return d.g0; })];
    const StepG2P a = g2p_view(sd, so, keep_order);
    const MacGrid g = mac_view(sd);
    g2p_brick<ST>(a, g, sd.bg, sd.gbx, sd.gby, sd.gbz, (int)(gid - sd.g0), start + sd.b0, perm);
}

// The trilinear form; the tricubic one is limited to two waves per SIMD
// (its LDS boxes allow two work-groups per CU anyway), which keeps the
// compiler from spending its whole register file on one wave.
__global__ __launch_bounds__(gbrick::NT) void k_flip_g2p_brick_tl(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order,
    uint32_t nbrick, int per_xcd, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ perm) {
    g2p_brick_kernel<0>(scenes, nsc, so, keep_order, nbrick, per_xcd, start, perm);
}

__global__ __launch_bounds__(gbrick::NT, 2) void k_flip_g2p_brick_tc(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order,
    uint32_t nbrick, int per_xcd, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ perm) {
    g2p_brick_kernel<1>(scenes, nsc, so, keep_order, nbrick, per_xcd, start, perm);
}

void layout_scenes(SceneDesc *sc, int nsc, SceneTotals &tot) {
    tot = SceneTotals();
    for (int i = 0; i < nsc; i++) {
        SceneDesc &d = sc[i];
        BinGrid &bg = d.bg;
        bg.b[0] = d.ni + 2;
        bg.b[1] = d.nj + 2;
        bg.b[2] = d.nk + 2;
        bg.pad[0] = bg.pad[1] = bg.pad[2] = 1;
        bg.off[0] = bg.off[1] = bg.off[2] = 0.0f;
        bg.inv_dx = 1.0 / d.dx;
        bg.nbin = (uint32_t)bg.b[0] * (uint32_t)bg.b[1] * (uint32_t)bg.b[2];
        d.p0 = tot.n;
        d.b0 = (uint32_t)tot.nbin;
        d.blk0 = (uint32_t)tot.nblk;
        d.nbx = (d.ni + brick::BX - 1) / brick::BX;
        d.nby = (d.nj + brick::BY - 1) / brick::BY;
        d.nbz = (d.nk + brick::BZ - 1) / brick::BZ;
        d.k0 = (uint32_t)tot.np2g;
        d.gbx = (d.ni + gbrick::GB - 1) / gbrick::GB;
        d.gby = (d.nj + gbrick::GB - 1) / gbrick::GB;
        d.gbz = (d.nk + gbrick::GB - 1) / gbrick::GB;
        d.g0 = (uint32_t)tot.ng2p;
        d.nbb = 8u * (uint32_t)(d.gbx * d.gby * d.gbz);
        d.kb0 = (uint32_t)tot.nbb;
        d.f0 = tot.nface;
        d.r0 = tot.nrow;
        d.rm0 = (uint32_t)tot.nrm;
        d.rmw = (d.nbx + 63) / 64;
        tot.nrm += (int64_t)d.rmw * d.nby * d.nbz;
        MacGrid g;
        g.ni = d.ni;
        g.nj = d.nj;
        g.nk = d.nk;
        g.dx = d.dx;
        set_limits(g);
        for (int c = 0; c < 3; c++) {
            d.lim[c] = g.lim[c];
            d.limf[c] = g.limf[c];
        }
        tot.n += d.n;
        tot.nbin += (int64_t)bg.nbin + 1;
        tot.nblk += sort_blocks(d.n);
        tot.np2g += (int64_t)d.nbx * d.nby * d.nbz;
        tot.ng2p += (int64_t)d.gbx * d.gby * d.gbz;
        tot.nbb += (int64_t)d.nbb + 1;
        // combine waves: each takes combine_rows(nx) consecutive face rows
        // along x
        int64_t waves[3];
        for (int c = 0; c < 3; c++) {
            const int64_t rows = (int64_t)(d.nj + (c == 1)) * (d.nk + (c == 2));
            const int64_t rpw = combine_rows(d.ni + (c == 0));
            waves[c] = (rows + rpw - 1) / rpw;
        }
        d.cw[0] = waves[0];
        d.cw[1] = waves[0] + waves[1];
        tot.nrow += waves[0] + waves[1] + waves[2];
        tot.nface += (int64_t)(d.ni + 1) * d.nj * d.nk + (int64_t)d.ni * (d.nj + 1) * d.nk +
                     (int64_t)d.ni * d.nj * (d.nk + 1);
    }
}

void sorted_copy(const SceneDesc *scenes_dev, int nsc, const uint32_t *perm, int64_t n,
                 float *const xs[3], float *const vs[3], hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sorted_copy, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, scenes_dev,
                       nsc, perm, n, xs[0], xs[1], xs[2], vs[0], vs[1], vs[2]);
    GFSX_LAUNCH_CHECK();
}

void flip_g2p_direct(const StepBatch &s, const uint32_t *perm, hipStream_t st) {
    // one scene only
    const SceneDesc &d = s.scenes[0];
    Sorted so;
    for (int c = 0; c < 3; c++) {
        so.x[c] = s.xs[c];
        so.v[c] = s.vs[c];
    }
    const StepG2P b = g2p_view(d, so, s.keep_order);
    const MacGrid g = mac_view(d);
    if (s.stencil == 0)
        hipLaunchKernelGGL(k_flip_g2p<0>, dim3(grid_for(d.n, kBlock)), dim3(kBlock), 0, st, b, g,
                           perm);
    else
        hipLaunchKernelGGL(k_flip_g2p<1>, dim3(grid_for(d.n, kBlock)), dim3(kBlock), 0, st, b, g,
                           perm);
    GFSX_LAUNCH_CHECK();
}

void flip_step(const StepBatch &s, const SortBuffers &sb, hipStream_t st, hipEvent_t *ev) {
    if (ev) GFSX_HIP(hipEventRecord(ev[0], st));
    const int nsc = (int)s.scenes.size();
    // cells ordered inside by (octant, index), so that each half-cell is a
    // run for the P2G gathers.  The brick flags are set by the sort for this
    // step's particles only.
    fill_u8(sb.brick_flag, 0, s.tot.np2g, st);
    cell_sort(sb, s.tot.n, st, true, true);
    Sorted so;
    for (int c = 0; c < 3; c++) {
        so.x[c] = s.xs[c];
        so.v[c] = s.vs[c];
    }
    sorted_copy(s.scenes_dev, nsc, (const uint32_t *)sb.perm, s.tot.n, s.xs, s.vs, st);
    if (ev) GFSX_HIP(hipEventRecord(ev[1], st));

    if (s.p2g_variant != 1) {
        const uint32_t nbrick = (uint32_t)s.tot.np2g;
        const int per_xcd = (int)((nbrick + 7) / 8);
        constexpr int cap = brick::CAP;
        auto *kern = s.p2g_variant == 2 ? k_flip_p2g_brick<2, cap>
                   : s.p2g_variant == 3 ? k_flip_p2g_brick<3, cap> : k_flip_p2g_brick<0, cap>;
        constexpr int dcap = brick::DENSE_CAP;
        auto *kern_dense = s.p2g_variant == 2 ? k_flip_p2g_dense<2, dcap>
                         : s.p2g_variant == 3 ? k_flip_p2g_dense<3, dcap>
                                              : k_flip_p2g_dense<0, dcap>;
        fill_u32(s.dense_count, 0u, 2, st);
        hipLaunchKernelGGL(kern, dim3((unsigned)(8 * (int64_t)per_xcd)), dim3(brick::NTC), 0, st,
                           s.scenes_dev, nsc, so, nbrick, per_xcd, (const uint32_t *)sb.start,
                           (const uint32_t *)sb.hrank, (const uint8_t *)sb.soct,
                           (const uint8_t *)sb.brick_flag, s.dense, s.dense_count);
        GFSX_LAUNCH_CHECK();
        // bricks holding more than cap particles, from the dense list
        const unsigned ndense = (unsigned)std::min<int64_t>(nbrick, brick::DENSE_BLOCKS);
        hipLaunchKernelGGL(kern_dense, dim3(ndense), dim3(brick::NT), 0, st, s.scenes_dev, nsc,
                           so, (const uint32_t *)sb.start, (const uint32_t *)sb.hrank, s.dense,
                           s.dense_count);
    } else {
        // the direct gather: one scene only
        const SceneDesc &d = s.scenes[0];
        const StepP2G &a = d.p2g;
        const int64_t F = a.nface[0] + a.nface[1] + a.nface[2];
        hipLaunchKernelGGL(k_flip_p2g, dim3(grid_for(F, kBlock)), dim3(kBlock), 0, st, a, d.bg,
                           d.dx, (float)(1.0 / (d.dx * d.dx)), (const uint32_t *)sb.count,
                           (const uint32_t *)sb.start, (const uint32_t *)nullptr);
    }
    GFSX_LAUNCH_CHECK();
    if (s.epi) fs_epilogue(*s.epi, st);
    if (ev) GFSX_HIP(hipEventRecord(ev[2], st));

    if (s.tot.n <= 0 || s.grids_only) {
        if (ev) GFSX_HIP(hipEventRecord(ev[3], st));
        return;
    }
    if (s.g2p_variant == 0) {
        const uint32_t nbrick = (uint32_t)s.tot.ng2p;
        const int per_xcd = (int)((nbrick + 7) / 8);
        const dim3 grid((unsigned)(8 * per_xcd + gbrick::NEXTRA)), block(gbrick::NT);
        auto *kern = s.stencil == 0 ? k_flip_g2p_brick_tl : k_flip_g2p_brick_tc;
        hipLaunchKernelGGL(kern, grid, block, 0, st, s.scenes_dev, nsc, so, s.keep_order, nbrick,
                           per_xcd, (const uint32_t *)sb.start, (const uint32_t *)sb.perm);
        GFSX_LAUNCH_CHECK();
    } else {
        flip_g2p_direct(s, (const uint32_t *)sb.perm, st);
    }
    if (s.epi && s.epi->solid) fs_collide(*s.epi, st);
    if (ev) GFSX_HIP(hipEventRecord(ev[3], st));
}

}  // namespace gfsx
