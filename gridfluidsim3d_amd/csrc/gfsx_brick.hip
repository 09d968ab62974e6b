// gfsx_brick.hip -- the default fused FLIP step (p2g_variant 4) on gfx950.
//
//  * brick sort : a stable counting sort of the particles by P2G brick
//                 (4^3 cells).  Every wave of 64 consecutive input particles
//                 splits into runs of equal bricks; a bin takes its runs in
//                 wave order and each run in lane order, which is exactly the
//                 order of a sequential stable counting sort.  Only the runs
//                 (a few per wave when the input is close to brick order, as
//                 the previous step leaves it) are ordered, never the
//                 particles; positions and velocities are then written
//                 straight to their slots.
//  * scatter P2G: one work-group per brick adds K and K v of its own
//                 particles to the 2x2x2 nodes of each component around them,
//                 in a tile of 64-bit fixed-point sums in LDS (integer sums:
//                 the result does not depend on the order of the additions).
//  * combine    : one work-group per brick adds, in a fixed order, the tile
//                 entries of the at most 2x2x2 bricks that reach each of its
//                 faces, then normalises, snapshots and adds the body force.
//  * G2P        : one work-group per 8^3-cell brick (eight P2G bricks, one
//                 contiguous range of the sort) stages the grid boxes in LDS
//                 and runs PIC/FLIP and RK4 for its particles.
//
// Reference behaviour followed: ScalarField::addPointValue / applyWeightField
// (scalarfield.cpp:90-106,167-201), the FluidSimulation stages
// (fluidsimulation.cpp:2010-2017,2526-2595,2766-2805,3104-3129,3181-3209),
// MACVelocityField::_interpolate{,Linear}U/V/W (macvelocityfield.cpp:350-575)
// and ParticleAdvector::_RK4 (particleadvector.cpp:1045-1054).

#include "gfsx_internal.h"
#include "gfsx_sample.h"
#include "gfsx_device.h"

#include <type_traits>

namespace gfsx {

namespace bsort {
constexpr int ITEMS = 4;                          // particles per thread
constexpr int64_t TILE = (int64_t)ITEMS * kBlock; // particles per block (as sort_blocks)
constexpr int WAVES = (int)(TILE / 64);           // waves per block tile
constexpr uint32_t SMALL = 16;                    // runs per bin ordered by one thread
constexpr uint32_t MID = 256;                     // ... by one wave
constexpr int HUGE_WORDS = 8192;                  // bitmap words per pass (k_bsort_order_huge)
constexpr int MID_BLOCKS = 2048, HUGE_BLOCKS = 256;
constexpr int HT_BITS = 8, HT = 1 << HT_BITS;     // bin table entries per work-group
constexpr int HT_PROBES = 16;                     // entries a run tries
constexpr uint32_t ALL_RUNS = 0xFFFFFFFFu;        // nchunk: every run a chunk of its own
}  // namespace bsort

constexpr int kSlowBlocks = 1024;   // work-groups of k_flip_g2p_slow

// ---------------------------------------------------------------------------
// Brick keys
// ---------------------------------------------------------------------------

// P2G brick along one axis of the cell floor(p/dx) (FP64, as the binning of
// the cell sort and Grid3d::positionToGridIndex, grid3d.h:58-63).  Cells
// -1 and n (the pad layers) belong to the first and last brick; anything
// further out (and NaN) clears ok.
__device__ __forceinline__ int brick_axis(float p, double inv_dx, int n, bool &ok) {
    const double c = floor((double)p * inv_dx);
    ok = ok && c >= -1.0 && c <= (double)n;
    return ok ? (min(max((int)c, 0), n - 1) >> 2) : 0;
}

// Local bin of brick (bx,by,bz): its G2P brick's index times 8 plus its
// octant inside that G2P brick, so the 8 bins of a G2P brick are
// consecutive.
__device__ __forceinline__ uint32_t brick_local(const SceneDesc &sd, int bx, int by, int bz) {
    const uint32_t g = (uint32_t)(bx >> 1) +
                       (uint32_t)sd.gbx * ((uint32_t)(by >> 1) + (uint32_t)sd.gby * (uint32_t)(bz >> 1));
    return (g << 3) | (uint32_t)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2));
}

__device__ __forceinline__ void brick_coords(const SceneDesc &sd, uint32_t lb, int &bx, int &by,
                                             int &bz) {
    const uint32_t g = lb >> 3, sub = lb & 7u;
    const uint32_t gbx = (uint32_t)sd.gbx, gby = (uint32_t)sd.gby;
    const uint32_t gx = g % gbx, gy = (g / gbx) % gby, gz = g / (gbx * gby);
    bx = (int)(2 * gx + (sub & 1u));
    by = (int)(2 * gy + ((sub >> 1) & 1u));
    bz = (int)(2 * gz + (sub >> 2));
}

// Local bin of a particle: its brick, or nbb (the scene's overflow bin).
__device__ __forceinline__ uint32_t brick_key(const SceneDesc &sd, float x, float y, float z) {
    bool ok = true;
    const double inv = sd.bg.inv_dx;
    const int bx = brick_axis(x, inv, sd.ni, ok);
    const int by = brick_axis(y, inv, sd.nj, ok);
    const int bz = brick_axis(z, inv, sd.nk, ok);
    return ok ? brick_local(sd, bx, by, bz) : sd.nbb;
}

// The runs of equal keys among the valid lanes of a wave (all its lanes
// with that key), in the order of each run's lowest lane: a lane receives
// its run's index j, its rank in the run (lane order) and the run's lowest
// lane; at_leader(is_leader, key, length, j) is called by every lane once
// per run.  Returns the number of runs.
template <class F>
__device__ __forceinline__ int wave_runs(uint32_t key, bool valid, int &j, int &rank,
                                         int &leader, F &&at_leader) {
    const int lane = threadIdx.x & 63;
    uint64_t rem = __ballot(valid);
    int r = 0;
    j = 0;
    rank = 0;
    leader = 0;
    while (rem) {
        const int lead = __ffsll((unsigned long long)rem) - 1;
        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, lead);
        const uint64_t m = __ballot(key == k) & rem;
        if ((m >> lane) & 1ull) {
            j = r;
            rank = __popcll(m & ((1ull << lane) - 1ull));
            leader = lead;
        }
        at_leader(lane == lead, k, (uint32_t)__popcll(m), r);
        rem &= ~m;
        r++;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Brick sort
// ---------------------------------------------------------------------------

// Count every bin's particles and chunks.  A chunk is what one work-group
// (TILE particles, 16 waves) holds of one bin: the input is close to brick
// order, so a work-group's particles fall in a few bins.  The runs of each
// wave (its particles of one bin) are gathered per bin in an LDS table: a
// 16-bit mask of the waves with a run of that bin and each run's length,
// so a run's offset in its chunk is the length of the bin's runs in the
// waves before it (wave order, then lane order: the order of a stable
// sort).  The work-group adds (1 << 32) | particles to each of its bins
// with one atomic; the old value's high word is the chunk's provisional
// rank among the bin's chunks (ordered later by chunk id).  Chunk records
// (bin, provisional rank) and lengths go to runrec / rlen at
// blk * TILE + c, c the chunk's index among the work-group's bins in table
// order; nchunk[blk] = their number.  A work-group whose runs do not all
// find a table entry takes each run as a chunk of its own, c = wave * 64 +
// run, and sets nchunk[blk] = ALL_RUNS.  Each particle's chunk and offset
// in it go to jr (c << 10 | offset) at its binning position.
__global__ __launch_bounds__(kBlock) void k_bsort_bin(const SceneDesc *__restrict__ scenes, int nsc,
                                                     unsigned long long *__restrict__ count,
                                                     uint2 *__restrict__ runrec,
                                                     uint32_t *__restrict__ rlen,
                                                     uint8_t *__restrict__ nruns,
                                                     uint32_t *__restrict__ nchunk,
                                                     uint32_t *__restrict__ jr) {
    using namespace bsort;
    static_assert(HT == kBlock && WAVES == 16, "one table entry per thread, 16-bit wave masks");
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    __shared__ uint32_t s_key[HT];       // bin, or EMPTY
    __shared__ uint32_t s_cnt[HT];       // particles
    __shared__ uint32_t s_mask[HT];      // waves with a run of the bin
    __shared__ uint16_t s_len[HT][WAVES];   // those runs' lengths
    __shared__ uint32_t s_cc[HT];        // chunk index of a used entry
    __shared__ uint32_t s_wsum[kBlock / 64];
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const uint32_t blk = blockIdx.x;
    s_key[t] = EMPTY;
    s_cnt[t] = 0u;
    s_mask[t] = 0u;
    if (t == 0) s_fail = 0;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, blk, [](const SceneDesc &d) { return d.blk0; })];
    const int64_t n = sd.n;
    const uint32_t kb0 = sd.kb0;
    const int64_t base = (int64_t)(blk - sd.blk0) * TILE + t;
    float px[ITEMS], py[ITEMS], pz[ITEMS];
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        const int64_t ii = i < n ? i : 0;
        px[m] = gld(sd.xin[0], ii);
        py[m] = gld(sd.xin[1], ii);
        pz[m] = gld(sd.xin[2], ii);
    }
    // a lane leads at most one run per item; every lane keeps its run's
    // index, its rank in it and its leader
    uint32_t lbin[ITEMS], llen[ITEMS];
    int lrun[ITEMS], lj[ITEMS], lrank[ITEMS], llead[ITEMS];
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        const bool valid = base + (int64_t)m * kBlock < n;
        const uint32_t key = valid ? brick_key(sd, px[m], py[m], pz[m]) : 0u;
        const uint32_t wave = blk * WAVES + (uint32_t)(m * (kBlock / 64) + wid);
        lrun[m] = -1;
        lbin[m] = 0;
        llen[m] = 0;
        const int nr = wave_runs(key, valid, lj[m], lrank[m], llead[m],
                                 [&](bool lead, uint32_t k, uint32_t len, int r) {
                                     if (lead) {
                                         lbin[m] = kb0 + k;
                                         llen[m] = len;
                                         lrun[m] = r;
                                     }
                                 });
        if (lane == 0) nruns[wave] = (uint8_t)nr;
        if (!valid) lj[m] = -1;
    }
    __syncthreads();   // the table is clear
    // every leader finds its bin's entry (linear probing from a hash of the
    // bin) and enters its run there
    int lslot[ITEMS];
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        lslot[m] = -1;
        if (lrun[m] < 0) continue;
        const uint32_t h = (lbin[m] * 2654435761u) >> (32 - HT_BITS);
        for (int p = 0; p < HT_PROBES; p++) {
            const uint32_t e = (h + (uint32_t)p) & (uint32_t)(HT - 1);
            const uint32_t prev = atomicCAS(&s_key[e], EMPTY, lbin[m]);
            if (prev == EMPTY || prev == lbin[m]) {
                lslot[m] = (int)e;
                break;
            }
        }
        const int wv = m * (kBlock / 64) + wid;
        if (lslot[m] >= 0) {
            atomicAdd(&s_cnt[lslot[m]], llen[m]);
            atomicOr(&s_mask[lslot[m]], 1u << wv);
            s_len[lslot[m]][wv] = (uint16_t)llen[m];
        } else {
            s_fail = 1;
        }
    }
    __syncthreads();
    const bool all_runs = s_fail != 0;
    const size_t c0 = (size_t)blk * TILE;
    if (!all_runs) {
        // each entry's run lengths become the runs' offsets in the chunk:
        // the lengths of the bin's runs in earlier waves
        if (s_key[t] != EMPTY) {
            const uint32_t mk = s_mask[t];
            uint32_t at = 0;
#pragma unroll
            for (int w = 0; w < WAVES; w++) {
                if ((mk >> w) & 1u) {
                    const uint32_t l = s_len[t][w];
                    s_len[t][w] = (uint16_t)at;
                    at += l;
                }
            }
        }
        // the used entries in table order are the chunks
        const bool used = s_key[t] != EMPTY;
        uint32_t tot;
        const uint32_t cc = block_exclusive_sum(used ? 1u : 0u, s_wsum, tot);
        s_cc[t] = cc;
        if (used) {
            const uint32_t g = s_key[t];
            const unsigned long long old = atomicAdd(&count[g], (1ull << 32) | s_cnt[t]);
            runrec[c0 + cc] = make_uint2(g, (uint32_t)(old >> 32));
            rlen[c0 + cc] = s_cnt[t];
        }
        if (t == 0) nchunk[blk] = tot;
    } else {
#pragma unroll
        for (int m = 0; m < ITEMS; m++) {
            if (lrun[m] < 0) continue;
            const size_t id = c0 + (size_t)(m * (kBlock / 64) + wid) * 64 + (size_t)lrun[m];
            const unsigned long long old = atomicAdd(&count[lbin[m]], (1ull << 32) | llen[m]);
            runrec[id] = make_uint2(lbin[m], (uint32_t)(old >> 32));
            rlen[id] = llen[m];
        }
        if (t == 0) nchunk[blk] = ALL_RUNS;
    }
    __syncthreads();   // s_cc
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        const int wv = m * (kBlock / 64) + wid;
        // the leader: its run's chunk and the run's offset in it
        uint32_t rcc = 0, roff = 0;
        if (lrun[m] >= 0) {
            if (!all_runs) {
                rcc = s_cc[lslot[m]];
                roff = s_len[lslot[m]][wv];
            } else {
                rcc = (uint32_t)(wv * 64 + lrun[m]);
            }
        }
        const uint32_t cc = (uint32_t)__shfl((int)rcc, llead[m], 64);
        const uint32_t off = (uint32_t)__shfl((int)roff, llead[m], 64);
        if (lj[m] >= 0) jr[c0 + (size_t)wv * 64 + lane] = (cc << 10) | (off + (uint32_t)lrank[m]);
    }
}

// Device-wide exclusive scan of 64-bit values (tiles of 4096, three phases).
__device__ __forceinline__ unsigned long long wave_inclusive_sum64(unsigned long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ unsigned long long block_exclusive_sum64(unsigned long long v,
                                                                    unsigned long long *wsum,
                                                                    unsigned long long &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long inc = wave_inclusive_sum64(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    unsigned long long woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
        const unsigned long long s = wsum[w];
        woff += (w < wid) ? s : 0ull;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return woff + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_tile_reduce64(const unsigned long long *in, int64_t m,
                                                         unsigned long long *sums) {
    __shared__ unsigned long long wsum[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    unsigned long long s = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        const int64_t idx = base + it * kBlock + threadIdx.x;
        if (idx < m) s += in[idx];
    }
    unsigned long long total;
    block_exclusive_sum64(s, wsum, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_tile_scan64(const unsigned long long *in, int64_t m,
                                                       const unsigned long long *tile_off,
                                                       unsigned long long *out) {
    __shared__ unsigned long long tile[kScanTile];
    __shared__ unsigned long long wsum[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        const int e = it * kBlock + threadIdx.x;
        const int64_t idx = base + e;
        tile[e] = idx < m ? in[idx] : 0ull;
    }
    __syncthreads();
    unsigned long long sum = 0;
    const int first = threadIdx.x * kScanItems;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) sum += tile[first + i];
    unsigned long long total;
    const unsigned long long toff = tile_off ? tile_off[blockIdx.x] : 0ull;
    unsigned long long run = block_exclusive_sum64(sum, wsum, total) + toff;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const unsigned long long v = tile[first + i];
        tile[first + i] = run;
        run += v;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kScanItems; it++) {
        const int e = it * kBlock + threadIdx.x;
        const int64_t idx = base + e;
        if (idx < m) out[idx] = tile[e];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[m] = toff + total;
}

int64_t scan64_scratch_words(int64_t m) {
    const int64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb <= 1) return 0;
    return 2 * nb + 1 + scan64_scratch_words(nb);
}

// out[i] = sum(in[0..i)) for i in [0, m]
static void exclusive_scan64(const unsigned long long *in, unsigned long long *out, int64_t m,
                             unsigned long long *scratch, hipStream_t s) {
    const int64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_tile_scan64, dim3(1), dim3(kBlock), 0, s, in, m,
                           (const unsigned long long *)nullptr, out);
        GFSX_LAUNCH_CHECK();
        return;
    }
    unsigned long long *sums = scratch, *sums_scan = scratch + nb;
    hipLaunchKernelGGL(k_tile_reduce64, dim3((unsigned)nb), dim3(kBlock), 0, s, in, m, sums);
    GFSX_LAUNCH_CHECK();
    exclusive_scan64(sums, sums_scan, nb, scratch + 2 * nb + 1, s);
    hipLaunchKernelGGL(k_tile_scan64, dim3((unsigned)nb), dim3(kBlock), 0, s, in, m,
                       (const unsigned long long *)sums_scan, out);
    GFSX_LAUNCH_CHECK();
}

// Particle and run starts of every bin from the scanned counts.
// Also clears the order kernels' two list counters and G2P's slow-list
// count (zeros: three words, no separate fill launches).
__global__ void k_bsort_split(const unsigned long long *__restrict__ start, uint32_t m,
                              uint32_t *__restrict__ pstart, uint32_t *__restrict__ rstart,
                              uint32_t *__restrict__ lcount, uint32_t *__restrict__ slow_count) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b == 0) {
        lcount[0] = 0u;
        lcount[1] = 0u;
        *slow_count = 0u;
    }
    if (b >= m) return;
    const unsigned long long v = start[b];
    pstart[b] = (uint32_t)v;
    rstart[b] = (uint32_t)(v >> 32);
}

// Every chunk into its bin's chunk list, at its provisional rank (one
// thread per binning work-group).  nlist bounds the list.
__global__ __launch_bounds__(kBlock) void k_bsort_list(int64_t nblk, const uint32_t *__restrict__ nchunk,
                                                      const uint8_t *__restrict__ nruns,
                                                      const uint2 *__restrict__ runrec,
                                                      const uint32_t *__restrict__ rstart,
                                                      uint32_t *__restrict__ list, uint32_t nlist) {
    using namespace bsort;
    const int64_t b = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (b >= nblk) return;
    const uint32_t c0 = (uint32_t)b * (uint32_t)TILE;
    auto put = [&](uint32_t id) {
        const uint2 rec = runrec[id];
        const uint32_t at = rstart[rec.x] + rec.y;
        if (at < nlist) list[at] = id;
    };
    const uint32_t nc = nchunk[b];
    if (nc != ALL_RUNS) {
        for (uint32_t c = 0; c < nc; c++) put(c0 + c);
    } else {
        for (int w = 0; w < WAVES; w++) {
            const int nr = nruns[(size_t)b * WAVES + w];
            for (int r = 0; r < nr; r++) put(c0 + (uint32_t)(w * 64 + r));
        }
    }
}

// Sorted run ids of one bin (ascending = wave order) -> each run's first
// sort slot: the bin's first slot plus the lengths of the runs before it.
template <int N>
__device__ __forceinline__ void write_bases(const uint32_t (&v)[N], uint32_t len, uint32_t at,
                                            const uint32_t *__restrict__ rlen,
                                            uint32_t *__restrict__ runbase) {
#pragma unroll
    for (int m = 0; m < N; m++) {
        if ((uint32_t)m < len) {
            runbase[v[m]] = at;
            at += rlen[v[m]];
        }
    }
}

// Bins of up to SMALL runs, one thread each (bitonic network in registers);
// larger bins are listed for the wave and work-group passes.
__global__ __launch_bounds__(kBlock) void k_bsort_order_small(
    uint32_t nbin, const uint32_t *__restrict__ rstart, const uint32_t *__restrict__ pstart,
    const uint32_t *__restrict__ list, const uint32_t *__restrict__ rlen,
    uint32_t *__restrict__ runbase, uint32_t *__restrict__ mid, uint32_t *__restrict__ huge,
    uint32_t *__restrict__ lcount) {
    using namespace bsort;
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    uint32_t s = 0, len = 0;
    if (b < nbin) {
        s = rstart[b];
        len = rstart[b + 1] - s;
    }
    // larger bins to the wave and work-group passes: one atomic per wave
    // and list
    {
        const int lane = threadIdx.x & 63;
        const uint64_t below = (1ull << lane) - 1ull;
        const bool to_mid = len > SMALL && len <= MID, to_huge = len > MID;
        const uint64_t mm = __ballot(to_mid), mh = __ballot(to_huge);
        uint32_t bm = 0, bh = 0;
        if (lane == 0) {
            if (mm) bm = atomicAdd(&lcount[0], (uint32_t)__popcll(mm));
            if (mh) bh = atomicAdd(&lcount[1], (uint32_t)__popcll(mh));
        }
        bm = __shfl(bm, 0, 64);
        bh = __shfl(bh, 0, 64);
        if (to_mid) mid[bm + (uint32_t)__popcll(mm & below)] = b;
        if (to_huge) huge[bh + (uint32_t)__popcll(mh & below)] = b;
        if (len > SMALL) len = 0;
    }
    const uint32_t at = len > 0 ? pstart[b] : 0u;
    if (len == 1) {
        runbase[list[s]] = at;
        return;
    }
    const bool need16 = __any(len > 8u), need8 = __any(len > 1u);
    if (need16) {
        uint32_t v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = (uint32_t)m < len ? list[s + m] : 0xFFFFFFFFu;
        bitonic_sort<16>(v);
        write_bases<16>(v, len, at, rlen, runbase);
    } else if (need8) {
        uint32_t v[8];
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = (uint32_t)m < len ? list[s + m] : 0xFFFFFFFFu;
        bitonic_sort<8>(v);
        write_bases<8>(v, len, at, rlen, runbase);
    }
}

// Bins of SMALL+1 .. MID runs: one wave each, K run ids per lane sorted by a
// wave bitonic network, bases from a wave prefix sum of the lengths.
template <int K>
__device__ __forceinline__ void wave_order_runs(uint32_t s, uint32_t len, uint32_t at,
                                                const uint32_t *__restrict__ list,
                                                const uint32_t *__restrict__ rlen,
                                                uint32_t *__restrict__ runbase) {
    const int lane = threadIdx.x & 63;
    uint64_t v[K];
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        v[r] = e < len ? (uint64_t)list[s + e] : ~0ull;
    }
    wave_bitonic<K>(v);
    uint32_t l[K], tot = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        l[r] = e < len ? rlen[(uint32_t)v[r]] : 0u;
        tot += l[r];
    }
    uint32_t pos = at + wave_inclusive_sum(tot) - tot;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t e = (uint32_t)(lane * K + r);
        if (e < len) runbase[(uint32_t)v[r]] = pos;
        pos += l[r];
    }
}

__global__ __launch_bounds__(kBlock) void k_bsort_order_mid(
    const uint32_t *__restrict__ rstart, const uint32_t *__restrict__ pstart,
    const uint32_t *__restrict__ list, const uint32_t *__restrict__ rlen,
    uint32_t *__restrict__ runbase, const uint32_t *__restrict__ mid,
    const uint32_t *__restrict__ lcount) {
    const uint32_t nmid = lcount[0];
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t c = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); c < nmid; c += waves) {
        const uint32_t b = mid[c];
        const uint32_t s = rstart[b], len = rstart[b + 1] - s, at = pstart[b];
        if (len <= 64u) wave_order_runs<1>(s, len, at, list, rlen, runbase);
        else wave_order_runs<4>(s, len, at, list, rlen, runbase);
    }
}

// Bins of more than MID chunks: one work-group each.  Chunk ids are
// blk * TILE + c with c below 256 (one chunk per bin and work-group) or
// wave * 64 + run (one per bin and wave), so a bin holds at most one chunk
// per id >> 6, and ordering its chunks means ordering those: the
// work-group sets one bit per id >> 6 in an LDS bitmap (a window of
// HUGE_WORDS*32 at a time), ranks each chunk by the population count below
// its bit, writes the chunks in that order to list2, then takes the bases
// from a prefix sum of the lengths in that order.
__global__ __launch_bounds__(kBlock) void k_bsort_order_huge(
    int64_t nwave, const uint32_t *__restrict__ rstart, const uint32_t *__restrict__ pstart,
    const uint32_t *__restrict__ list, uint32_t *__restrict__ list2,
    const uint32_t *__restrict__ rlen, uint32_t *__restrict__ runbase,
    const uint32_t *__restrict__ huge, const uint32_t *__restrict__ lcount) {
    using namespace bsort;
    __shared__ uint32_t bits[HUGE_WORDS], pre[HUGE_WORDS];
    __shared__ uint32_t wsum[kBlock / 64];
    const int t = threadIdx.x;
    constexpr int PER = HUGE_WORDS / kBlock;
    const uint32_t nh = lcount[1];
    for (uint32_t c = blockIdx.x; c < nh; c += gridDim.x) {
        const uint32_t b = huge[c];
        const uint32_t s = rstart[b], len = rstart[b + 1] - s;
        uint32_t done = 0;
        for (int64_t w0 = 0; w0 < nwave; w0 += (int64_t)HUGE_WORDS * 32) {
            for (int e = t; e < HUGE_WORDS; e += kBlock) bits[e] = 0u;
            __syncthreads();
            for (uint32_t e = t; e < len; e += kBlock) {
                const int64_t w = (int64_t)(list[s + e] >> 6) - w0;
                if (w >= 0 && w < (int64_t)HUGE_WORDS * 32)
                    atomicOr(&bits[w >> 5], 1u << (w & 31));
            }
            __syncthreads();
            uint32_t cnt = 0;
#pragma unroll
            for (int k = 0; k < PER; k++) cnt += __popc(bits[t * PER + k]);
            uint32_t tot;
            uint32_t run = done + block_exclusive_sum(cnt, wsum, tot);
#pragma unroll
            for (int k = 0; k < PER; k++) {
                pre[t * PER + k] = run;
                run += __popc(bits[t * PER + k]);
            }
            __syncthreads();
            for (uint32_t e = t; e < len; e += kBlock) {
                const uint32_t id = list[s + e];
                const int64_t w = (int64_t)(id >> 6) - w0;
                if (w >= 0 && w < (int64_t)HUGE_WORDS * 32) {
                    const int word = (int)(w >> 5);
                    const uint32_t below = bits[word] & ((1u << (w & 31)) - 1u);
                    list2[s + pre[word] + __popc(below)] = id;
                }
            }
            done += tot;
            __syncthreads();
        }
        uint32_t at = pstart[b];
        for (uint32_t e0 = 0; e0 < len; e0 += kBlock) {
            const uint32_t e = e0 + t;
            const uint32_t id = e < len ? list2[s + e] : 0u;
            const uint32_t l = e < len ? rlen[id] : 0u;
            uint32_t tot;
            const uint32_t ex = block_exclusive_sum(l, wsum, tot);
            if (e < len) runbase[id] = at + ex;
            at += tot;
        }
        __syncthreads();
    }
}

// Positions and velocities (when out.x[0] is set) and the particle index
// (when perm is set) written to their sort slots: the chunk's first slot
// plus the particle's offset in the chunk, both as k_bsort_bin recorded
// them (jr).  nslot bounds the slots.
struct SortedOut {
    float *x[3], *v[3];
};

__global__ __launch_bounds__(kBlock) void k_bsort_scatter(const SceneDesc *__restrict__ scenes, int nsc,
                                                         const uint32_t *__restrict__ jr,
                                                         const uint32_t *__restrict__ runbase,
                                                         SortedOut out, uint32_t *__restrict__ perm,
                                                         uint32_t nslot) {
    using namespace bsort;
    const uint32_t blk = blockIdx.x;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, blk, [](const SceneDesc &d) { return d.blk0; })];
    const int64_t n = sd.n;
    const int64_t base = (int64_t)(blk - sd.blk0) * TILE + threadIdx.x;
    const size_t c0 = (size_t)blk * TILE;
    const bool copies = out.x[0] != nullptr;
    uint32_t e[ITEMS];
    float p[ITEMS][6];
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        const int64_t i = base + (int64_t)m * kBlock;
        const bool valid = i < n;
        const int64_t ii = valid ? i : 0;
        e[m] = valid ? gld(jr, c0 + (size_t)(m * kBlock) + threadIdx.x) : 0xFFFFFFFFu;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            p[m][c] = copies ? gld(sd.xin[c], ii) : 0.0f;
            p[m][3 + c] = copies ? gld(sd.vin[c], ii) : 0.0f;
        }
    }
#pragma unroll
    for (int m = 0; m < ITEMS; m++) {
        if (e[m] == 0xFFFFFFFFu) continue;
        const uint32_t dst = gld(runbase, c0 + (e[m] >> 10)) + (e[m] & 1023u);
        if (dst >= nslot) continue;
        if (copies) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                out.x[c][dst] = p[m][c];
                out.v[c][dst] = p[m][3 + c];
            }
        }
        if (perm) perm[dst] = (uint32_t)(sd.p0 + base + (int64_t)m * kBlock);
    }
}

// ---------------------------------------------------------------------------
// Scatter P2G into per-brick tiles
// ---------------------------------------------------------------------------
//
// The work-group of brick b takes the particles of its own cells (the first
// and last brick of an axis also own the pad cells beyond the grid).  Per
// particle and component, the 2x2x2 nodes around it are the candidates of
// the reference's 3x3x3 node loop that can lie within R = dx
// (scalarfield.cpp:175-197): along the component's own axis the nodes of
// its cell, across it the nodes of its half-shifted cell.  Tile of
// component c: 5 nodes along c (4b .. 4b+4) and 6 across (4b-1 .. 4b+4),
// which holds every in-grid node a particle of the brick reaches.
//
// Sums: K and K |v| as 64-bit fixed point, K |v| into one of two sums by
// the sign of v, so every term is non-negative and keeps the 24 significant
// bits of its float product whatever the other particles' speeds.  The
// scales are powers of two per brick: W terms K * 2^eW with N * 2^eW <= 2^62
// (N particles, K <= 1), S terms K |v| * 2^eS with N * max|v| * 2^eS <= 2^62,
// so no sum overflows.  The 32 fractional bits matter: a face reached only
// at the edge of the radius has a tiny W, and S / W keeps its precision only
// while the resolution stays far below W.  A velocity that is not finite,
// or of magnitude 2^90 or more, is not added: it marks the nodes within its
// radius NaN, +inf or -inf, what the reference's float sum gives there (the
// reference adds only inside the radius, scalarfield.cpp:191-196).
namespace bscat {
constexpr int TN = 5 * 6 * 6;   // nodes per component tile
constexpr int NREP = 2;         // tile copies (by lane parity)
constexpr int NT = 256;
constexpr int PRE = 2;          // particles per thread kept in registers
constexpr float kBig = 1.2379400392853803e27f;   // 2^90
constexpr uint32_t BAD_NAN = 1u, BAD_PINF = 2u, BAD_NINF = 4u;
__host__ __device__ constexpr int tsx(int c) { return c == 0 ? 5 : 6; }
__host__ __device__ constexpr int tsy(int c) { return c == 1 ? 5 : 6; }
__host__ __device__ constexpr int tsz(int c) { return c == 2 ? 5 : 6; }
}  // namespace bscat
static_assert(3 * bscat::TN == kP2gTileNodes, "tile buffer size");

// Tile entries of a scene in HBM: per component one array of
// (nbx*SX) x (nby*SY) x (nbz*SZ) entries (x fastest), brick (bx,by,bz)'s
// tile at (bx*SX, by*SY, bz*SZ), so that the tiles of neighbouring bricks
// along x sit side by side and a face row reads them in sequence.  A scene
// owns the entries from kb0 * kP2gTileNodes on (it has at least
// nbx*nby*nbz <= nbb bins), component c at c * nbx*nby*nbz * TN after that.
struct TileGrid {
    size_t base;          // first entry of component c
    int tx, ty;           // entries along x and y
};

__device__ __forceinline__ TileGrid tile_grid(const SceneDesc &sd, int c) {
    using namespace bscat;
    TileGrid t;
    const size_t nbrick = (size_t)sd.nbx * (size_t)sd.nby * (size_t)sd.nbz;
    t.base = (size_t)sd.kb0 * kP2gTileNodes + (size_t)c * nbrick * TN;
    t.tx = sd.nbx * tsx(c);
    t.ty = sd.nby * tsy(c);
    return t;
}

__device__ __forceinline__ size_t tile_at(const TileGrid &t, int ex, int ey, int ez) {
    return t.base + (size_t)ex + (size_t)t.tx * ((size_t)ey + (size_t)t.ty * (size_t)ez);
}

// floor(y * 2^32) of y in [0, 2^31) as 64 bits: the integer part is the
// truncating conversion of y (y >= 0), the fraction y - floor(y) is exact
// (v_fract), and so is the fraction times 2^32.
__device__ __forceinline__ unsigned long long fix64(float y) {
    const uint32_t hi = (uint32_t)y;
    const float lo = __builtin_amdgcn_fractf(y) * 4294967296.0f;
    return ((unsigned long long)hi << 32) | (unsigned long long)(uint32_t)lo;
}

// Add one component of one particle.  acc: the component's three sums
// (S of v >= 0, S of v < 0, W) of this lane's tile copy; (tx,ty,tz) the tile
// position of corner (0,0,0); q*[0..1] the squared distances over R^2 to the
// two nodes per axis.  kw = K * 2^eW / 2^32 = a^2 (c0 + c1 a), a = 1 - t;
// ratio = 2^(eS - eW).  FAST: every corner in the tile and
// v finite.
template <int C, bool FAST>
__device__ __forceinline__ void bscatter_comp(unsigned long long *acc, uint32_t *sbad, int tx,
                                              int ty, int tz, const float qx[2],
                                              const float qy[2], const float qz[2], float v,
                                              float ratio, float c0, float c1, bool bad) {
    using namespace bscat;
    constexpr int SX = tsx(C), SY = tsy(C), SZ = tsz(C);
    const int base = tx + SX * (ty + SY * tz);
    unsigned long long *Sa = acc + (v < 0.0f ? TN : 0) + base;
    unsigned long long *Wa = acc + 2 * TN + base;
    const float vr = fabsf(v) * ratio;
    const uint32_t bits = v != v ? BAD_NAN : (v > 0.0f ? BAD_PINF : BAD_NINF);
    float yz[2][2];
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) yz[dy][dz] = qy[dy] + qz[dz];
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                if (!FAST) {
                    if ((unsigned)(tx + dx) >= (unsigned)SX || (unsigned)(ty + dy) >= (unsigned)SY ||
                        (unsigned)(tz + dz) >= (unsigned)SZ)
                        continue;
                }
                const int off = dx + SX * (dy + SY * dz);
                const float tt = qx[dx] + yz[dy][dz];
                // about half the corners lie beyond the radius (K = 0): those
                // lanes issue no LDS traffic.  Inside it a = 1 - t lies in
                // (0, 1] (t is a sum of squares), so no clamp is needed.
                if (tt < 1.0f) {
                    const float a = 1.0f - tt;
                    const float kw = a * a * fmaf(c1, a, c0);
                    atomicAdd(Wa + off, fix64(kw));
                    if (FAST || !bad) {
                        atomicAdd(Sa + off, fix64(kw * vr));
                    } else {
                        const int nb = C * TN + base + off;   // 4 flag bits per node
                        atomicOr(&sbad[nb >> 3], bits << ((nb & 7) * 4));
                    }
                }
            }
}

// OFF32: fewer than 2^30 sort slots, so that particle arrays are indexed by
// 32-bit byte offsets (the array addresses stay in scalar registers).
template <bool OFF32>
__global__ __launch_bounds__(bscat::NT) void k_p2g_bscatter(
    const SceneDesc *__restrict__ scenes, int nsc, uint32_t nbin, int per_xcd,
    const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ gperm, SortedOut copy,
    float2 *__restrict__ tiles, unsigned long long *__restrict__ tile_rows,
    unsigned long long *__restrict__ fluid) {
    using namespace bscat;
    __shared__ unsigned long long s_acc[NREP][3][3][TN];
    __shared__ uint32_t s_bad[(3 * TN + 7) / 8];
    __shared__ float s_red[NT / 64][3];
    __shared__ unsigned long long s_fluid;
    // consecutive bricks on one XCD: the combine reads neighbours' tiles
    const uint32_t gid = (blockIdx.x & 7u) * (uint32_t)per_xcd + (blockIdx.x >> 3);
    if (gid >= nbin) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &d) { return d.kb0; })];
    const uint32_t lb = gid - sd.kb0;
    if (lb >= sd.nbb) {
        // the overflow bin reaches no face; its particles' sorted copies
        // (when asked for) are written here
        if (copy.x[0]) {
            const uint32_t p0 = (uint32_t)sd.p0;
            for (uint32_t q = gld(pstart, gid) + t, q1 = gld(pstart, gid + 1); q < q1; q += NT) {
                const uint32_t i = gld(gperm, q) - p0;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    copy.x[c][q] = gld(sd.xin[c], i);
                    copy.v[c][q] = gld(sd.vin[c], i);
                }
            }
        }
        return;
    }
    const uint32_t q0 = gld(pstart, gid), N = gld(pstart, gid + 1) - q0;
    if (N == 0) {
        if (t == 0) fluid[gid] = 0ull;
        return;
    }
    int bxi, byi, bzi;
    brick_coords(sd, lb, bxi, byi, bzi);
    const int o[3] = {4 * bxi, 4 * byi, 4 * bzi};
    const int ni = sd.ni, nj = sd.nj, nk = sd.nk;
    const double inv_dx = sd.bg.inv_dx;
    // slot q's particle: the sorted copies at q, or (gperm set) the scene's
    // inputs at gperm[q] - p0
    const float *xp = gperm ? sd.xin[0] : sd.p2g.x, *yp = gperm ? sd.xin[1] : sd.p2g.y,
                *zp = gperm ? sd.xin[2] : sd.p2g.z;
    const float *vp[3] = {gperm ? sd.vin[0] : sd.p2g.vel[0], gperm ? sd.vin[1] : sd.p2g.vel[1],
                          gperm ? sd.vin[2] : sd.p2g.vel[2]};
    const uint32_t p0 = (uint32_t)sd.p0;
    auto at = [&](uint32_t q) { return gperm ? gld(gperm, q) - p0 : q; };
    auto ld = [&](const float *a, uint32_t i) { return OFF32 ? gld_at(a, i * 4u) : gld(a, i); };
    auto st = [&](float *a, uint32_t i, float v) {
        if (OFF32)
            gst_at(a, i * 4u, v);
        else
            gst(a, i, v);
    };
    // with copy.x[0] set, every particle read through gperm is also written
    // to its sort slot (for G2P to read in order)
    const bool copies = copy.x[0] != nullptr;
    auto put = [&](uint32_t q, float x, float y, float z, const float v[3]) {
        st(copy.x[0], q, x);
        st(copy.x[1], q, y);
        st(copy.x[2], q, z);
#pragma unroll
        for (int c = 0; c < 3; c++) st(copy.v[c], q, v[c]);
    };
    {
        unsigned long long *z = &s_acc[0][0][0][0];
        for (int e = t; e < NREP * 3 * 3 * TN; e += NT) z[e] = 0ull;
        for (int e = t; e < (3 * TN + 7) / 8; e += NT) s_bad[e] = 0u;
        if (t == 0) s_fluid = 0ull;
    }

    // pass 1: the largest |v| below 2^90 per component; the first PRE
    // particles of each thread stay in registers
    float px[PRE], py[PRE], pz[PRE], pv[PRE][3];
    float vm[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < PRE; m++) {
        const uint32_t s = (uint32_t)(t + m * NT);
        const uint32_t q = at(q0 + (s < N ? s : 0u));
        px[m] = ld(xp, q);
        py[m] = ld(yp, q);
        pz[m] = ld(zp, q);
#pragma unroll
        for (int c = 0; c < 3; c++) pv[m][c] = ld(vp[c], q);
    }
#pragma unroll
    for (int m = 0; m < PRE; m++) {
        if ((uint32_t)(t + m * NT) < N) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float av = fabsf(pv[m][c]);
                if (av < kBig) vm[c] = fmaxf(vm[c], av);
            }
            if (copies) put(q0 + (uint32_t)(t + m * NT), px[m], py[m], pz[m], pv[m]);
        }
    }
    for (uint32_t s = (uint32_t)(t + PRE * NT); s < N; s += NT) {
        const uint32_t q = at(q0 + s);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float av = fabsf(ld(vp[c], q));
            if (av < kBig) vm[c] = fmaxf(vm[c], av);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = vm[c];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
        if (lane == 0) s_red[wave][c] = v;
    }
    __syncthreads();
    const int eW = ilogb(4611686018427387904.0 / (double)N);   // 2^62 / N
    int eS[3];
    float ratio[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) v = fmaxf(v, s_red[w][c]);
        eS[c] = v > 0.0f ? min(ilogb(4611686018427387904.0 / ((double)N * (double)v)), eW + 100) : eW;
        ratio[c] = ldexpf(1.0f, eS[c] - eW);
    }
    const float c0 = ldexpf(5.0f / 9.0f, eW - 32), c1 = ldexpf(4.0f / 9.0f, eW - 32);

    // pass 2: scatter
    const int rep = lane & (NREP - 1);
    unsigned long long fl = 0ull;
    auto scatter = [&](float x, float y, float z, const float v[3]) {
        const AxisCoords ca[3] = {axis_coords(x, inv_dx), axis_coords(y, inv_dx),
                                  axis_coords(z, inv_dx)};
        // squared distances over R^2 to the two nodes per axis: unshifted
        // (the component's own axis) and half-shifted (the other two)
        float qa[3][2], qb[3][2];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float ta = ca[a].ta, tb = ca[a].tb;
            const float ua = 1.0f - ta, ub = 1.0f - tb;
            qa[a][0] = ta * ta;
            qa[a][1] = ua * ua;
            qb[a][0] = tb * tb;
            qb[a][1] = ub * ub;
        }
        // the particle's cell is fluid (fluidsimulation.cpp:2010-2017)
        if ((unsigned)ca[0].ia < (unsigned)ni && (unsigned)ca[1].ia < (unsigned)nj &&
            (unsigned)ca[2].ia < (unsigned)nk)
            fl |= 1ull << ((ca[0].ia - o[0]) + 4 * (ca[1].ia - o[1]) + 16 * (ca[2].ia - o[2]));
        int tc[3][3];
        bool fast = true;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                tc[c][a] = a == c ? ca[a].ia - o[a] : ca[a].ib - o[a] + 1;
                fast = fast && (unsigned)tc[c][a] <= (a == c ? 3u : 4u);
            }
        bool bad[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bad[c] = !(fabsf(v[c]) < kBig);
            fast = fast && !bad[c];
        }
        unsigned long long *acc = &s_acc[rep][0][0][0];
        if (fast) {
            bscatter_comp<0, true>(acc, s_bad, tc[0][0], tc[0][1], tc[0][2], qa[0], qb[1], qb[2],
                                   v[0], ratio[0], c0, c1, false);
            bscatter_comp<1, true>(acc + 3 * TN, s_bad, tc[1][0], tc[1][1], tc[1][2], qb[0], qa[1],
                                   qb[2], v[1], ratio[1], c0, c1, false);
            bscatter_comp<2, true>(acc + 6 * TN, s_bad, tc[2][0], tc[2][1], tc[2][2], qb[0], qb[1],
                                   qa[2], v[2], ratio[2], c0, c1, false);
        } else {
            bscatter_comp<0, false>(acc, s_bad, tc[0][0], tc[0][1], tc[0][2], qa[0], qb[1], qb[2],
                                    v[0], ratio[0], c0, c1, bad[0]);
            bscatter_comp<1, false>(acc + 3 * TN, s_bad, tc[1][0], tc[1][1], tc[1][2], qb[0],
                                    qa[1], qb[2], v[1], ratio[1], c0, c1, bad[1]);
            bscatter_comp<2, false>(acc + 6 * TN, s_bad, tc[2][0], tc[2][1], tc[2][2], qb[0],
                                    qb[1], qa[2], v[2], ratio[2], c0, c1, bad[2]);
        }
    };
#pragma unroll
    for (int m = 0; m < PRE; m++)
        if ((uint32_t)(t + m * NT) < N) scatter(px[m], py[m], pz[m], pv[m]);
    for (uint32_t s = (uint32_t)(t + PRE * NT); s < N; s += NT) {
        const uint32_t q = at(q0 + s);
        const float v[3] = {ld(vp[0], q), ld(vp[1], q), ld(vp[2], q)};
        const float x = ld(xp, q), y = ld(yp, q), z = ld(zp, q);
        if (copies) put(q0 + s, x, y, z, v);
        scatter(x, y, z, v);
    }
    {
        uint32_t lo = (uint32_t)fl, hi = (uint32_t)(fl >> 32);
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            lo |= __shfl_xor(lo, d, 64);
            hi |= __shfl_xor(hi, d, 64);
        }
        if (lane == 0 && (lo | hi)) atomicOr(&s_fluid, ((unsigned long long)hi << 32) | lo);
    }
    __syncthreads();

    // the copies summed, back to float, out to the tile grids: one
    // component at a time, so the tile shape is a constant
    auto flush = [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        constexpr int SX = tsx(C), SY = tsy(C), SZ = tsz(C);
        const TileGrid tg = tile_grid(sd, C);
        const int es = eS[C];
        for (int nd = t; nd < TN; nd += NT) {
            unsigned long long Sp = 0, Sn = 0, W = 0;
#pragma unroll
            for (int r = 0; r < NREP; r++) {
                Sp += s_acc[r][C][0][nd];
                Sn += s_acc[r][C][1][nd];
                W += s_acc[r][C][2][nd];
            }
            // the sums hold K |v| 2^eS and K 2^eW
            float Sf = (float)ldexp((double)(long long)(Sp - Sn), -es);
            const float Wf = (float)ldexp((double)W, -eW);
            const int e = C * TN + nd;
            const uint32_t badb = (s_bad[e >> 3] >> ((e & 7) * 4)) & 0xFu;
            if (badb) {
                if ((badb & BAD_NAN) || ((badb & BAD_PINF) && (badb & BAD_NINF))) Sf = __builtin_nanf("");
                else Sf = (badb & BAD_PINF) ? INFINITY : -INFINITY;
            }
            const int tz = nd / (SX * SY), ty = (nd / SX) % SY, tx = nd % SX;
            tiles[tile_at(tg, bxi * SX + tx, byi * SY + ty, bzi * SZ + tz)] = make_float2(Sf, Wf);
        }
    };
    flush(std::integral_constant<int, 0>());
    flush(std::integral_constant<int, 1>());
    flush(std::integral_constant<int, 2>());
    if (t == 0) {
        // the tile is written: its brick's bit in the occupied-brick mask
        atomicOr(&tile_rows[sd.rm0 + (uint32_t)sd.rmw * ((uint32_t)byi + (uint32_t)sd.nby * (uint32_t)bzi) +
                            (uint32_t)(bxi >> 6)],
                 1ull << (bxi & 63));
        fluid[gid] = s_fluid;
    }
}

// ---------------------------------------------------------------------------
// Combine: every face from the tiles that hold it
// ---------------------------------------------------------------------------
//
// One wave per combine_rows(nx) consecutive face rows along x (component
// c, node rows j, k), each row 64 faces at a time in array order.  Node
// 4b+r of an axis lies in the tile of brick b at position r (own axis:
// nodes 4b .. 4b+4) or r+1 (across: 4b-1 .. 4b+4); also in brick b-1's tile
// (r = 0: position 4 along the own axis, 5 across) or, across, in brick
// b+1's (r = 3: position 0).  The entries are added lower brick first, z
// then y then x, so each face's sum has a fixed order; a brick without
// particles (its bit clear in the occupied-brick mask of its row of bricks)
// adds nothing.  The cells on either side of a face along its axis are
// fluid when their bit is set in the mask their brick's scatter wrote
// (fluidmaterialgrid.cpp:119-143).

// Bricks and tile positions holding node n of an axis with nb bricks:
// cnt entries (brick offset db, position tp), lower brick first.
struct AxisTiles {
    int b, cnt, db[2], tp[2];
};

__device__ __forceinline__ AxisTiles axis_tiles(int n, int nb, bool own) {
    AxisTiles a;
    a.b = min(n >> 2, nb - 1);
    const int r = n - 4 * a.b;
    a.db[1] = 0;
    a.tp[1] = 0;
    if (own) {
        const bool two = r == 0 && a.b > 0;
        a.cnt = two ? 2 : 1;
        a.db[0] = two ? -1 : 0;
        a.tp[0] = two ? 4 : r;
    } else if (r == 0 && a.b > 0) {
        a.cnt = 2;
        a.db[0] = -1;
        a.tp[0] = 5;
        a.tp[1] = 1;
    } else if (r == 3 && a.b + 1 < nb) {
        a.cnt = 2;
        a.db[0] = 0;
        a.tp[0] = 4;
        a.db[1] = 1;
    } else {
        a.cnt = 1;
        a.db[0] = 0;
        a.tp[0] = r + 1;
    }
    return a;
}

// Whether cell (i,j,k) (in the grid) holds a particle.
__device__ __forceinline__ bool cell_fluid(const SceneDesc &sd,
                                           const unsigned long long *__restrict__ fluid, int i,
                                           int j, int k) {
    const uint32_t bin = sd.kb0 + brick_local(sd, i >> 2, j >> 2, k >> 2);
    return (gld(fluid, bin) >> ((i & 3) + 4 * (j & 3) + 16 * (k & 3))) & 1ull;
}

__global__ __launch_bounds__(kBlock) void k_p2g_bcombine(
    const SceneDesc *__restrict__ scenes, const int64_t *__restrict__ r0s, int nsc, int64_t nwave,
    const float2 *__restrict__ tiles, const unsigned long long *__restrict__ tile_rows,
    const unsigned long long *__restrict__ fluid) {
    using namespace bscat;
    const int lane = threadIdx.x & 63;
    // the wave index, known to the compiler as the same for all lanes, so
    // that the scene and row set-up below is scalar work
    const int64_t wv = (int64_t)blockIdx.x * (kBlock / 64) +
                       __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wv >= nwave) return;
    // the scene: a search of the dense wave offsets (a few cache lines for
    // every wave of the launch), not of the descriptors
    int si = 0;
    if (nsc > 1) {
        int lo = 0, hi = nsc - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (r0s[mid] <= wv) lo = mid;
            else hi = mid - 1;
        }
        si = lo;
    }
    const SceneDesc &sd = scenes[si];
    int64_t lw = wv - sd.r0;
    const int c = lw >= sd.cw[1] ? 2 : (lw >= sd.cw[0] ? 1 : 0);
    if (c > 0) lw -= sd.cw[c - 1];
    const int ni = sd.ni, nj = sd.nj, nk = sd.nk;
    const int nxc = ni + (c == 0), nyc = nj + (c == 1), nzc = nk + (c == 2);
    // this wave's rows, in turn: [r, r + rpw) of the component's nyc * nzc;
    // every index below is the same for all lanes
    const int rpw = combine_rows(nxc);
    const uint32_t nrows = (uint32_t)nyc * (uint32_t)nzc;
    const uint32_t r0 = (uint32_t)lw * (uint32_t)rpw;
    const uint32_t r1 = min(r0 + (uint32_t)rpw, nrows);
    int J = (int)(r0 % (uint32_t)nyc), K = (int)(r0 / (uint32_t)nyc);
    const TileGrid tg = tile_grid(sd, c);
    const int sx = tsx(c), sy = tsy(c), sz = tsz(c);
    const FaceOut fo = face_out(sd.p2g, c);
    const unsigned long long *te = reinterpret_cast<const unsigned long long *>(tiles);
    for (uint32_t r = r0; r < r1; r++) {
        const AxisTiles ay = axis_tiles(J, sd.nby, c == 1), az = axis_tiles(K, sd.nbz, c == 2);
        // per (z, y) brick row of the face row: the tile row it reads, and
        // that row's occupied-brick mask (one word, loaded once for the
        // wave, when the grid has at most 64 bricks along x)
        size_t trow[2][2];
        uint32_t mrow[2][2];
        unsigned long long mw[2][2];
#pragma unroll
        for (int iz = 0; iz < 2; iz++)
#pragma unroll
            for (int iy = 0; iy < 2; iy++) {
                const int by = ay.b + ay.db[iy], bz = az.b + az.db[iz];
                const bool ok = iz < az.cnt && iy < ay.cnt;
                trow[iz][iy] = tile_at(tg, 0, by * sy + ay.tp[iy], bz * sz + az.tp[iz]);
                mrow[iz][iy] = sd.rm0 + (uint32_t)sd.rmw * ((uint32_t)by + (uint32_t)sd.nby * (uint32_t)bz);
                mw[iz][iy] = ok && sd.rmw == 1 ? tile_rows[mrow[iz][iy]] : 0ull;
            }
        for (int I = lane; I < nxc; I += 64) {
            const AxisTiles ax = axis_tiles(I, sd.nbx, c == 0);
            const bool two = ax.cnt == 2;
            const int bx0 = ax.b + ax.db[0], bx1 = two ? ax.b + ax.db[1] : bx0;
            const size_t e0 = (size_t)(bx0 * sx + ax.tp[0]), e1 = (size_t)(bx1 * sx + ax.tp[1]);
            // the entries of the bricks that hold particles (a brick without
            // particles wrote no tile), every load issued before any is used
            uint32_t f0[2][2], f1[2][2];
#pragma unroll
            for (int iz = 0; iz < 2; iz++)
#pragma unroll
                for (int iy = 0; iy < 2; iy++) {
                    const bool ok = iz < az.cnt && iy < ay.cnt;
                    unsigned long long m0 = mw[iz][iy], m1 = mw[iz][iy];
                    if (sd.rmw > 1 && ok) {
                        m0 = gld(tile_rows, (size_t)mrow[iz][iy] + (size_t)(bx0 >> 6));
                        m1 = gld(tile_rows, (size_t)mrow[iz][iy] + (size_t)(bx1 >> 6));
                    }
                    f0[iz][iy] = ok ? (uint32_t)(m0 >> (bx0 & 63)) & 1u : 0u;
                    f1[iz][iy] = ok && two ? (uint32_t)(m1 >> (bx1 & 63)) & 1u : 0u;
                }
            unsigned long long v0[2][2], v1[2][2];   // (S, W) float pairs
#pragma unroll
            for (int iz = 0; iz < 2; iz++)
#pragma unroll
                for (int iy = 0; iy < 2; iy++) {
                    v0[iz][iy] = f0[iz][iy] ? gld(te, trow[iz][iy] + e0) : 0ull;
                    v1[iz][iy] = f1[iz][iy] ? gld(te, trow[iz][iy] + e1) : 0ull;
                }
            // the entries added lower brick first: z, then y, then x
            float S = 0.0f, W = 0.0f;
#pragma unroll
            for (int iz = 0; iz < 2; iz++)
#pragma unroll
                for (int iy = 0; iy < 2; iy++) {
                    if (f0[iz][iy]) {
                        S += __uint_as_float((uint32_t)v0[iz][iy]);
                        W += __uint_as_float((uint32_t)(v0[iz][iy] >> 32));
                    }
                    if (f1[iz][iy]) {
                        S += __uint_as_float((uint32_t)v1[iz][iy]);
                        W += __uint_as_float((uint32_t)(v1[iz][iy] >> 32));
                    }
                }
            const int64_t f = (int64_t)I + (int64_t)nxc * ((int64_t)J + (int64_t)nyc * K);
            float val = W > 0.0f ? S / W : S;
            if (fo.old) gst(fo.old, (size_t)f, val);
            if (fo.wgt) gst(fo.wgt, (size_t)f, W);
            if (fo.grav_on) {
                // the cells on either side along axis c
                const int n = c == 0 ? I : (c == 1 ? J : K);
                bool fl = false;
                if (n < fo.nax) fl = cell_fluid(sd, fluid, I, J, K);
                if (n > 0)
                    fl = fl || cell_fluid(sd, fluid, I - (c == 0), J - (c == 1), K - (c == 2));
                if (fl) val += fo.grav_delta;
            }
            gst(fo.grid, (size_t)f, val);
        }
        if (++J == nyc) {
            J = 0;
            K++;
        }
    }
}

// ---------------------------------------------------------------------------
// G2P over brick ranges
// ---------------------------------------------------------------------------

// PIC/FLIP blend (fluidsimulation.cpp:3115-3127) and RK4
// (particleadvector.cpp:1045-1054) of one particle through the staged
// boxes of a brick of GBE^3 cells, for a brick whose field keeps every RK
// stage within 0.99 cells of the start.  Returns false, computing nothing,
// when the start position lies in the grid but outside the brick's cells:
// then a stage may leave the boxes.  P2: inv_dx is a power of two.
// INTERIOR: the brick lies at least two cells inside the grid on every
// side, so a start cell in the brick and stages within 0.99 cells of the
// start keep every sample position in the grid (no grid tests); a start
// outside the grid is then left to the caller as well.
template <int ST, int GBE, bool P2, bool INTERIOR = false>
__device__ __forceinline__ bool g2p_particle_in(const StepG2P &a, const MacGrid &g,
                                                const float *const cur[3],
                                                const float *const old[3], int lx, int ly,
                                                int lz, int lox, int loy, int loz, int c0x,
                                                int c0y, int c0z, const float x0[3],
                                                const float v0[3], float x1[3], float vnew[3]) {
    const bool ing = x0[0] >= 0.0f && x0[1] >= 0.0f && x0[2] >= 0.0f && x0[0] < g.limf[0] &&
                     x0[1] < g.limf[1] && x0[2] < g.limf[2];
    if (ing) {
        // x0 >= 0 here: with P2 the FP32 product is exact
        const float fi = (float)g.inv_dx;
        const int ix = (P2 ? (int)floorf(x0[0] * fi) : (int)floor((double)x0[0] * g.inv_dx)) - c0x;
        const int iy = (P2 ? (int)floorf(x0[1] * fi) : (int)floor((double)x0[1] * g.inv_dx)) - c0y;
        const int iz = (P2 ? (int)floorf(x0[2] * fi) : (int)floor((double)x0[2] * g.inv_dx)) - c0z;
        if (!((unsigned)ix < (unsigned)GBE && (unsigned)iy < (unsigned)GBE &&
              (unsigned)iz < (unsigned)GBE))
            return false;
    } else if (INTERIOR) {
        return false;
    }
    // outside the grid every sample is 0 and the stages stay at x0
    float k1[3], vo[3];
    brick_sample_in<ST, true, GBE, P2, INTERIOR>(g, cur[0], cur[1], cur[2], old[0], old[1], old[2], lx, ly,
                                       lz, lox, loy, loz, x0[0], x0[1], x0[2], k1, vo);
    flip_blend(a, v0, k1, vo, vnew);
    float xq[3], k2[3], k3[3], k4[3], unused[3];
    auto stage = [&](const float h, const float k[3], float out[3]) {
#pragma unroll
        for (int c = 0; c < 3; c++) xq[c] = x0[c] + h * k[c];
        brick_sample_in<ST, false, GBE, P2, INTERIOR>(g, cur[0], cur[1], cur[2], nullptr, nullptr, nullptr,
                                            lx, ly, lz, 0, 0, 0, xq[0], xq[1], xq[2], out, unused);
    };
    stage(a.rk.half, k1, k2);
    stage(a.rk.half, k2, k3);
    stage(a.rk.full, k3, k4);
#pragma unroll
    for (int c = 0; c < 3; c++)
        x1[c] = x0[c] + a.rk.sixth * (((k1[c] + 2.0f * k2[c]) + 2.0f * k3[c]) + k4[c]);
    return true;
}

// The same for any start position, every sample checked against its box
// (brick_sample: a stencil that leaves the box reads global memory; the
// same bits as the boxes give).
template <int ST, int GBE, bool P2>
__device__ __forceinline__ void g2p_particle_checked(const StepG2P &a, const MacGrid &g,
                                                     const float *const cur[3],
                                                     const float *const old[3], int lx, int ly,
                                                     int lz, int lox, int loy, int loz,
                                                     const float x0[3], const float v0[3],
                                                     float x1[3], float vnew[3]) {
    float k1[3], vo[3];
    brick_sample<ST, true, GBE, P2>(g, a.old[0], a.old[1], a.old[2], cur[0], cur[1], cur[2],
                                    old[0], old[1], old[2], lx, ly, lz, lox, loy, loz, x0[0],
                                    x0[1], x0[2], k1, vo);
    flip_blend(a, v0, k1, vo, vnew);
    float xq[3], k2[3], k3[3], k4[3], unused[3];
    auto stage = [&](const float h, const float k[3], float out[3]) {
#pragma unroll
        for (int c = 0; c < 3; c++) xq[c] = x0[c] + h * k[c];
        brick_sample<ST, false, GBE, P2>(g, nullptr, nullptr, nullptr, cur[0], cur[1], cur[2],
                                         nullptr, nullptr, nullptr, lx, ly, lz, 0, 0, 0, xq[0],
                                         xq[1], xq[2], out, unused);
    };
    stage(a.rk.half, k1, k2);
    stage(a.rk.half, k2, k3);
    stage(a.rk.full, k3, k4);
#pragma unroll
    for (int c = 0; c < 3; c++)
        x1[c] = x0[c] + a.rk.sixth * (((k1[c] + 2.0f * k2[c]) + 2.0f * k3[c]) + k4[c]);
}

// Work the brick kernel leaves to k_flip_g2p_slow: single particles whose
// start cell lies outside their brick (in the grid).
struct G2PSlow {
    uint32_t *slots;      // sort slots
    uint32_t *count;      // number of slots
};

// max(a, b), NaN if either is NaN (a wave's largest |value| per component;
// a NaN value makes its brick take the checked path).
__device__ __forceinline__ float nan_max(float a, float b) {
    return (a != a || b != b) ? __builtin_nanf("") : fmaxf(a, b);
}

// One work-group per G2P brick of GBE^3 cells from cell (c0x, c0y, c0z):
// its particles are slots [q0, q1) of the brick sort (8 consecutive bins
// for GBE = 8, one bin for GBE = 4).  When the largest |value| of the
// staged new field keeps every RK stage within 0.99 cells of the start
// position (trilinear weights are convex; Catmull-Rom weights sum to at
// most 1.25 in magnitude per axis, so at most 2 overall), a start cell
// inside the brick keeps every stencil inside the staged boxes and no
// sample tests them.  Otherwise every sample is checked against its box.
template <int ST, int GBE, bool P2, bool SIMPLE>
__device__ __forceinline__ void g2p_bbrick(const StepG2P &a, const MacGrid &g, int c0x, int c0y,
                                           int c0z, uint32_t q0, uint32_t q1,
                                           const uint32_t *__restrict__ perm, G2PSlow slow) {
    using namespace gbrick;
    using B = Box<ST, GBE>;
    // SIMPLE: sorted copies in, sorted order out, no particle ids: slot q
    // reads its particle at q and writes it at q - p0
    auto ld = [&](uint32_t q, float x[3], float v[3]) {
        if (SIMPLE) {
            const uint32_t off = q * 4u;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                x[c] = gld_at(a.xin[c], off);
                v[c] = gld_at(a.vin[c], off);
            }
        } else {
            load_particle(a, q, x, v);
        }
    };
    auto put = [&](uint32_t q, const float x1[3], const float vn[3]) {
        if (SIMPLE) {
            const uint32_t o = (q - a.p0) * 4u;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                gst_at(a.xout[c], o, x1[c]);
                gst_at(a.vout[c], o, vn[c]);
            }
        } else {
            store_particle(a, perm, q, x1, vn);
        }
    };
    __shared__ float s_cur[3][B::E * B::E * B::E];
    __shared__ float s_old[3][B::EO * B::EO * B::EO];
    __shared__ uint8_t s_fast[NT / 64];   // per wave: its stages stay near their starts
    if (q1 == q0) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int lx = c0x - B::LO, ly = c0y - B::LO, lz = c0z - B::LO;
    const int lox = c0x - B::LOO, loy = c0y - B::LOO, loz = c0z - B::LOO;
    float m[3], mo[3];
    {
        float *const bc[3] = {s_cur[0], s_cur[1], s_cur[2]};
        float *const bo[3] = {s_old[0], s_old[1], s_old[2]};
        stage_mac_boxes<B::E>(bc, g.f, g.ni, g.nj, g.nk, lx, ly, lz, m);
        stage_mac_boxes<B::EO>(bo, a.old, g.ni, g.nj, g.nk, lox, loy, loz, mo);
    }
    const float reach = fabsf(a.rk.full) * (float)g.inv_dx * (ST == 0 ? 1.0f : 2.0f);
    bool wfast = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = m[c];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v = nan_max(v, __shfl_xor(v, d, 64));
        wfast = wfast && reach * v <= 0.99f;
    }
    if (lane == 0) s_fast[wave] = wfast ? 1 : 0;
    __syncthreads();
    bool fast = true;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) fast = fast && s_fast[w] != 0;
    const float *const cur[3] = {s_cur[0], s_cur[1], s_cur[2]};
    const float *const old[3] = {s_old[0], s_old[1], s_old[2]};
    if (!fast) {
        // a stage may leave the boxes: every sample checked
        for (uint32_t q = q0 + (uint32_t)t; q < q1; q += NT) {
            float x0[3], v0[3];
            ld(q, x0, v0);
            float x1[3], vnew[3];
            g2p_particle_checked<ST, GBE, P2>(a, g, cur, old, lx, ly, lz, lox, loy, loz, x0, v0,
                                              x1, vnew);
            put(q, x1, vnew);
        }
        return;
    }
    // the next particle is loaded while this one is computed
    auto run = [&](auto interior) {
        constexpr bool INTERIOR = decltype(interior)::value;
        float xn[3] = {0.0f, 0.0f, 0.0f}, vn[3] = {0.0f, 0.0f, 0.0f};
        if (q0 + (uint32_t)t < q1) ld(q0 + t, xn, vn);
        for (uint32_t q = q0 + (uint32_t)t; q < q1; q += NT) {
            const float x0[3] = {xn[0], xn[1], xn[2]};
            const float v0[3] = {vn[0], vn[1], vn[2]};
            if (q + NT < q1) ld(q + NT, xn, vn);
            float x1[3], vnew[3];
            const bool done = g2p_particle_in<ST, GBE, P2, INTERIOR>(
                a, g, cur, old, lx, ly, lz, lox, loy, loz, c0x, c0y, c0z, x0, v0, x1, vnew);
            if (done) put(q, x1, vnew);
            // the rest to the slow list, one atomic per wave (by the lowest
            // lane that has one: lanes past the end have left the loop)
            const uint64_t mk = __ballot(!done);
            if (mk) {
                const int lead = __ffsll((unsigned long long)mk) - 1;
                uint32_t base = 0;
                if (lane == lead) base = atomicAdd(slow.count, (uint32_t)__popcll(mk));
                base = __shfl(base, lead, 64);
                if (!done) slow.slots[base + (uint32_t)__popcll(mk & ((1ull << lane) - 1ull))] = q;
            }
        }
    };
    // two cells of margin to every grid face (a stage moves under a cell;
    // the margin also covers the rounding of positions)
    const bool interior = c0x >= 2 && c0y >= 2 && c0z >= 2 && c0x + GBE + 2 <= g.ni &&
                          c0y + GBE + 2 <= g.nj && c0z + GBE + 2 <= g.nk;
    if (interior)
        run(std::true_type{});
    else
        run(std::false_type{});
}

// One work-group per G2P brick of 8^3 cells of any scene.
template <int ST, bool P2, bool SIMPLE>
__device__ __forceinline__ void g2p_bb_kernel(const SceneDesc *__restrict__ scenes, int nsc,
                                              Sorted so, int keep_order, uint32_t nbrick,
                                              int per_xcd, const uint32_t *__restrict__ pstart,
                                              const uint32_t *__restrict__ perm, G2PSlow slow) {
    // consecutive bricks share box halos: keep them on one XCD
    const uint32_t gid = (blockIdx.x & 7u) * (uint32_t)per_xcd + (blockIdx.x >> 3);
    if (gid >= nbrick) return;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &d) { return d.g0; })];
    const StepG2P a = g2p_view(sd, so, keep_order);
    const MacGrid g = mac_view(sd);
    const int lg = (int)(gid - sd.g0);
    const int gx = lg % sd.gbx, gy = (lg / sd.gbx) % sd.gby, gz = lg / (sd.gbx * sd.gby);
    const uint32_t b0 = sd.kb0 + 8u * (uint32_t)lg;
    g2p_bbrick<ST, 8, P2, SIMPLE>(a, g, gx * 8, gy * 8, gz * 8, pstart[b0], pstart[b0 + 8], perm,
                                  slow);
}

// One work-group per P2G brick (4^3 cells, one bin of the sort) of any scene.
template <int ST, bool P2, bool SIMPLE>
__device__ __forceinline__ void g2p_b4_kernel(const SceneDesc *__restrict__ scenes, int nsc,
                                              Sorted so, int keep_order, uint32_t nbin,
                                              int per_xcd, const uint32_t *__restrict__ pstart,
                                              const uint32_t *__restrict__ perm, G2PSlow slow) {
    const uint32_t gid = (blockIdx.x & 7u) * (uint32_t)per_xcd + (blockIdx.x >> 3);
    if (gid >= nbin) return;
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, gid, [](const SceneDesc &d) { return d.kb0; })];
    const uint32_t lb = gid - sd.kb0;
    if (lb >= sd.nbb) return;   // the overflow bin: k_flip_g2p_slow
    int bx, by, bz;
    brick_coords(sd, lb, bx, by, bz);
    const StepG2P a = g2p_view(sd, so, keep_order);
    const MacGrid g = mac_view(sd);
    g2p_bbrick<ST, 4, P2, SIMPLE>(a, g, bx * 4, by * 4, bz * 4, pstart[gid], pstart[gid + 1], perm,
                                  slow);
}

// Trilinear: five work-groups per CU fit in LDS; five waves per SIMD.
template <bool P2, bool SIMPLE>
__global__ __launch_bounds__(gbrick::NT, 5) void k_flip_g2p_bb_tl(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order, uint32_t nbrick,
    int per_xcd, const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ perm,
    G2PSlow slow) {
    g2p_bb_kernel<0, P2, SIMPLE>(scenes, nsc, so, keep_order, nbrick, per_xcd, pstart, perm, slow);
}

// Tricubic: three waves per SIMD (168 VGPRs) run about 10% faster than two
// (256) at the cost of a few bytes of spill.
template <bool P2, bool SIMPLE>
__global__ __launch_bounds__(gbrick::NT, 3) void k_flip_g2p_bb_tc(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order, uint32_t nbrick,
    int per_xcd, const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ perm,
    G2PSlow slow) {
    g2p_bb_kernel<1, P2, SIMPLE>(scenes, nsc, so, keep_order, nbrick, per_xcd, pstart, perm, slow);
}

template <bool P2, bool SIMPLE>
__global__ __launch_bounds__(gbrick::NT, 4) void k_flip_g2p_b4_tl(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order, uint32_t nbin,
    int per_xcd, const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ perm,
    G2PSlow slow) {
    g2p_b4_kernel<0, P2, SIMPLE>(scenes, nsc, so, keep_order, nbin, per_xcd, pstart, perm, slow);
}

template <bool P2, bool SIMPLE>
__global__ __launch_bounds__(gbrick::NT, 3) void k_flip_g2p_b4_tc(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order, uint32_t nbin,
    int per_xcd, const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ perm,
    G2PSlow slow) {
    g2p_b4_kernel<1, P2, SIMPLE>(scenes, nsc, so, keep_order, nbin, per_xcd, pstart, perm, slow);
}

// The particles the brick kernel left, sampled from global memory
// (sample_mac: the same bits as the boxes give): the single slots, and each
// scene's overflow bin (particles beyond the padded grid, where every
// sample is 0).
template <int ST>
__device__ __forceinline__ void g2p_direct_one(const SceneDesc *__restrict__ scenes, int nsc,
                                               Sorted so, int keep_order,
                                               const uint32_t *__restrict__ perm, uint32_t q) {
    const SceneDesc &sd = scenes[scene_of(scenes, nsc, (uint64_t)q, [](const SceneDesc &d) { return d.p0; })];
    const StepG2P a = g2p_view(sd, so, keep_order);
    const MacGrid g = mac_view(sd);
    float x0[3], v0[3];
    load_particle(a, q, x0, v0);
    float k1[3], vo[3], vnew[3], x1[3];
    sample_mac<ST, true>(g, a.old[0], a.old[1], a.old[2], x0[0], x0[1], x0[2], k1, vo);
    flip_blend(a, v0, k1, vo, vnew);
    rk_step<ST>(g, 4, a.rk, x0, k1, x1);
    store_particle(a, perm, q, x1, vnew);
}

template <int ST>
__global__ __launch_bounds__(kBlock) void k_flip_g2p_slow(
    const SceneDesc *__restrict__ scenes, int nsc, Sorted so, int keep_order,
    const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ perm, G2PSlow slow) {
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t t0 = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t ns = *slow.count;
    for (uint32_t e = t0; e < ns; e += stride)
        g2p_direct_one<ST>(scenes, nsc, so, keep_order, perm, slow.slots[e]);
    // each scene's overflow bin by one work-group at a time
    for (int sc = (int)blockIdx.x; sc < nsc; sc += (int)gridDim.x) {
        const uint32_t ovf = scenes[sc].kb0 + scenes[sc].nbb;
        const uint32_t q1 = pstart[ovf + 1];
        for (uint32_t q = pstart[ovf] + threadIdx.x; q < q1; q += kBlock)
            g2p_direct_one<ST>(scenes, nsc, so, keep_order, perm, q);
    }
}

// ---------------------------------------------------------------------------
// The step
// ---------------------------------------------------------------------------

// A scene table entry (and, when r0 is set, its combine wave offset)
// written by a kernel from its argument (no copy from host memory: nothing
// to stage, nothing to allocate).
__global__ void k_put_scene(SceneDesc *dst, SceneDesc d, int64_t *r0) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *dst = d;
        if (r0) *r0 = d.r0;
    }
}

void put_scenes(SceneDesc *dst, const SceneDesc *src, int n, hipStream_t st, int64_t *r0) {
    for (int i = 0; i < n; i++) {
        hipLaunchKernelGGL(k_put_scene, dim3(1), dim3(64), 0, st, dst + i, src[i],
                           r0 ? r0 + i : nullptr);
        GFSX_LAUNCH_CHECK();
    }
}

void flip_step_brick(const StepBatch &s, const BrickSort &bs, hipStream_t st, hipEvent_t *ev) {
    using namespace bsort;
    if (ev) GFSX_HIP(hipEventRecord(ev[0], st));
    const int nsc = (int)s.scenes.size();
    const uint32_t nbin = bs.nbb_total;
    const unsigned nblk = (unsigned)s.tot.nblk;
    const int64_t n = s.tot.n;
    fill_u32((uint32_t *)bs.count, 0u, 2 * (int64_t)nbin, st);
    if (n > 0) {
        hipLaunchKernelGGL(k_bsort_bin, dim3(nblk), dim3(kBlock), 0, st, s.scenes_dev, nsc, bs.count,
                           bs.runrec, bs.rlen, bs.nruns, bs.nchunk, bs.jr);
        GFSX_LAUNCH_CHECK();
    }
    exclusive_scan64(bs.count, bs.start, nbin, bs.scan_scratch, st);
    hipLaunchKernelGGL(k_bsort_split, dim3(grid_for((int64_t)nbin + 1, kBlock)), dim3(kBlock), 0, st,
                       (const unsigned long long *)bs.start, nbin + 1, bs.pstart, bs.rstart, bs.lcount,
                       bs.slow_count);
    GFSX_LAUNCH_CHECK();
    SortedOut so_out;
    for (int c = 0; c < 3; c++) {
        so_out.x[c] = s.gather ? nullptr : s.xs[c];
        so_out.v[c] = s.gather ? nullptr : s.vs[c];
    }
    if (s.gather && !bs.perm) throw Error("gfsx: the particle gather needs the sort's permutation");
    if (n > 0) {
        hipLaunchKernelGGL(k_bsort_list, dim3(grid_for((int64_t)nblk, kBlock)), dim3(kBlock), 0, st,
                           (int64_t)nblk, (const uint32_t *)bs.nchunk, (const uint8_t *)bs.nruns,
                           (const uint2 *)bs.runrec, (const uint32_t *)bs.rstart, bs.list,
                           (uint32_t)n);
        GFSX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bsort_order_small, dim3(grid_for(nbin, kBlock)), dim3(kBlock), 0, st,
                           nbin, (const uint32_t *)bs.rstart, (const uint32_t *)bs.pstart,
                           (const uint32_t *)bs.list, (const uint32_t *)bs.rlen, bs.runbase, bs.mid,
                           bs.huge, bs.lcount);
        GFSX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bsort_order_mid, dim3(MID_BLOCKS), dim3(kBlock), 0, st,
                           (const uint32_t *)bs.rstart, (const uint32_t *)bs.pstart,
                           (const uint32_t *)bs.list, (const uint32_t *)bs.rlen, bs.runbase,
                           (const uint32_t *)bs.mid, (const uint32_t *)bs.lcount);
        GFSX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bsort_order_huge, dim3(HUGE_BLOCKS), dim3(kBlock), 0, st, bs.nwave,
                           (const uint32_t *)bs.rstart, (const uint32_t *)bs.pstart,
                           (const uint32_t *)bs.list, bs.list2, (const uint32_t *)bs.rlen,
                           bs.runbase, (const uint32_t *)bs.huge, (const uint32_t *)bs.lcount);
        GFSX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bsort_scatter, dim3(nblk), dim3(kBlock), 0, st, s.scenes_dev, nsc,
                           (const uint32_t *)bs.jr, (const uint32_t *)bs.runbase, so_out, bs.perm,
                           (uint32_t)n);
        GFSX_LAUNCH_CHECK();
    }
    if (ev) GFSX_HIP(hipEventRecord(ev[1], st));

    const int per_xcd = (int)((nbin + 7) / 8);
    // gather with sorted copies: the P2G kernel writes them
    SortedOut p2g_copy;
    for (int c = 0; c < 3; c++) {
        p2g_copy.x[c] = s.gather && s.xs[0] ? s.xs[c] : nullptr;
        p2g_copy.v[c] = s.gather && s.xs[0] ? s.vs[c] : nullptr;
    }
    fill_u32((uint32_t *)s.tile_rows, 0u, 2 * std::max<int64_t>(s.tot.nrm, 1), st);
    hipLaunchKernelGGL(s.tot.n < ((int64_t)1 << 30) ? k_p2g_bscatter<true> : k_p2g_bscatter<false>,
                       dim3((unsigned)(8 * (int64_t)per_xcd)), dim3(bscat::NT), 0, st,
                       s.scenes_dev, nsc, nbin, per_xcd, (const uint32_t *)bs.pstart,
                       s.gather ? (const uint32_t *)bs.perm : nullptr, p2g_copy, s.tiles,
                       s.tile_rows, bs.fluid);
    GFSX_LAUNCH_CHECK();
    if (nsc > 1 && !s.r0s) throw Error("gfsx: the combine of several scenes needs their wave offsets");
    hipLaunchKernelGGL(k_p2g_bcombine, dim3(grid_for(s.tot.nrow, kBlock / 64)), dim3(kBlock), 0, st,
                       s.scenes_dev, s.r0s, nsc, s.tot.nrow, (const float2 *)s.tiles,
                       (const unsigned long long *)s.tile_rows, (const unsigned long long *)bs.fluid);
    GFSX_LAUNCH_CHECK();
    if (s.epi) fs_epilogue(*s.epi, st);
    if (ev) GFSX_HIP(hipEventRecord(ev[2], st));

    if (n <= 0 || s.grids_only) {
        if (ev) GFSX_HIP(hipEventRecord(ev[3], st));
        return;
    }
    Sorted so;
    for (int c = 0; c < 3; c++) {
        so.x[c] = s.xs[c];
        so.v[c] = s.vs[c];
    }
    if (s.gather && !s.xs[0]) {
        if (s.g2p_variant != 0) throw Error("gfsx: the particle gather needs the brick G2P");
        so.perm = bs.perm;
    }
    // the brick G2P without the permutation read, particle ids or kept order
    bool simple = !so.perm && !s.keep_order && s.tot.n < ((int64_t)1 << 30);
    for (const auto &d : s.scenes) simple = simple && !d.id_out;
    if (s.g2p_variant == 0) {
        const uint32_t nbrick = (uint32_t)s.tot.ng2p;
        const int gper = (int)((nbrick + 7) / 8);
        G2PSlow slow;
        slow.slots = bs.slow_slots;
        slow.count = bs.slow_count;
        if (s.g2p_brick == 4) {
            const int bper = (int)((nbin + 7) / 8);
            auto *kern = s.stencil == 0
                             ? (s.p2 ? (simple ? k_flip_g2p_b4_tl<true, true> : k_flip_g2p_b4_tl<true, false>)
                                     : (simple ? k_flip_g2p_b4_tl<false, true> : k_flip_g2p_b4_tl<false, false>))
                             : (s.p2 ? (simple ? k_flip_g2p_b4_tc<true, true> : k_flip_g2p_b4_tc<true, false>)
                                     : (simple ? k_flip_g2p_b4_tc<false, true> : k_flip_g2p_b4_tc<false, false>));
            hipLaunchKernelGGL(kern, dim3((unsigned)(8 * bper)), dim3(gbrick::NT), 0, st,
                               s.scenes_dev, nsc, so, s.keep_order, nbin, bper,
                               (const uint32_t *)bs.pstart, (const uint32_t *)bs.perm, slow);
        } else {
            auto *kern = s.stencil == 0
                             ? (s.p2 ? (simple ? k_flip_g2p_bb_tl<true, true> : k_flip_g2p_bb_tl<true, false>)
                                     : (simple ? k_flip_g2p_bb_tl<false, true> : k_flip_g2p_bb_tl<false, false>))
                             : (s.p2 ? (simple ? k_flip_g2p_bb_tc<true, true> : k_flip_g2p_bb_tc<true, false>)
                                     : (simple ? k_flip_g2p_bb_tc<false, true> : k_flip_g2p_bb_tc<false, false>));
            hipLaunchKernelGGL(kern, dim3((unsigned)(8 * gper)), dim3(gbrick::NT), 0, st,
                               s.scenes_dev, nsc, so, s.keep_order, nbrick, gper,
                               (const uint32_t *)bs.pstart, (const uint32_t *)bs.perm, slow);
        }
        GFSX_LAUNCH_CHECK();
        auto *slowk = s.stencil == 0 ? k_flip_g2p_slow<0> : k_flip_g2p_slow<1>;
        hipLaunchKernelGGL(slowk, dim3(kSlowBlocks), dim3(kBlock), 0, st, s.scenes_dev, nsc, so,
                           s.keep_order, (const uint32_t *)bs.pstart, (const uint32_t *)bs.perm,
                           slow);
        GFSX_LAUNCH_CHECK();
    } else {
        flip_g2p_direct(s, bs.perm, st);
    }
    if (s.epi && s.epi->solid) fs_collide(*s.epi, st);
    if (ev) GFSX_HIP(hipEventRecord(ev[3], st));
}

}  // namespace gfsx
