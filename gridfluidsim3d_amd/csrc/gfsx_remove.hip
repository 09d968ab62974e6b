// gfsx_remove.hip -- particle removal after the step, on gfx950:
// FluidSimulation::_removeMarkerParticlesInSolidCells and
// _removeMarkerParticles (fluidsimulation.cpp:1985-1988, 3213-3243).
//
// Particles in solid cells are dropped; a cell keeps at most max_per_cell
// particles.  The reference keeps the first ones of a rand() shuffle of the
// whole array; here a cell keeps the max_per_cell particles with the
// smallest SplitMix64 key of (seed, input index), a seeded order that is the
// same on every run and every machine.  After the cell sort (integer work),
// k_remove_mark keeps the particles of cells that are not over-full and
// lists the over-full cells; k_remove_select finds, per listed cell, the
// max_per_cell-th smallest key by a radix select over the cell's keys (eight
// byte passes, each a 256-bucket histogram in LDS: linear in the cell's
// count) and keeps the particles whose key does not exceed it; a scan of the
// keep flags gives each kept particle its output index and k_remove_compact
// copies the kept particles in input order.

#include "gfsx_internal.h"

#include <algorithm>

namespace gfsx {

namespace {

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>((n + kBlock - 1) / kBlock, 1);
}

// SplitMix64 (Steele, Lea, Flood 2014; constants of Vigna's splitmix64.c):
// a bijection of i for a fixed seed, so keys never tie
__device__ __forceinline__ uint64_t removal_key(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_remove_mark(RemoveArgs r) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= r.n) return;
    const uint32_t b = bin_of(r.x[0][p], r.x[1][p], r.x[2][p], r.bg);
    uint32_t k = 1;
    if (b != r.bg.nbin) {   // outside the grid: kept
        if (r.solid && r.solid[b]) {
            k = 0;
        } else {
            const uint32_t q0 = r.start[b], q1 = r.start[b + 1];
            if (q1 - q0 > (uint32_t)r.max_per_cell) {
                // decided by k_remove_select; the cell's first particle in
                // sort order lists it
                k = 0;
                if (r.perm[q0] == (uint32_t)p) r.full[atomicAdd(r.nfull, 1u)] = b;
            }
        }
    }
    r.keep[p] = k;
}

// One work-group per over-full cell (grid-stride over the list): the
// max_per_cell-th smallest key of the cell, byte by byte from the top, then
// keep[p] = 1 for the particles whose key is not above it.  Keys are
// distinct (a bijection of the index), so exactly max_per_cell are kept.
static_assert(kBlock == 256, "one histogram bucket per thread");

__global__ __launch_bounds__(kBlock) void k_remove_select(RemoveArgs r) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_rank;
    const uint32_t nfull = *r.nfull;
    const int t = threadIdx.x;
    for (uint32_t c = blockIdx.x; c < nfull; c += gridDim.x) {
        const uint32_t b = r.full[c];
        const uint32_t q0 = r.start[b], q1 = r.start[b + 1];
        if (t == 0) {
            s_prefix = 0ull;
            s_rank = (uint32_t)r.max_per_cell;   // 1-based rank still to find
        }
        for (int pass = 7; pass >= 0; pass--) {
            const int sh = 8 * pass;
            hist[t] = 0u;   // kBlock == 256 buckets
            __syncthreads();
            const uint64_t prefix = s_prefix;
            for (uint32_t q = q0 + t; q < q1; q += kBlock) {
                const uint64_t key = removal_key(r.seed, (uint64_t)r.perm[q]);
                // the bytes above this pass match the prefix found so far
                const bool match = pass == 7 || (key >> (sh + 8)) == (prefix >> (sh + 8));
                if (match) atomicAdd(&hist[(key >> sh) & 255u], 1u);
            }
            __syncthreads();
            if (t == 0) {
                uint32_t rank = s_rank, d = 0;
                while (hist[d] < rank) {
                    rank -= hist[d];
                    d++;
                }
                s_rank = rank;
                s_prefix = prefix | ((uint64_t)d << sh);
            }
            __syncthreads();
        }
        const uint64_t kth = s_prefix;
        for (uint32_t q = q0 + t; q < q1; q += kBlock) {
            const uint32_t p = r.perm[q];
            if (removal_key(r.seed, (uint64_t)p) <= kth) r.keep[p] = 1u;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_remove_compact(RemoveArgs r) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= r.n) return;
    if (!r.keep[p]) return;
    const uint32_t o = r.dst[p];
    for (int c = 0; c < 3; c++) {
        r.xo[c][o] = r.x[c][p];
        r.vo[c][o] = r.v[c][p];
    }
    if (r.id_out) r.id_out[o] = r.id_in ? r.id_in[p] : (uint32_t)p;
}

}  // namespace

void remove_mark(const RemoveArgs &r, hipStream_t s) {
    if (r.n <= 0) return;
    fill_u32(r.nfull, 0u, 1, s);
    hipLaunchKernelGGL(k_remove_mark, dim3(blocks_for(r.n)), dim3(kBlock), 0, s, r);
    GFSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_remove_select, dim3(1024), dim3(kBlock), 0, s, r);
    GFSX_LAUNCH_CHECK();
}

void remove_compact(const RemoveArgs &r, hipStream_t s) {
    if (r.n <= 0) return;
    hipLaunchKernelGGL(k_remove_compact, dim3(blocks_for(r.n)), dim3(kBlock), 0, s, r);
    GFSX_LAUNCH_CHECK();
}

}  // namespace gfsx
