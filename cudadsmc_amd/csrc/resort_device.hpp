// resort_device.hpp -- device side of the incremental re-sort shared by its
// own kernels (primitives.hip) and the fused finalize (fused.hip), which runs
// the re-sort's first phase alongside its per-trap reductions.
#pragma once

#include "dsmc_device.hpp"
#include "dsmc_internal.hpp"

namespace dsmc {

constexpr int kStayTile = kResortTile;         // slots per tile (and bucket)
constexpr int kStayShift = 11;
constexpr int kTileWords = kStayTile / 64;     // mask words per tile
constexpr int kPrepWords = 256;                // mask words per prep block (8 tiles)
constexpr int kOverflowCount = 3;              // counts[]: movers in the buckets' overflow list
static_assert((1 << kStayShift) == kStayTile, "stay tile");
static_assert(kTileWords == 32, "a tile's mask words are half a wave");
static_assert(kPrepWords % kTileWords == 0, "prep blocks hold whole tiles");

// Exclusive prefix of v over the 32 lanes of this lane's half-wave.
__device__ __forceinline__ uint32_t halfwave_exclusive_scan(uint32_t v)
{
    const int hl = threadIdx.x & 31;
    uint32_t s = v;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const uint32_t u = __shfl_up(s, d, 32);
        if (hl >= d) s += u;
    }
    return s - v;
}

// Prep blocks needed for n_cap slots (at least one, for n = 0).
__host__ __device__ inline int resort_prep_blocks(int n_cap)
{
    const int words = (n_cap + 63) / 64;
    return words > 0 ? (words + kPrepWords - 1) / kPrepWords : 1;
}

// Block k (kPrepWords threads) of the re-sort's first phase, over mask words
// [256 k, 256 k + 256) of the move's mover and removal masks (n slots):
//   * the movers and removed atoms before the block are the sums of the
//     move's per-block counts below it (every block adds them up itself, so
//     no block waits for another), those before each word a block scan;
//   * pre_m / pre_d: movers and removed atoms before each tile, entry
//     `tiles` the totals; counts = {stays, movers, removed, 0 overflow}
//     (last block);
//   * mov_v: the movers' slots in slot order;
//   * the bucket counts are cleared.
// per_tile = move blocks per tile (kStayTile / move block size).
__device__ __forceinline__ void resort_prep_block(
    int k, int n, const uint32_t *__restrict__ blk_moved, const uint32_t *__restrict__ blk_dead,
    int per_tile, const uint64_t *__restrict__ mask_m, const uint64_t *__restrict__ mask_d,
    uint32_t *__restrict__ pre_m, uint32_t *__restrict__ pre_d, int *__restrict__ counts,
    uint32_t *__restrict__ mov_v, uint32_t *__restrict__ bcnt, uint32_t *tmp,
    unsigned long long *tmp64)
{
    const int tid = (int)threadIdx.x;
    const int words = (n + 63) >> 6;
    const int tiles = n > 0 ? (n - 1) / kStayTile + 1 : 1;
    const int nblk = words > 0 ? (words + kPrepWords - 1) / kPrepWords : 1;
    if (k >= nblk) return;   // whole block leaves together
    const int g = k * kPrepWords + tid;   // this thread's word
    if (g <= (n >> kStayShift)) bcnt[g] = 0u;   // buckets 0 .. n / 2048
    uint64_t mm = 0ull, md = 0ull;
    if (g < words) { mm = mask_m[g]; md = mask_d[g]; }
    // move blocks before this block: 8 k tiles of per_tile each, a multiple
    // of 8 entries, read as 16-byte vectors, eight of each array per thread
    // in flight
    uint32_t bm = 0, bd = 0;
    {
        const uint4 *vm = (const uint4 *)blk_moved, *vd = (const uint4 *)blk_dead;
        const int nv = k * (kPrepWords / kTileWords) * per_tile / 4;
        for (int v0 = tid; v0 < nv; v0 += 8 * kPrepWords) {
            uint4 a[8], d[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int v = v0 + u * kPrepWords;
                a[u] = v < nv ? vm[v] : make_uint4(0, 0, 0, 0);
                d[u] = v < nv ? vd[v] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                bm += a[u].x + a[u].y + a[u].z + a[u].w;
                bd += d[u].x + d[u].y + d[u].z + d[u].w;
            }
        }
    }
    // two block scans: the earlier blocks' totals packed as (movers, removed)
    // in 64 bits, and the words' counts as (movers, removed) in 16-bit halves
    // (a block holds at most 16384 slots)
    static_assert(kPrepWords * 64 < 65536, "16-bit halves");
    unsigned long long tot64;
    (void)block_exclusive_scan<kPrepWords>(((unsigned long long)bm << 32) | bd, tmp64, &tot64);
    bm = (uint32_t)(tot64 >> 32);
    bd = (uint32_t)tot64;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<kPrepWords>(
        ((uint32_t)__popcll(mm) << 16) | (uint32_t)__popcll(md), tmp, &tot);
    const uint32_t tm = tot >> 16, td = tot & 0xFFFFu;
    const uint32_t om = bm + (ex >> 16);
    const uint32_t od = bd + (ex & 0xFFFFu);
    if ((g & (kTileWords - 1)) == 0 && (g >> 5) < tiles) {
        pre_m[g >> 5] = om;
        pre_d[g >> 5] = od;
    }
    if (k == nblk - 1 && tid == 0) {
        const uint32_t M = bm + tm, D = bd + td;
        pre_m[tiles] = M;
        pre_d[tiles] = D;
        counts[0] = n - (int)M - (int)D;
        counts[1] = (int)M;
        counts[2] = (int)D;
        counts[kOverflowCount] = 0;
    }
    uint32_t o = om;
    for (uint64_t b = mm; b; b &= b - 1)
        mov_v[o++] = (uint32_t)g * 64u + (uint32_t)(__ffsll((unsigned long long)b) - 1);
}

}  // namespace dsmc
