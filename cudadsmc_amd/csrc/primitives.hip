// primitives.hip -- data-movement primitives of the DSMC step written for
// wave64 / LDS on gfx950:
//   * stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits,
//     reduce-then-scan per pass (upsweep histogram, per-digit scan, downsweep
//     scatter).  Within a tile the rank of each key among equal digits comes
//     from 8 wave ballots (lanes whose digit matches) plus a per-wave running
//     count in LDS, so the order of equal keys is the input order (stable);
//     the scatter is staged through LDS so global writes are digit runs.
//     Replaces thrust::sort_by_key (collisions.cu:282-284).
//   * exclusive scan of int32 (thrust::exclusive_scan, main.cu:440-442).
//   * stable stream compaction map from a stencil (thrust::remove_if,
//     main.cu:493-498), ballot + popcount ranks.
//   * fills and a Philox block generator (for known-answer tests).
// None of these kernels waits on another block, so none can hang.
#include "dsmc_device.hpp"
#include "dsmc_internal.hpp"
#include "resort_device.hpp"

namespace dsmc {

static_assert(kSortBlock == kRadix, "downsweep assigns one digit per thread");

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortBlock) void k_radix_upsweep(
    const uint32_t *__restrict__ keys, const int *__restrict__ n_dev, int shift,
    uint32_t *__restrict__ hist, int nb)
{
    __shared__ uint32_t h[kRadix];
    const int tid = threadIdx.x;
    for (int i = tid; i < kRadix; i += kSortBlock) h[i] = 0;
    __syncthreads();
    const int n = *n_dev;
    const int tile0 = blockIdx.x * kSortTile;
    if (tile0 < n) {
        const int lane = tid & (kWave - 1), w = tid / kWave;
        const int base = tile0 + w * (kWave * kSortItems);
        const uint64_t lt = lanemask_lt();
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const int idx = base + k * kWave + lane;
            const bool valid = idx < n;
            const uint32_t d = valid ? (keys[idx] >> shift) & (kRadix - 1) : 0u;
            const uint64_t peers = match_digit8(d, __ballot(valid));
            if (valid && (peers & lt) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
        }
    }
    // tiles past the end are not scanned (block 0 always writes: n may be 0)
    if (tile0 >= n && blockIdx.x > 0) return;
    __syncthreads();
    for (int d = tid; d < kRadix; d += kSortBlock) hist[(size_t)d * nb + blockIdx.x] = h[d];
}

// One block per digit: exclusive scan of that digit's per-block counts.
__global__ __launch_bounds__(256) void k_radix_scan(uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ totals,
                                                    const int *__restrict__ n_dev, int nb)
{
    __shared__ uint32_t tmp[256 / kWave + 1];
    const int n = *n_dev;
    const int nb_used = min(nb, ((n > 0 ? n : 1) + kSortTile - 1) / kSortTile);  // tiles with data
    uint32_t *row = hist + (size_t)blockIdx.x * nb;
    const int per = (nb_used + 255) / 256;
    const int b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i) {
        const int j = b0 + i;
        if (j < nb_used) s += row[j];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan<256>(s, tmp, &total);
    for (int i = 0; i < per; ++i) {
        const int j = b0 + i;
        if (j < nb_used) { const uint32_t c = row[j]; row[j] = ex; ex += c; }
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSortBlock) void k_radix_downsweep(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out,
    const int *__restrict__ n_dev, int shift, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ totals, int nb)
{
    constexpr int W = kSortBlock / kWave;
    __shared__ uint32_t wcnt[W][kRadix];
    __shared__ uint32_t lstart[kRadix];
    __shared__ int delta[kRadix];
    __shared__ uint32_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    __shared__ uint32_t tmp[W + 1];

    const int n = *n_dev;
    const int tile0 = blockIdx.x * kSortTile;
    if (tile0 >= n) return;  // whole block leaves together

    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    for (int i = tid; i < W * kRadix; i += kSortBlock) (&wcnt[0][0])[i] = 0;
    __syncthreads();

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    const int base = tile0 + w * (kWave * kSortItems);
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const int idx = base + k * kWave + lane;
        const bool valid = idx < n;
        key[k] = valid ? keys_in[idx] : 0xFFFFFFFFu;
        val[k] = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = (key[k] >> shift) & (kRadix - 1);
        const uint64_t peers = match_digit8(d, __ballot(valid));
        const uint32_t prev = wcnt[w][d];
        // Lanes of one wave issue the read above before the leader's write below
        // (in-order LDS per wave; the write depends on the read).
        if (valid && (peers & lt) == 0) wcnt[w][d] = prev + (uint32_t)__popcll(peers);
        rank[k] = prev + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();

    {   // exclusive offsets across waves, per digit; block count of the digit
        const int d = tid;
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < W; ++ww) { const uint32_t c = wcnt[ww][d]; wcnt[ww][d] = run; run += c; }
        lstart[d] = run;
    }
    __syncthreads();
    {
        const int d = tid;
        const uint32_t cnt = lstart[d];
        uint32_t tot;
        const uint32_t lex = block_exclusive_scan<kSortBlock>(cnt, tmp, &tot);
        const uint32_t gex = block_exclusive_scan<kSortBlock>(totals[d], tmp, &tot);
        const uint32_t g = gex + hist[(size_t)d * nb + blockIdx.x];
        lstart[d] = lex;
        delta[d] = (int)(g - lex);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const int idx = base + k * kWave + lane;
        if (idx < n) {
            const uint32_t d = (key[k] >> shift) & (kRadix - 1);
            const uint32_t p = lstart[d] + wcnt[w][d] + rank[k];
            skey[p] = key[k];
            sval[p] = val[k];
        }
    }
    __syncthreads();
    const int cnt = min(kSortTile, n - tile0);
    for (int p = tid; p < cnt; p += kSortBlock) {
        const uint32_t k = skey[p];
        const uint32_t d = (k >> shift) & (kRadix - 1);
        const uint32_t g = (uint32_t)(p + delta[d]);
        keys_out[g] = k;
        vals_out[g] = sval[p];
    }
}

size_t sort_scratch_bytes(int n_cap)
{
    const int nb = div_up(n_cap > 0 ? n_cap : 1, kSortTile);
    return (size_t)kRadix * nb * sizeof(uint32_t) + (size_t)kRadix * sizeof(uint32_t) + 256;
}

void sort_scratch_bind(SortScratch &s, int n_cap, void *mem)
{
    s.n_cap = n_cap;
    s.nb = div_up(n_cap > 0 ? n_cap : 1, kSortTile);
    s.hist = (uint32_t *)mem;
    s.totals = s.hist + (size_t)kRadix * s.nb;
}

void radix_sort(const uint32_t *keys, const uint32_t *vals, uint32_t *keys_alt,
                uint32_t *vals_alt, uint32_t *keys_out, uint32_t *vals_out,
                const int *n_dev, int key_bits, const SortScratch &s,
                hipStream_t stream, Profiler *prof)
{
    const int passes = key_bits <= kRadixBits ? 1 : div_up(key_bits, kRadixBits);
    const uint32_t *sk = keys;
    const uint32_t *sv = vals;
    const double bytes_pass = (double)s.n_cap * 16.0;
    for (int p = 0; p < passes; ++p) {
        const int q = passes - 1 - p;  // passes still to go after this one
        uint32_t *dk = (q % 2 == 0) ? keys_out : keys_alt;
        uint32_t *dv = (q % 2 == 0) ? vals_out : vals_alt;
        const int shift = p * kRadixBits;
        DSMC_LAUNCH(prof, "radix_upsweep", (double)s.n_cap * 4.0, k_radix_upsweep, s.nb,
                    kSortBlock, stream, sk, n_dev, shift, s.hist, s.nb);
        DSMC_LAUNCH(prof, "radix_scan", (double)kRadix * s.nb * 8.0, k_radix_scan, kRadix, 256,
                    stream, s.hist, s.totals, n_dev, s.nb);
        DSMC_LAUNCH(prof, "radix_downsweep", bytes_pass, k_radix_downsweep, s.nb, kSortBlock,
                    stream, sk, sv, dk, dv, n_dev, shift, s.hist, s.totals, s.nb);
        sk = dk;
        sv = dv;
    }
}

// ---------------------------------------------------------------------------
// Incremental re-sort.  Input: keys key_new[0..n) of slots that were in
// stable sorted order of key_old[0..n) (the previous sort's keys), and the
// move kernel's per-64-slot masks of the slots whose key changed ("movers")
// and of removed atoms (dead key, which sorts last).  Output: the stable sort
// of key_new by (key, slot) -- the same permutation a full radix sort gives
// -- in a buffer other than key_old.  Between time steps only the atoms that
// crossed a cell face change key (a few percent); the others ("stays") keep
// their order, so nothing is sorted but the movers, and the stays are never
// copied out of slot order:
//   1-2. prep: movers and removed atoms before every 2048-slot tile (prefix
//      sums of the move kernel's per-block counts) and the movers' slots in
//      slot order (resort_prep_block; normally run by the fused finalize
//      right after the move, alongside its per-trap sums);
//   3. locate: each mover's insertion point q in slot space, the first slot
//      j with (key_old[j], j) >= (key, slot) (binary search of key_old, whose
//      stays equal key_new, through samples of it that the move kernel
//      left): the mover goes after every stay below q and before every stay
//      from q on.  Movers are bucketed by the tile of q: each takes the next
//      place of its bucket's fixed room of kBucketCap entries (one atomic per
//      bucket and wave), or, beyond the room, of an overflow list;
//   4. merge: one block per tile adds up the bucket counts of the earlier
//      tiles, ranks its bucket's movers by counting (buckets hold a few
//      dozen; a bucket beyond the LDS table of kBucketLds movers is gathered
//      from its room and the overflow list and sorted in scratch), and
//      writes the tile's stays and movers to
//      their final positions: with S = stays before the tile plus movers in
//      earlier buckets, a stay goes to S + (stays of the tile before it) +
//      (bucket movers with q <= its slot), a mover of rank r to S + (stays
//      of the tile below q) + r.  The tile's removed atoms go to the end of
//      the output, in slot order.
// No kernel waits on another block.
// ---------------------------------------------------------------------------
constexpr int kBucketLds = 1024;               // movers a bucket ranks in LDS
constexpr int kBucketCap = 1024;               // fixed room of a bucket (>= kBucketLds)
constexpr int kBucketBits = 13;
constexpr int kMaxBuckets = 1 << kBucketBits;  // tiles the locate's LDS sample table holds

// Stand-alone first phase (when the fused finalize did not run it): see
// resort_prep_block.
__global__ __launch_bounds__(kPrepWords) void k_ins_prep(
    const uint32_t *__restrict__ blk_moved, const uint32_t *__restrict__ blk_dead, int per_tile,
    const uint64_t *__restrict__ mask_m, const uint64_t *__restrict__ mask_d,
    const int *__restrict__ n_dev, uint32_t *__restrict__ pre_m, uint32_t *__restrict__ pre_d,
    int *__restrict__ counts, uint32_t *__restrict__ mov_v, uint32_t *__restrict__ bcnt)
{
    __shared__ uint32_t tmp[kPrepWords / kWave + 1];
    __shared__ unsigned long long tmp64[kPrepWords / kWave + 1];
    resort_prep_block((int)blockIdx.x, *n_dev, blk_moved, blk_dead, per_tile, mask_m, mask_d, pre_m,
                      pre_d, counts, mov_v, bcnt, tmp, tmp64);
}

// (key, slot) order; slots are distinct.
__device__ __forceinline__ bool kv_less(uint32_t ka, uint32_t va, uint32_t kb, uint32_t vb)
{
    return ka < kb || (ka == kb && va < vb);
}

// Insertion point q of every mover in slot space: the first slot j with
// (key_old[j], j) >= (key, slot) (key_old is sorted, so this is a prefix).
// The move kernel sampled key_old at every tile start (tsamp, 2048 slots
// apart) and every 64 slots (samp).  The search narrows q to one tile by a
// binary search of the tile samples in LDS, to 64 slots by the 32 samples of
// that tile (one 128-byte line, loaded at once), and then by a binary search
// of those 64 slots of key_old (two lines): the same few steps wherever the
// mover lands (a move along z crosses about Nc^2 cells' worth of atoms).
// Each mover then takes the next place of q's bucket (the 2048-slot tile of
// q): a bucket has a fixed room of kBucketCap movers, and the movers beyond
// it go to an overflow list.
__global__ __launch_bounds__(256) void k_ins_locate(
    const uint32_t *__restrict__ key_new, const uint32_t *__restrict__ key_old,
    const uint32_t *__restrict__ samp, const uint32_t *__restrict__ tsamp,
    const uint32_t *__restrict__ mov_v, int n_cap, const int *__restrict__ n_dev,
    int *__restrict__ counts, uint32_t *__restrict__ fb_k, uint32_t *__restrict__ fb_v,
    uint32_t *__restrict__ fb_q, uint32_t *__restrict__ ov_k, uint32_t *__restrict__ ov_v,
    uint32_t *__restrict__ ov_q, uint32_t *__restrict__ bcnt)
{
    __shared__ __align__(16) uint32_t lt[kMaxBuckets];
    // the first round's mover slot and the tile samples are loaded with the
    // counts (entries past them are never used); the samples four per
    // 16-byte load, up to four loads per thread in flight
    const int jf = blockIdx.x * 256 + threadIdx.x;
    uint32_t vn = jf < n_cap ? mov_v[jf] : 0u;
    const int ntc = (n_cap + kStayTile - 1) >> kStayShift;
    {
        const uint4 *t4 = (const uint4 *)tsamp;
        uint4 *l4 = (uint4 *)lt;
        const int nv = (ntc + 3) >> 2;
        for (int v0 = threadIdx.x; v0 < nv; v0 += 4 * 256) {
            uint4 c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int v = v0 + u * 256;
                c[u] = v < nv ? t4[v] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (v0 + u * 256 < nv) l4[v0 + u * 256] = c[u];
        }
    }
    const int s = *n_dev, m = counts[1];
    if (blockIdx.x * 256 >= m) return;   // whole block leaves together
    const int nt = (s + kStayTile - 1) >> kStayShift;   // tiles (samples in tsamp)
    const int nw = (s + 63) >> 6;                       // 64-slot samples in samp
    uint32_t kn = jf < m ? key_new[vn] : 0u;
    __syncthreads();
    // the grid covers 1/16 of the slots; more movers take further rounds
    for (int j0 = 0; j0 < m; j0 += (int)gridDim.x * 256) {
        const int j = j0 + blockIdx.x * 256 + threadIdx.x;
        const bool valid = j < m;
        const uint32_t v = vn, k = kn;
        if (j0 + (int)gridDim.x * 256 < m) {   // the next round's mover
            const int jn = j + (int)gridDim.x * 256;
            vn = jn < m ? mov_v[jn] : 0u;
            kn = jn < m ? key_new[vn] : 0u;
        }
        int lo = 0;
        if (valid) {
            // tiles whose first slot lies below (k, v)
            int hi = nt;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kv_less(lt[mid], (uint32_t)mid << kStayShift, k, v)) lo = mid + 1; else hi = mid;
            }
            if (lo > 0) {
                // the tile's 32 samples (one 128-byte line, loaded at once):
                // those below (k, v) are a prefix, at least the first
                const int w0 = (lo - 1) * kTileWords;
                const uint4 *s4 = (const uint4 *)(samp + w0);
                uint4 sw[kTileWords / 4];
#pragma unroll
                for (int u = 0; u < kTileWords / 4; ++u) sw[u] = s4[u];
                int c = 0;
#pragma unroll
                for (int u = 0; u < kTileWords / 4; ++u) {
                    const int w = w0 + 4 * u;
                    c += (w + 0 < nw && kv_less(sw[u].x, (uint32_t)(w + 0) << 6, k, v));
                    c += (w + 1 < nw && kv_less(sw[u].y, (uint32_t)(w + 1) << 6, k, v));
                    c += (w + 2 < nw && kv_less(sw[u].z, (uint32_t)(w + 2) << 6, k, v));
                    c += (w + 3 < nw && kv_less(sw[u].w, (uint32_t)(w + 3) << 6, k, v));
                }
                // q lies after the last such sample g and at or before g + 64
                const int g = (w0 + c - 1) << 6;
                lo = g + 1;
                hi = min(g + 64, s);
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (kv_less(key_old[mid], (uint32_t)mid, k, v)) lo = mid + 1; else hi = mid;
                }
            }
        }
        // bucket slot: lanes of a wave with the same bucket share one atomic
        // (nearby movers mostly share buckets)
        const uint32_t bkt = (uint32_t)lo >> kStayShift;
        const uint64_t peers = match_bits<kBucketBits>(bkt, __ballot(valid));
        const int leader = __ffsll((unsigned long long)peers) - 1;
        uint32_t base = 0;
        if (valid && (int)(threadIdx.x & (kWave - 1)) == leader)
            base = atomicAdd(&bcnt[bkt], (uint32_t)__popcll(peers));
        base = __shfl(base, leader < 0 ? 0 : leader, kWave);
        if (valid) {
            const uint32_t r = base + (uint32_t)__popcll(peers & lanemask_lt());
            if (r < (uint32_t)kBucketCap) {   // the bucket's fixed room
                const size_t o = (size_t)bkt * kBucketCap + r;
                fb_k[o] = k; fb_v[o] = v; fb_q[o] = (uint32_t)lo;
            } else {   // beyond it: the overflow list
                const int o = atomicAdd(&counts[kOverflowCount], 1);
                ov_k[o] = k; ov_v[o] = v; ov_q[o] = (uint32_t)lo;
            }
        }
    }
}

// Sort of one bucket that does not fit the LDS table, by one block, in
// global scratch: chunks of kBucketLds sorted in LDS (bitonic network), then
// pairwise merges of runs (merge path per thread).  Used only when a great
// many atoms land between the same two stays (e.g. nearly all atoms change
// cell in one step); O(b log b) work for one block.  Leaves the sorted bucket
// in (k0, v0, q0) or (k1, v1, q1); returns the one that holds it.
__device__ int big_bucket_sort(uint32_t *k0, uint32_t *v0, uint32_t *q0, uint32_t *k1, uint32_t *v1,
                               uint32_t *q1, int b, uint32_t *lk, uint32_t *lv, uint32_t *lq)
{
    const int tid = threadIdx.x;
    for (int c0 = 0; c0 < b; c0 += kBucketLds) {
        const int c = min(kBucketLds, b - c0);
        for (int e = tid; e < kBucketLds; e += 256) {
            const bool in = e < c;
            lk[e] = in ? k0[c0 + e] : 0xFFFFFFFFu;
            lv[e] = in ? v0[c0 + e] : 0xFFFFFFFFu;
            lq[e] = in ? q0[c0 + e] : 0xFFFFFFFFu;
        }
        __syncthreads();
        for (int size = 2; size <= kBucketLds; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int pr = tid; pr < kBucketLds / 2; pr += 256) {
                    const int i = 2 * pr - (pr & (stride - 1));   // lower index of the pair
                    const int j = i + stride;
                    const bool up = (i & size) == 0;
                    const bool gt = kv_less(lk[j], lv[j], lk[i], lv[i]);
                    if (gt == up) {
                        uint32_t t = lk[i]; lk[i] = lk[j]; lk[j] = t;
                        t = lv[i]; lv[i] = lv[j]; lv[j] = t;
                        t = lq[i]; lq[i] = lq[j]; lq[j] = t;
                    }
                }
                __syncthreads();
            }
        }
        for (int e = tid; e < c; e += 256) { k0[c0 + e] = lk[e]; v0[c0 + e] = lv[e]; q0[c0 + e] = lq[e]; }
        __syncthreads();
    }
    int src = 0;
    for (int wdt = kBucketLds; wdt < b; wdt <<= 1) {
        const uint32_t *sk = src ? k1 : k0, *sv = src ? v1 : v0, *sq = src ? q1 : q0;
        uint32_t *dk = src ? k0 : k1, *dv = src ? v0 : v1, *dq = src ? q0 : q1;
        for (int lo = 0; lo < b; lo += 2 * wdt) {
            const int na = min(wdt, b - lo), nb = max(0, min(wdt, b - lo - wdt));
            const int len = na + nb;
            const int per = (len + 255) / 256;
            const int d0 = min(tid * per, len), d1 = min(d0 + per, len);
            // merge path split of output d0 between run a = [lo, lo+na), b = [lo+na, ...)
            int l = max(0, d0 - nb), h = min(d0, na);
            while (l < h) {
                const int mid = (l + h) >> 1;
                const int jb = lo + na + d0 - 1 - mid;
                if (kv_less(sk[lo + mid], sv[lo + mid], sk[jb], sv[jb])) l = mid + 1; else h = mid;
            }
            int ia = l, ib = d0 - l;
            for (int d = d0; d < d1; ++d) {
                const bool ta = ia < na && (ib >= nb || kv_less(sk[lo + ia], sv[lo + ia],
                                                                 sk[lo + na + ib], sv[lo + na + ib]));
                const int sidx = ta ? lo + ia : lo + na + ib;
                dk[lo + d] = sk[sidx]; dv[lo + d] = sv[sidx]; dq[lo + d] = sq[sidx];
                ia += ta; ib += !ta;
            }
        }
        __syncthreads();   // this pass's writes are visible to the whole block
        src ^= 1;
    }
    return src;
}

// One block per tile t (2048 slots).  The tile's output is a contiguous
// window from w0: position p of the tile (slot g0 + p) is preceded there by
// E(p) = the stays below p plus the bucket movers inserted below p, so a stay
// at p goes to w0 + E(p) + h(p), h(p) = movers inserted at p, and a mover
// inserted at p to w0 + E(p) + (its rank among the movers inserted at p).  E
// is an exclusive scan of st(p) + h(p) over the tile, h a histogram of the
// bucket in LDS.  Buckets larger than the LDS table are sorted in global
// scratch by the block; their movers then go to w0 + (stays below p) + (rank
// in the bucket).
constexpr int kMergePer = kStayTile / 256;   // slots per thread

__global__ __launch_bounds__(256) void k_ins_merge(
    const uint32_t *__restrict__ key_new, const uint64_t *__restrict__ mask_m,
    const uint64_t *__restrict__ mask_d, const int *__restrict__ n_dev, int n_cap,
    const uint32_t *__restrict__ pre_m, const uint32_t *__restrict__ pre_d,
    const uint32_t *__restrict__ fb_k, const uint32_t *__restrict__ fb_v,
    const uint32_t *__restrict__ fb_q, const uint32_t *__restrict__ ov_k,
    const uint32_t *__restrict__ ov_v, const uint32_t *__restrict__ ov_q,
    const uint32_t *__restrict__ bcnt, const int *__restrict__ counts, uint32_t dead_key,
    uint32_t *__restrict__ srt_k, uint32_t *__restrict__ srt_v,
    uint32_t *__restrict__ srt_q, uint32_t *__restrict__ alt_k, uint32_t *__restrict__ alt_v,
    uint32_t *__restrict__ alt_q, uint32_t *__restrict__ out_k, uint32_t *__restrict__ out_v)
{
    // the bucket's movers: (key << 32 | slot), the order they are ranked in,
    // and insertion points
    __shared__ __align__(16) uint64_t lkv[kBucketLds];
    __shared__ __align__(16) uint32_t lq[kBucketLds];
    __shared__ uint32_t H[kStayTile];      // histogram h, then the scan E
    __shared__ uint64_t stw[kTileWords];   // stay bits of the tile's words
    __shared__ uint32_t wst[kTileWords];   // stays of the tile before each word
    __shared__ uint32_t tmp[256 / kWave + 1];
    __shared__ int n_big;   // movers gathered for a big bucket
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int g0 = t * kStayTile;
    const int p0 = tid * kMergePer;   // positions p0 .. p0 + 7 in the scan
    // Every load that does not need the alive count is issued together with
    // it, in one round: the tile's keys at positions tid + 256 u (coalesced;
    // up to the capacity, those past the alive slots are never used), its
    // mask words, counts and bucket size, the bucket counts of the earlier
    // tiles, and this thread's first mover of the bucket's room (used when
    // the bucket holds it).
    uint32_t kk[kMergePer];
#pragma unroll
    for (int u = 0; u < kMergePer; ++u) {
        const int p = tid + 256 * u;
        kk[u] = g0 + p < n_cap ? key_new[g0 + p] : 0u;
    }
    const int n = *n_dev;
    const uint32_t pm = pre_m[t], pd = pre_d[t];
    const int b = (int)bcnt[t];
    const int n_dead = counts[2];
    uint64_t wm = 0ull, wd = 0ull;   // mask words of the tile (in bounds up to the last tile + 1)
    if (tid < kTileWords) {
        wm = mask_m[(g0 >> 6) + tid];
        wd = mask_d[(g0 >> 6) + tid];
    }
    const size_t fb0 = (size_t)t * kBucketCap;   // the bucket's fixed room
    const uint32_t mk = fb_k[fb0 + tid], mv = fb_v[fb0 + tid], mq = fb_q[fb0 + tid];
    uint32_t bsum = 0;
    {
        const uint4 *c4 = (const uint4 *)bcnt;
        const int nv = t >> 2;
        for (int v0 = tid; v0 < nv; v0 += 4 * 256) {
            uint4 c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int v = v0 + u * 256;
                c[u] = v < nv ? c4[v] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) bsum += c[u].x + c[u].y + c[u].z + c[u].w;
        }
        if (tid < (t & 3)) bsum += bcnt[4 * nv + tid];
    }
    if (t > (n >> kStayShift)) return;   // whole block leaves together
    const int ns = min(kStayTile, n - g0);   // slots of the tile (0 for the bucket at n)
    if (tid < kTileWords) {
        const int l0 = tid * 64;
        uint64_t st = 0ull, dd = 0ull;
        uint32_t od = 0u;
        if (l0 < ns) {
            const uint64_t in = ns - l0 >= 64 ? ~0ull : (1ull << (ns - l0)) - 1ull;
            st = ~(wm | wd) & in;
            dd = wd & in;
            od = (uint32_t)(n - n_dead) + pd;
        }
        stw[tid] = st;
        wst[tid] = halfwave_exclusive_scan((uint32_t)__popcll(st));
        // removed atoms: to the end of the output, in slot order
        od += halfwave_exclusive_scan((uint32_t)__popcll(dd));
        for (uint64_t bb = dd; bb; bb &= bb - 1) {
            out_k[od] = dead_key;
            out_v[od++] = (uint32_t)(g0 + l0) + (uint32_t)(__ffsll((unsigned long long)bb) - 1);
        }
    }
#pragma unroll
    for (int u = 0; u < kMergePer; ++u) H[tid + 256 * u] = 0u;
    const bool small = b <= kBucketLds;
    // movers in earlier buckets
    uint32_t before;
    (void)block_exclusive_scan<256>(bsum, tmp, &before);   // syncs: H cleared, stay words in place
    // first output slot of the tile: stays of earlier tiles + earlier movers
    const uint32_t w0 = (uint32_t)g0 - pm - pd + before;
    const uint32_t *gk = nullptr, *gv = nullptr, *gq = nullptr;   // sorted big bucket
    if (small) {
        if (tid < b) {
            lkv[tid] = (uint64_t)mk << 32 | mv;
            lq[tid] = mq;
            atomicAdd(&H[mq - (uint32_t)g0], 1u);
        }
        for (int e = tid + 256; e < b; e += 256) {
            const uint32_t q = fb_q[fb0 + e];
            lkv[e] = (uint64_t)fb_k[fb0 + e] << 32 | fb_v[fb0 + e];
            lq[e] = q;
            atomicAdd(&H[q - (uint32_t)g0], 1u);
        }
    } else {
        // gather the bucket (its room, then its movers in the overflow list)
        // in scratch and sort it there
        uint32_t *ak = srt_k + before, *av = srt_v + before, *aq = srt_q + before;
        uint32_t *ck = alt_k + before, *cv = alt_v + before, *cq = alt_q + before;
        const int nr = min(b, kBucketCap);   // movers in the room
        for (int e = tid; e < nr; e += 256) {
            ak[e] = fb_k[fb0 + e]; av[e] = fb_v[fb0 + e]; aq[e] = fb_q[fb0 + e];
        }
        if (tid == 0) n_big = nr;
        __syncthreads();
        const int no = counts[kOverflowCount];
        for (int e = tid; e < no; e += 256) {
            const uint32_t q = ov_q[e];
            if ((q >> kStayShift) == (uint32_t)t) {
                const int o = atomicAdd(&n_big, 1);
                ak[o] = ov_k[e]; av[o] = ov_v[e]; aq[o] = q;
            }
        }
        __syncthreads();
        uint32_t *lk = (uint32_t *)lkv, *lv = lk + kBucketLds;   // scratch of the sort
        const int which = big_bucket_sort(ak, av, aq, ck, cv, cq, b, lk, lv, lq);
        gk = which ? ck : ak; gv = which ? cv : av; gq = which ? cq : aq;
        for (int r = tid; r < b; r += 256) atomicAdd(&H[gq[r] - (uint32_t)g0], 1u);
    }
    __syncthreads();
    // E: exclusive scan of st(p) + h(p) over the tile, 8 consecutive
    // positions per thread
    const uint32_t st8 = (uint32_t)(stw[p0 >> 6] >> (p0 & 63)) & 0xFFu;
    uint32_t e[kMergePer], run = 0;
#pragma unroll
    for (int u = 0; u < kMergePer; ++u) {
        e[u] = run;
        run += ((st8 >> u) & 1u) + H[p0 + u];
    }
    uint32_t total;
    const uint32_t base = block_exclusive_scan<256>(run, tmp, &total);   // syncs: H read
#pragma unroll
    for (int u = 0; u < kMergePer; ++u) H[p0 + u] = base + e[u];
    __syncthreads();   // E complete
    // stays, at positions tid + 256 u (coalesced stores): E(p) + h(p) =
    // E(p + 1) - st(p) = E(p + 1) - 1
#pragma unroll
    for (int u = 0; u < kMergePer; ++u) {
        const int p = tid + 256 * u;
        if (p >= ns || !((stw[p >> 6] >> (p & 63)) & 1ull)) continue;
        const uint32_t o = w0 + (p + 1 < kStayTile ? H[p + 1] : total) - 1u;
        out_k[o] = kk[u];
        out_v[o] = (uint32_t)(g0 + p);
    }
    if (small) {
        for (int x = tid; x < b; x += 256) {
            const uint32_t q = lq[x], p = q - (uint32_t)g0;
            const uint32_t ep = H[p];
            const uint32_t stp = (uint32_t)(stw[p >> 6] >> (p & 63)) & 1u;
            const uint32_t hp = (p + 1 < (uint32_t)kStayTile ? H[p + 1] : total) - ep - stp;
            uint32_t r = 0;   // rank among the movers inserted at the same point
            const uint64_t kx = lkv[x];
            if (hp > 1) {
                // four movers per step, without branches: (key, slot) order is
                // the order of key << 32 | slot
                const uint4 *lq4 = (const uint4 *)lq;
                const ulonglong2 *k2 = (const ulonglong2 *)lkv;
#pragma unroll 2
                for (int f = 0; f < b; f += 4) {
                    const uint4 qq = lq4[f >> 2];
                    const ulonglong2 ka = k2[f >> 1], kb = k2[(f >> 1) + 1];
                    r += (uint32_t)((qq.x == q) & (ka.x < kx));
                    r += (uint32_t)((qq.y == q) & (ka.y < kx) & (f + 1 < b));
                    r += (uint32_t)((qq.z == q) & (kb.x < kx) & (f + 2 < b));
                    r += (uint32_t)((qq.w == q) & (kb.y < kx) & (f + 3 < b));
                }
            }
            const uint32_t o = w0 + ep + r;
            out_k[o] = (uint32_t)(kx >> 32);
            out_v[o] = (uint32_t)kx;
        }
    } else {
        for (int r = tid; r < b; r += 256) {
            const int p = (int)(gq[r] - (uint32_t)g0);
            const uint64_t below = (p & 63) ? stw[p >> 6] & (~0ull >> (64 - (p & 63))) : 0ull;
            const uint32_t o = w0 + wst[p >> 6] + (uint32_t)__popcll(below) + (uint32_t)r;
            out_k[o] = gk[r];
            out_v[o] = gv[r];
        }
    }
}

// Per-slot arrays (movers, big-bucket scratch, radix alternates), the
// buckets' fixed rooms (kBucketCap movers for each of the tiles + 1 buckets),
// and per-tile counts.
static size_t resort_room(size_t tiles) { return (tiles + 1) * (size_t)kBucketCap; }

size_t resort_scratch_bytes(int n_cap)
{
    const size_t n = (size_t)(n_cap > 0 ? n_cap : 1);
    const size_t tiles = (size_t)div_up((long long)n, kStayTile);
    return 10 * n * sizeof(uint32_t) + 3 * resort_room(tiles) * sizeof(uint32_t) +
           4 * (tiles + 2) * sizeof(uint32_t) + 8 * sizeof(int) + 24 * 256;
}

void resort_scratch_bind(ResortScratch &r, int n_cap, void *mem)
{
    auto align = [](char *p) { return (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255); };
    const size_t n = (size_t)(n_cap > 0 ? n_cap : 1);
    r.n_cap = (int)n;
    r.tiles = div_up((long long)n, kStayTile);
    char *p = align((char *)mem);
    uint32_t **arrays[10] = {&r.mov_k, &r.mov_v, &r.mov_q, &r.mov_r, &r.srt_k,
                             &r.srt_v, &r.srt_q, &r.alt_k, &r.alt_v, &r.spare};
    for (auto a : arrays) { *a = (uint32_t *)p; p = align(p + n * sizeof(uint32_t)); }
    uint32_t **rooms[3] = {&r.bk_k, &r.bk_v, &r.bk_q};
    for (auto a : rooms) { *a = (uint32_t *)p; p = align(p + resort_room(r.tiles) * sizeof(uint32_t)); }
    r.tile_cnt = (uint32_t *)p; p = align(p + (size_t)(r.tiles + 2) * sizeof(uint32_t));
    r.tile_dead = (uint32_t *)p; p = align(p + (size_t)(r.tiles + 2) * sizeof(uint32_t));
    r.bcnt = (uint32_t *)p; p = align(p + (size_t)(r.tiles + 2) * sizeof(uint32_t));
    r.counts = (int *)p;
}

bool resort_supported(int n_cap)
{
    // the locate keeps one LDS sample per tile (n / 2048 + 1 buckets)
    return (n_cap > 0 ? n_cap : 1) / kStayTile + 1 <= kMaxBuckets;
}

size_t resort_mask_words(int n_cap)
{
    // whole tiles, so that every word of a move grid of blocks up to 2048
    // slots is inside
    return (size_t)div_up(n_cap > 0 ? n_cap : 1, kStayTile) * kTileWords + kTileWords;
}

ResortPrepArgs resort_prep_args(const ResortScratch &r, const int *n_dev, int move_block,
                                const uint32_t *blk_moved, const uint32_t *blk_dead,
                                const uint64_t *mask_m, const uint64_t *mask_d)
{
    ResortPrepArgs a;
    a.blocks = resort_prep_blocks(r.n_cap);
    a.per_tile = kStayTile / move_block;
    a.n_dev = n_dev;
    a.blk_moved = blk_moved; a.blk_dead = blk_dead;
    a.mask_m = mask_m; a.mask_d = mask_d;
    a.pre_m = r.tile_cnt; a.pre_d = r.tile_dead; a.mov_v = r.mov_v; a.bcnt = r.bcnt;
    a.counts = r.counts;
    return a;
}

void resort_incremental(const uint32_t *key_new, const uint32_t *key_old, const int *n_dev,
                        uint32_t dead_key, const uint32_t *blk_moved, const uint32_t *blk_dead,
                        const uint64_t *mask_m, const uint64_t *mask_d,
                        const uint32_t *key_samp, const uint32_t *key_tsamp, int move_block,
                        bool prepped, const ResortScratch &r, uint32_t *out_k, uint32_t *out_v,
                        hipStream_t stream, Profiler *prof)
{
    const double n = (double)r.n_cap;
    if (!prepped)
        DSMC_LAUNCH(prof, "resort_prep", 0.0, k_ins_prep, resort_prep_blocks(r.n_cap), kPrepWords,
                    stream, blk_moved, blk_dead, kStayTile / move_block, mask_m, mask_d, n_dev,
                    r.tile_cnt, r.tile_dead, r.counts, r.mov_v, r.bcnt);
    // the overflow list reuses the mover arrays that the locate does not read
    DSMC_LAUNCH(prof, "resort_locate", 0.0, k_ins_locate, div_up(div_up(r.n_cap, 16), 256), 256,
                stream, key_new, key_old, key_samp, key_tsamp, r.mov_v, r.n_cap, n_dev, r.counts,
                r.bk_k, r.bk_v, r.bk_q, r.mov_k, r.mov_r, r.mov_q, r.bcnt);
    DSMC_LAUNCH(prof, "resort_merge", n * 12.0, k_ins_merge, div_up(r.n_cap, kStayTile) + 1, 256,
                stream, key_new, mask_m, mask_d, n_dev, r.n_cap, r.tile_cnt, r.tile_dead, r.bk_k, r.bk_v,
                r.bk_q, r.mov_k, r.mov_r, r.mov_q, r.bcnt, r.counts, dead_key, r.srt_k, r.srt_v,
                r.srt_q, r.alt_k, r.alt_v, r.spare, out_k, out_v);
}

// ---------------------------------------------------------------------------
// Exclusive scan (int32)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int *__restrict__ in, int n,
                                                            int *__restrict__ partial)
{
    __shared__ int tmp[kScanBlock / kWave + 1];
    const int tile0 = blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = tile0 + k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    int total;
    (void)block_exclusive_scan<kScanBlock>(s, tmp, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single block: exclusive scan of partial[0..nb) in place; partial[nb] = total.
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(int *__restrict__ partial, int nb)
{
    __shared__ int tmp[kScanBlock / kWave + 1];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += kScanBlock) {
        const int i = c0 + threadIdx.x;
        const int v = i < nb ? partial[i] : 0;
        int total;
        const int ex = block_exclusive_scan<kScanBlock>(v, tmp, &total);
        if (i < nb) partial[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_final(const int *__restrict__ in,
                                                           int *__restrict__ out, int n,
                                                           const int *__restrict__ partial)
{
    __shared__ int tmp[kScanBlock / kWave + 1];
    const int t0 = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (t0 + k < n) ? in[t0 + k] : 0;
        s += v[k];
    }
    int total;
    int run = block_exclusive_scan<kScanBlock>(s, tmp, &total) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (t0 + k < n) out[t0 + k] = run;
        run += v[k];
    }
}

void exclusive_scan(const int *in, int *out, int n, int *partial, hipStream_t stream,
                    Profiler *prof)
{
    if (n <= 0) return;
    const int nb = div_up(n, kScanTile);
    DSMC_LAUNCH(prof, "scan_reduce", (double)n * 4.0, k_scan_reduce, nb, kScanBlock, stream,
                in, n, partial);
    DSMC_LAUNCH(prof, "scan_partials", (double)nb * 8.0, k_scan_partials, 1, kScanBlock,
                stream, partial, nb);
    DSMC_LAUNCH(prof, "scan_final", (double)n * 8.0, k_scan_final, nb, kScanBlock, stream, in,
                out, n, (const int *)partial);
}

// ---------------------------------------------------------------------------
// Stable compaction map
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanBlock) void k_compact_count(const int *__restrict__ stencil,
                                                              const int *__restrict__ n_dev,
                                                              int *__restrict__ partial)
{
    __shared__ int tmp[kScanBlock / kWave + 1];
    const int n = *n_dev;
    const int tile0 = blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = tile0 + k * kScanBlock + threadIdx.x;
        if (i < n) s += stencil[i] == 0;
    }
    int total;
    (void)block_exclusive_scan<kScanBlock>(s, tmp, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_compact_scatter(
    const int *__restrict__ stencil, const int *__restrict__ n_dev,
    const int *__restrict__ partial, uint32_t *__restrict__ map, int *__restrict__ n_out, int nb)
{
    constexpr int W = kScanBlock / kWave;
    __shared__ int wtot[W];
    const int n = *n_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = partial[nb];
    const int tile0 = blockIdx.x * kScanTile;
    if (tile0 >= n) return;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int base = tile0 + w * (kWave * kScanItems);
    const uint64_t lt = lanemask_lt();
    uint32_t rank[kScanItems];
    bool keep[kScanItems];
    int run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k * kWave + lane;
        keep[k] = i < n && stencil[i] == 0;
        const uint64_t b = __ballot(keep[k]);
        rank[k] = run + __popcll(b & lt);
        run += __popcll(b);
    }
    if (lane == 0) wtot[w] = run;
    __syncthreads();
    int off = partial[blockIdx.x];
    for (int ww = 0; ww < w; ++ww) off += wtot[ww];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (keep[k]) map[off + rank[k]] = (uint32_t)(base + k * kWave + lane);
}

void compaction_map(const int *stencil, const int *n_dev, int n_cap, uint32_t *map,
                    int *n_out_dev, int *partial, hipStream_t stream, Profiler *prof)
{
    const int nb = div_up(n_cap > 0 ? n_cap : 1, kScanTile);
    DSMC_LAUNCH(prof, "compact_count", (double)n_cap * 4.0, k_compact_count, nb, kScanBlock,
                stream, stencil, n_dev, partial);
    DSMC_LAUNCH(prof, "scan_partials", (double)nb * 8.0, k_scan_partials, 1, kScanBlock, stream,
                partial, nb);
    DSMC_LAUNCH(prof, "compact_scatter", (double)n_cap * 8.0, k_compact_scatter, nb, kScanBlock,
                stream, stencil, n_dev, (const int *)partial, map, n_out_dev, nb);
}

__global__ __launch_bounds__(256) void k_gather_u32(uint32_t *__restrict__ dst,
                                                    const uint32_t *__restrict__ src,
                                                    const uint32_t *__restrict__ map,
                                                    const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < *n_dev) dst[i] = src[map[i]];
}

void gather_u32(uint32_t *dst, const uint32_t *src, const uint32_t *map, const int *n_dev,
                int n_cap, hipStream_t stream, Profiler *prof)
{
    DSMC_LAUNCH(prof, "gather_u32", (double)n_cap * 12.0, k_gather_u32, div_up(n_cap > 0 ? n_cap : 1, 256),
                256, stream, dst, src, map, n_dev);
}

// ---------------------------------------------------------------------------
// Fills and Philox
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint32_t v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double *p, double v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

static int fill_grid(long long n) { return (int)std::min<long long>(4096, (n + 255) / 256 + 1); }

void fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t stream, Profiler *prof)
{
    if (n <= 0) return;
    DSMC_LAUNCH(prof, "fill", (double)n * 4.0, k_fill_u32, fill_grid(n), 256, stream, p, v, n);
}

void fill_f64(double *p, double v, long long n, hipStream_t stream, Profiler *prof)
{
    if (n <= 0) return;
    DSMC_LAUNCH(prof, "fill", (double)n * 8.0, k_fill_f64, fill_grid(n), 256, stream, p, v, n);
}

// Streaming copy for counter calibration: four 16-byte vectors per thread
// and round, loads issued together, non-temporal loads and stores (each byte
// is touched once), grid-stride over whole rounds.
constexpr int kCopyUnroll = 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const u32x4 *__restrict__ src,
                                                     u32x4 *__restrict__ dst, long long n16)
{
    const long long stride = (long long)gridDim.x * 256 * kCopyUnroll;
    for (long long i0 = (long long)blockIdx.x * 256 * kCopyUnroll + threadIdx.x; i0 < n16; i0 += stride) {
        u32x4 v[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const long long i = i0 + (long long)u * 256;
            if (i < n16) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const long long i = i0 + (long long)u * 256;
            if (i < n16) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

void stream_copy(void *dst, const void *src, size_t bytes, hipStream_t stream)
{
    const long long n16 = (long long)(bytes / 16);
    if (n16 <= 0) return;
    const int grid = (int)std::min<long long>((n16 + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll), 8192LL);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, stream, (const u32x4 *)src,
                       (u32x4 *)dst, n16);
    DSMC_HIP(hipGetLastError());
}

// Gather of the FP32 state's tiled 16-byte records (three planes of 64
// records per 64-slot tile) through a slot map: record planes of slot i of
// dst = those of slot map[i] of src.  The access pattern of the fused move's
// state gather, for calibrating the traffic counters on it: 4 bytes of map
// and 48 bytes of records read, 48 bytes written per slot.
__global__ __launch_bounds__(256) void k_gather_copy(const u32x4 *__restrict__ src,
                                                     u32x4 *__restrict__ dst,
                                                     const uint32_t *__restrict__ map, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = __builtin_nontemporal_load(map + i);
    const size_t qs = (size_t)(j / 64) * 192 + (j % 64), qd = (size_t)(i / 64) * 192 + (i % 64);
    const u32x4 a = __builtin_nontemporal_load(src + qs), b = __builtin_nontemporal_load(src + qs + 64),
                c = __builtin_nontemporal_load(src + qs + 128);
    __builtin_nontemporal_store(a, dst + qd);
    __builtin_nontemporal_store(b, dst + qd + 64);
    __builtin_nontemporal_store(c, dst + qd + 128);
}

void gather_copy(void *dst, const void *src, const uint32_t *map, int n, hipStream_t stream)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_copy, dim3(div_up(n, 256)), dim3(256), 0, stream, (const u32x4 *)src,
                       (u32x4 *)dst, map, n);
    DSMC_HIP(hipGetLastError());
}

__global__ void k_philox_fill(uint32_t *out, int n, U32x4 ctr, uint32_t k0, uint32_t k1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U32x4 c = { ctr.x + (uint32_t)i, ctr.y, ctr.z, ctr.w };
    U32x4 o = philox(c, k0, k1);
    out[4 * i + 0] = o.x;
    out[4 * i + 1] = o.y;
    out[4 * i + 2] = o.z;
    out[4 * i + 3] = o.w;
}

void philox_fill(uint32_t *out, int n, const uint32_t ctr[4], const uint32_t key[2],
                 hipStream_t stream)
{
    if (n <= 0) return;
    U32x4 c = { ctr[0], ctr[1], ctr[2], ctr[3] };
    hipLaunchKernelGGL(k_philox_fill, dim3(div_up(n, 256)), dim3(256), 0, stream, out, n, c,
                       key[0], key[1]);
    DSMC_HIP(hipGetLastError());
}

}  // namespace dsmc
