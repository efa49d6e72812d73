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
// stable sorted order of key_old[0..n) (the previous sort's keys).  Output:
// the stable sort of key_new by (key, slot) -- identical to a full radix sort
// -- built as
//   1. split: slots whose key did not change ("stay", already in order) and
//      the others ("movers"), both in slot order;
//   2. radix sort of the movers alone (stable, so ties stay in slot order);
//   3. merge path of the two sorted runs by (key, slot).
// Between time steps only atoms that crossed a cell face (a few percent)
// are movers, so the full-width sort runs on a few percent of the data.
// ---------------------------------------------------------------------------
constexpr int kSplitTile = kSortTile;    // slots per split block (256 x 8)
constexpr int kMergeTile = kSortTile;    // outputs per merge block
constexpr int kMergeItems = kMergeTile / kSortBlock;

__global__ __launch_bounds__(kSortBlock) void k_split_count(const uint32_t *__restrict__ key_new,
                                                            const uint32_t *__restrict__ key_old,
                                                            const int *__restrict__ n_dev,
                                                            uint32_t *__restrict__ tile_cnt)
{
    __shared__ uint32_t tmp[kSortBlock / kWave + 1];
    const int n = *n_dev;
    const int tile0 = blockIdx.x * kSplitTile;
    uint32_t c = 0;
    if (tile0 < n) {
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const int i = tile0 + k * kSortBlock + threadIdx.x;
            if (i < n) c += key_new[i] != key_old[i];
        }
    }
    uint32_t total;
    (void)block_exclusive_scan<kSortBlock>(c, tmp, &total);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSortBlock) void k_split_scatter(
    const uint32_t *__restrict__ key_new, const uint32_t *__restrict__ key_old,
    const int *__restrict__ n_dev, const uint32_t *__restrict__ tile_cnt,
    uint32_t *__restrict__ stay_k, uint32_t *__restrict__ stay_v, uint32_t *__restrict__ mov_k,
    uint32_t *__restrict__ mov_v, int *__restrict__ counts)
{
    constexpr int W = kSortBlock / kWave;
    __shared__ uint32_t tmp[W + 1];
    __shared__ uint32_t wcnt[W];
    const int n = *n_dev;
    const int tile0 = blockIdx.x * kSplitTile;
    const int last_tile = n > 0 ? (n - 1) / kSplitTile : 0;
    if ((int)blockIdx.x > last_tile) return;  // whole block leaves together
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    // movers in earlier tiles (each block sums the counts itself: no waiting)
    uint32_t before = 0;
    for (int t = tid; t < (int)blockIdx.x; t += kSortBlock) before += tile_cnt[t];
    uint32_t tot;
    (void)block_exclusive_scan<kSortBlock>(before, tmp, &tot);
    before = tot;
    // each wave owns a contiguous run of kWave * kSortItems slots
    const int base = tile0 + w * (kWave * kSortItems);
    const uint64_t lt = lanemask_lt();
    uint32_t kn[kSortItems];
    bool mv[kSortItems];
    uint32_t rank[kSortItems];   // movers before this slot within the wave
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const int i = base + k * kWave + lane;
        const bool valid = i < n;
        kn[k] = valid ? key_new[i] : 0u;
        mv[k] = valid && kn[k] != key_old[i];
        const uint64_t bal = __ballot(mv[k]);
        rank[k] = run + (uint32_t)__popcll(bal & lt);
        run += (uint32_t)__popcll(bal);
    }
    if (lane == 0) wcnt[w] = run;
    __syncthreads();
    uint32_t woff = 0;
    for (int ww = 0; ww < w; ++ww) woff += wcnt[ww];
    uint32_t block_movers = 0;
    for (int ww = 0; ww < W; ++ww) block_movers += wcnt[ww];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const int i = base + k * kWave + lane;
        if (i >= n) continue;
        const uint32_t mbefore = before + woff + rank[k];  // movers before slot i
        if (mv[k]) {
            mov_k[mbefore] = kn[k];
            mov_v[mbefore] = (uint32_t)i;
        } else {
            const uint32_t sp = (uint32_t)i - mbefore;
            stay_k[sp] = kn[k];
            stay_v[sp] = (uint32_t)i;
        }
    }
    if ((int)blockIdx.x == last_tile && tid == 0) {
        const int movers = (int)(before + block_movers);
        counts[0] = n - movers;
        counts[1] = movers;
    }
}

// (key, slot) order; slots of the two runs are distinct.
__device__ __forceinline__ bool kv_less(uint32_t ka, uint32_t va, uint32_t kb, uint32_t vb)
{
    return ka < kb || (ka == kb && va < vb);
}

// Merge-path split of every output tile boundary: split[t] = stay elements
// among the first min(t * kMergeTile, total) outputs.
__global__ __launch_bounds__(256) void k_merge_partition(
    const uint32_t *__restrict__ stay_k, const uint32_t *__restrict__ stay_v,
    const uint32_t *__restrict__ mov_k, const uint32_t *__restrict__ mov_v,
    const int *__restrict__ counts, int *__restrict__ split, int n_bounds)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n_bounds) return;
    const int s = counts[0], m = counts[1];
    const long long dd = (long long)t * kMergeTile;
    const int d = (int)(dd < (long long)(s + m) ? dd : (long long)(s + m));
    int lo = d - m > 0 ? d - m : 0, hi = d < s ? d : s;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int j = d - 1 - mid;
        if (kv_less(stay_k[mid], stay_v[mid], mov_k[j], mov_v[j])) lo = mid + 1; else hi = mid;
    }
    split[t] = lo;
}

__global__ __launch_bounds__(kSortBlock) void k_merge(
    const uint32_t *__restrict__ stay_k, const uint32_t *__restrict__ stay_v,
    const uint32_t *__restrict__ mov_k, const uint32_t *__restrict__ mov_v,
    const int *__restrict__ counts, const int *__restrict__ split, uint32_t *__restrict__ out_k,
    uint32_t *__restrict__ out_v)
{
    __shared__ uint32_t sk[kMergeTile];
    __shared__ uint32_t sv[kMergeTile];
    const int s = counts[0], m = counts[1], total = s + m;
    const int d0 = blockIdx.x * kMergeTile;
    if (d0 >= total) return;
    const int d1 = min(d0 + kMergeTile, total);
    const int i0 = split[blockIdx.x], i1 = split[blockIdx.x + 1];
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int na = i1 - i0, nb = j1 - j0;
    for (int q = threadIdx.x; q < na; q += kSortBlock) { sk[q] = stay_k[i0 + q]; sv[q] = stay_v[i0 + q]; }
    for (int q = threadIdx.x; q < nb; q += kSortBlock) { sk[na + q] = mov_k[j0 + q]; sv[na + q] = mov_v[j0 + q]; }
    __syncthreads();
    // this thread's outputs: [diag, diag + kMergeItems) of the tile
    const int diag = min((int)threadIdx.x * kMergeItems, na + nb);
    int lo = diag - nb > 0 ? diag - nb : 0, hi = diag < na ? diag : na;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int j = na + diag - 1 - mid;
        if (kv_less(sk[mid], sv[mid], sk[j], sv[j])) lo = mid + 1; else hi = mid;
    }
    int a = lo, b = na + diag - lo;
    uint32_t rk[kMergeItems], rv[kMergeItems];
#pragma unroll
    for (int q = 0; q < kMergeItems; ++q) {
        const bool take_a = a < na && (b >= na + nb || kv_less(sk[a], sv[a], sk[b], sv[b]));
        const int src = take_a ? a : b;
        rk[q] = src < na + nb ? sk[src] : 0u;
        rv[q] = src < na + nb ? sv[src] : 0u;
        a += take_a;
        b += !take_a;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMergeItems; ++q) {
        const int o = (int)threadIdx.x * kMergeItems + q;
        if (o < na + nb) { sk[o] = rk[q]; sv[o] = rv[q]; }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < na + nb; q += kSortBlock) {
        out_k[d0 + q] = sk[q];
        out_v[d0 + q] = sv[q];
    }
}

size_t resort_scratch_bytes(int n_cap)
{
    const size_t n = (size_t)(n_cap > 0 ? n_cap : 1);
    const size_t tiles = (size_t)div_up((long long)n, kSplitTile) + 1;
    return 8 * n * sizeof(uint32_t) + (tiles + 1) * sizeof(uint32_t) + (tiles + 2) * sizeof(int) +
           4 * sizeof(int) + 12 * 256;
}

void resort_scratch_bind(ResortScratch &r, int n_cap, void *mem)
{
    auto align = [](char *p) { return (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255); };
    const size_t n = (size_t)(n_cap > 0 ? n_cap : 1);
    r.n_cap = (int)n;
    r.tiles = div_up((long long)n, kSplitTile);
    char *p = align((char *)mem);
    uint32_t **arrays[8] = {&r.stay_k, &r.stay_v, &r.mov_k, &r.mov_v,
                            &r.movs_k, &r.movs_v, &r.alt_k, &r.alt_v};
    for (auto a : arrays) { *a = (uint32_t *)p; p = align(p + n * sizeof(uint32_t)); }
    r.tile_cnt = (uint32_t *)p; p = align(p + (size_t)(r.tiles + 1) * sizeof(uint32_t));
    r.split = (int *)p; p = align(p + (size_t)(r.tiles + 2) * sizeof(int));
    r.counts = (int *)p;
}

void resort_incremental(const uint32_t *key_new, const uint32_t *key_old, const int *n_dev,
                        int key_bits, const ResortScratch &r, const SortScratch &s,
                        uint32_t *out_k, uint32_t *out_v, hipStream_t stream, Profiler *prof)
{
    const double n = (double)r.n_cap;
    DSMC_LAUNCH(prof, "resort_count", n * 8.0, k_split_count, r.tiles, kSortBlock, stream,
                key_new, key_old, n_dev, r.tile_cnt);
    DSMC_LAUNCH(prof, "resort_split", n * 16.0, k_split_scatter, r.tiles, kSortBlock, stream,
                key_new, key_old, n_dev, r.tile_cnt, r.stay_k, r.stay_v, r.mov_k, r.mov_v,
                r.counts);
    radix_sort(r.mov_k, r.mov_v, r.alt_k, r.alt_v, r.movs_k, r.movs_v, r.counts + 1, key_bits, s,
               stream, prof);
    const int bounds = r.tiles + 1;  // merge tiles == split tiles (total <= n_cap)
    DSMC_LAUNCH(prof, "resort_partition", (double)bounds * 64.0, k_merge_partition,
                div_up(bounds, 256), 256, stream, r.stay_k, r.stay_v, r.movs_k, r.movs_v,
                r.counts, r.split, bounds);
    DSMC_LAUNCH(prof, "resort_merge", n * 16.0, k_merge, r.tiles, kSortBlock, stream, r.stay_k,
                r.stay_v, r.movs_k, r.movs_v, r.counts, r.split, out_k, out_v);
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

// Streaming copy (16-byte vectors, grid-stride): the bandwidth reference used
// to calibrate the HBM-traffic counters.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst, long long n16)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n16; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

void stream_copy(void *dst, const void *src, size_t bytes, hipStream_t stream)
{
    const long long n16 = (long long)(bytes / 16);
    if (n16 <= 0) return;
    const int grid = (int)std::min<long long>((n16 + 255) / 256, 256LL * 64);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, stream, (const uint4 *)src,
                       (uint4 *)dst, n16);
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
