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
    __syncthreads();
    for (int d = tid; d < kRadix; d += kSortBlock) hist[(size_t)d * nb + blockIdx.x] = h[d];
}

// One block per digit: exclusive scan of that digit's per-block counts.
__global__ __launch_bounds__(256) void k_radix_scan(uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ totals, int nb)
{
    __shared__ uint32_t tmp[256 / kWave + 1];
    uint32_t *row = hist + (size_t)blockIdx.x * nb;
    const int per = (nb + 255) / 256;
    const int b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i) {
        const int j = b0 + i;
        if (j < nb) s += row[j];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan<256>(s, tmp, &total);
    for (int i = 0; i < per; ++i) {
        const int j = b0 + i;
        if (j < nb) { const uint32_t c = row[j]; row[j] = ex; ex += c; }
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
                    stream, s.hist, s.totals, s.nb);
        DSMC_LAUNCH(prof, "radix_downsweep", bytes_pass, k_radix_downsweep, s.nb, kSortBlock,
                    stream, sk, sv, dk, dv, n_dev, shift, s.hist, s.totals, s.nb);
        sk = dk;
        sv = dv;
    }
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
