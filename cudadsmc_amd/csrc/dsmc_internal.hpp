// dsmc_internal.hpp -- host-side internals shared by the library's
// translation units: launch bookkeeping (optional per-kernel hipEvent timing),
// error propagation, and the device-primitive launchers (radix sort, scan,
// compaction) that both numerical forms use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace dsmc {

// Thrown inside the library only; every extern "C" entry point catches it and
// converts it into a dsmc_status plus a message.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void hip_check(hipError_t e, const char *what, const char *file, int line);
#define DSMC_HIP(call) ::dsmc::hip_check((call), #call, __FILE__, __LINE__)

// Per-kernel timing with hipEvents on the launch stream.
class Profiler {
public:
    bool enabled = false;
    void begin(hipStream_t s, const char *name, double bytes);
    void end(hipStream_t s);
    // Resolve pending events (synchronises) and fold them into the totals.
    void collect();
    void reset();
    struct Stat { long long launches = 0; double total_ms = 0.0; double bytes = 0.0; };
    const std::map<std::string, Stat> &stats() const { return stats_; }
    ~Profiler();

private:
    struct Rec { std::string name; double bytes; hipEvent_t b, e; };
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, Stat> stats_;
    hipEvent_t take();
};

// Launch `kernel` with optional profiling.  `bytes` is the algorithmic byte
// count of this launch (used only for reporting).
#define DSMC_LAUNCH(prof, name, bytes, kernel, grid, block, stream, ...)            \
    do {                                                                            \
        ::dsmc::Profiler *p_ = (prof);                                              \
        if (p_ && p_->enabled) p_->begin((stream), (name), (double)(bytes));        \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (stream), __VA_ARGS__); \
        DSMC_HIP(hipGetLastError());                                                \
        if (p_ && p_->enabled) p_->end((stream));                                   \
    } while (0)

// ---------------------------------------------------------------------------
// Device primitives (primitives.hip).
// ---------------------------------------------------------------------------
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Scratch for a radix sort of up to n_cap pairs.
struct SortScratch {
    int n_cap = 0;
    int nb = 0;                   // blocks per pass
    uint32_t *hist = nullptr;     // kRadix * nb
    uint32_t *totals = nullptr;   // kRadix
};
size_t sort_scratch_bytes(int n_cap);
void sort_scratch_bind(SortScratch &s, int n_cap, void *mem);

// Stable LSD radix sort of keys[0..n) with values (identity slot indices when
// vals == nullptr).  n is read from device memory.  The sorted result lands in
// (keys_out, vals_out); ping-pong through (keys_alt, vals_alt) for the
// intermediate passes.  key_bits selects the number of 8-bit passes.
void radix_sort(const uint32_t *keys, const uint32_t *vals, uint32_t *keys_alt,
                uint32_t *vals_alt, uint32_t *keys_out, uint32_t *vals_out,
                const int *n_dev, int key_bits, const SortScratch &s,
                hipStream_t stream, Profiler *prof);

// Incremental re-sort (see primitives.hip): the stable sort of key_new[0..n)
// by (key, slot), for slots that are in stable sorted order of key_old, into
// out (which must not alias key_old).  dead_key sorts last.  Slots are
// grouped in tiles of kResortTile.
constexpr int kResortTile = 2048;
struct ResortScratch {
    int n_cap = 0, tiles = 0;
    uint32_t *mov_k = nullptr, *mov_v = nullptr;
    uint32_t *mov_q = nullptr, *mov_r = nullptr;
    uint32_t *bk_k = nullptr, *bk_v = nullptr, *bk_q = nullptr;   // buckets' fixed rooms
    uint32_t *srt_k = nullptr, *srt_v = nullptr, *srt_q = nullptr;
    uint32_t *alt_k = nullptr, *alt_v = nullptr, *spare = nullptr;
    uint32_t *tile_cnt = nullptr, *tile_dead = nullptr, *bcnt = nullptr;
    int *counts = nullptr;       // [0] stays, [1] movers, [2] removed, [3] bucket overflow
};
// Arguments of the re-sort's first phase when another kernel runs it (the
// fused finalize); blocks = 0: not run there.
struct ResortPrepArgs {
    int blocks = 0;
    int per_tile = 0;              // move blocks per tile
    const int *n_dev = nullptr;    // slots the move wrote
    const uint32_t *blk_moved = nullptr, *blk_dead = nullptr;
    const uint64_t *mask_m = nullptr, *mask_d = nullptr;
    uint32_t *pre_m = nullptr, *pre_d = nullptr, *mov_v = nullptr, *bcnt = nullptr;
    int *counts = nullptr;
};
// The first phase of an incremental re-sort of n_dev slots that follows a
// move of move_block-slot blocks.
ResortPrepArgs resort_prep_args(const ResortScratch &r, const int *n_dev, int move_block,
                                const uint32_t *blk_moved, const uint32_t *blk_dead,
                                const uint64_t *mask_m, const uint64_t *mask_d);
size_t resort_scratch_bytes(int n_cap);
void resort_scratch_bind(ResortScratch &r, int n_cap, void *mem);
bool resort_supported(int n_cap);
// 64-bit words of each of the mover and removal masks for n_cap slots.
size_t resort_mask_words(int n_cap);
// blk_moved / blk_dead: per block of move_block slots, the atoms whose key
// changed (not removed) and the removed atoms; mask_m / mask_d: the same
// atoms as one bit per slot; key_samp / key_tsamp: key_old[64 w] for every
// 64-slot word w and key_old[kResortTile t] for every tile t below n (all
// written by the fused move).  prepped: the first phase (resort_prep_block)
// already ran on these masks (in the fused finalize).
void resort_incremental(const uint32_t *key_new, const uint32_t *key_old, const int *n_dev,
                        uint32_t dead_key, const uint32_t *blk_moved, const uint32_t *blk_dead,
                        const uint64_t *mask_m, const uint64_t *mask_d, const uint32_t *key_samp,
                        const uint32_t *key_tsamp, int move_block, bool prepped,
                        const ResortScratch &r, uint32_t *out_k, uint32_t *out_v,
                        hipStream_t stream, Profiler *prof);

// Exclusive scan of n ints (n known on the host); partial holds div_up(n, tile).
void exclusive_scan(const int *in, int *out, int n, int *partial, hipStream_t stream,
                    Profiler *prof);
inline int scan_partials(int n) { return div_up(n > 0 ? n : 1, kScanTile) + 1; }

// Stable compaction map: for stencil[i] == 0 (kept), map[j] = i for the j-th
// kept element; *n_out (device) = number kept.  n read from device memory;
// n_cap bounds the grid.  partial holds div_up(n_cap, tile) + 1 ints.
void compaction_map(const int *stencil, const int *n_dev, int n_cap, uint32_t *map,
                    int *n_out_dev, int *partial, hipStream_t stream, Profiler *prof);

// dst[i] = src[map[i]] for i < *n_dev.
void gather_u32(uint32_t *dst, const uint32_t *src, const uint32_t *map, const int *n_dev,
                int n_cap, hipStream_t stream, Profiler *prof);

void fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t stream, Profiler *prof);
void fill_f64(double *p, double v, long long n, hipStream_t stream, Profiler *prof);

// Streaming device-to-device copy of bytes (multiple of 16) for counter
// calibration.
void stream_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);
// Gather of n slots of the FP32 state's tiled records (48 bytes a slot, in
// whole 64-slot tiles) through map, for counter calibration.
void gather_copy(void *dst, const void *src, const uint32_t *map, int n, hipStream_t stream);

void philox_fill(uint32_t *out, int n, const uint32_t ctr[4], const uint32_t key[2],
                 hipStream_t stream);

}  // namespace dsmc
