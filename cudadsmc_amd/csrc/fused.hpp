// fused.hpp -- kernel parameter structs and host launchers of the FP32
// precision form (Kahan-compensated state, any number of traps).
#pragma once

#include "../../include/dsmc.h"
#include "dsmc_device.hpp"
#include "dsmc_internal.hpp"
#include "init_device.hpp"
#include "label_bits.hpp"
#include "plain.hpp"

namespace dsmc {

// FP32 state of a slot: each coordinate is a float plus the Kahan remainder
// that the previous updates could not represent (position p + rp, velocity
// v + rv).
struct FState {
    // in the order of the record planes below, so that each plane is one
    // 16-byte piece of the struct
    float px, py, pz, rpx;
    float vx, vy, vz, rvx;
    float rpy, rpz, rvy, rvz;
};

// The twelve floats of every slot are stored in tiles of 64 slots; a tile is
// three planes of 64 four-float records,
//   plane 0: (px, py, pz, rpx)   plane 1: (vx, vy, vz, rvx)
//   plane 2: (rpy, rpz, rvy, rvz)
// so a slot's state is three 16-byte accesses, a wave's access to a plane is
// 1 KiB contiguous, and the velocity (what a collision trial reads) is one
// 16-byte record; an accepted collision also writes the upper half of plane 2.
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kRecTile = 64;
constexpr int kRecPlanes = 3;
inline size_t rec_elems(size_t n) { return (n + kRecTile - 1) / kRecTile * kRecTile * kRecPlanes; }
__host__ __device__ __forceinline__ size_t rec_index(uint32_t j)
{
    return (size_t)(j / kRecTile) * (kRecTile * kRecPlanes) + (j % kRecTile);
}

// Device pointers of one FP32 state buffer.
struct FusedPtrs {
    f4    *r;        // per-slot state records (layout above), with the labels
    float *s[3];     // Bloch spin vector (spin runs only)
};

__device__ __forceinline__ void rec_unpack(const f4 &p0, const f4 &p1, const f4 &p2, FState &a)
{
    a.px = p0.x; a.py = p0.y; a.pz = p0.z; a.rpx = rem_value(p0.w);
    a.vx = p1.x; a.vy = p1.y; a.vz = p1.z; a.rvx = p1.w;
    a.rpy = rem_value(p2.x); a.rpz = rem_value(p2.y); a.rvy = p2.z; a.rvz = p2.w;
}

// the state of slot j (the position remainders' values), and its label
__device__ __forceinline__ void rec_load(const f4 *r, uint32_t j, FState &a, int *label = nullptr)
{
    const f4 *q = r + rec_index(j);
    const f4 p0 = q[0], p1 = q[kRecTile], p2 = q[2 * kRecTile];
    rec_unpack(p0, p1, p2, a);
    if (label) *label = rec_label(p0.w, p2.x, p2.y);
}

// Non-temporal: for the move's streaming read of the whole state (it is the
// last reader of the source buffer in a step; the cache is left to the
// re-sort's and the collide's data).
__device__ __forceinline__ void rec_load_stream(const f4 *r, uint32_t j, FState &a, int &label)
{
    const f4 *q = r + rec_index(j);
    const f4 p0 = __builtin_nontemporal_load(q), p1 = __builtin_nontemporal_load(q + kRecTile),
             p2 = __builtin_nontemporal_load(q + 2 * kRecTile);
    rec_unpack(p0, p1, p2, a);
    label = rec_label(p0.w, p2.x, p2.y);
}

// The move's write of the whole state: the position and remainder planes
// non-temporal (only the next move reads them all), the velocity plane
// through the cache (the next step's collide reads it for its trial pairs).
// The position remainders are already rounded (rem_round).
__device__ __forceinline__ void rec_store_stream(f4 *r, uint32_t j, const FState &a, int label)
{
    f4 *q = r + rec_index(j);
    __builtin_nontemporal_store(f4{ a.px, a.py, a.pz, rem_put(a.rpx, label_part(label, 0)) }, q);
    q[kRecTile] = f4{ a.vx, a.vy, a.vz, a.rvx };
    __builtin_nontemporal_store(f4{ rem_put(a.rpy, label_part(label, 1)),
                                    rem_put(a.rpz, label_part(label, 2)), a.rvy, a.rvz },
                                q + 2 * kRecTile);
}

__device__ __forceinline__ void rec_store(f4 *r, uint32_t j, const FState &a, int label)
{
    f4 *q = r + rec_index(j);
    q[0] = f4{ a.px, a.py, a.pz, rem_tag(a.rpx, label_part(label, 0)) };
    q[kRecTile] = f4{ a.vx, a.vy, a.vz, a.rvx };
    q[2 * kRecTile] = f4{ rem_tag(a.rpy, label_part(label, 1)), rem_tag(a.rpz, label_part(label, 2)),
                          a.rvy, a.rvz };
}

// Observable accumulators per trap (doubles).
enum { kAccEk = 0, kAccEp, kAccV2x, kAccV2y, kAccV2z, kAccEevap, kAccNevap, kAccN, kAccK };
// the move's records leave kAccN zero; the finalize sums there the atoms
// removed in the slots before the trap
constexpr int kAccDead = kAccN;

struct MoveParams {
    double dt;
    double mass;
    double gs;
    float  f_dt;
    float  f_half_mass;
    int    spin;
    int    substeps;
    int    interp;
    int    spin_loss;
    int    energy_cut;
    int    n_traps;
    int    obs_cap;      // records in the observables ring (per trap)
    uint32_t dead_key;
    // when count_moved is set: per move block, the atoms whose new key differs
    // from their key before the move (key_prev: the last sort's sorted keys,
    // indexed by output slot) and the removed atoms, and the same atoms as one
    // bit per slot (64-slot words), for the incremental re-sort of the next step;
    // key_samp / key_tsamp receive key_prev at every 64th and every
    // kResortTile-th slot below n (the re-sort's search index)
    int count_moved;
    const uint32_t *key_prev;
    uint32_t *blk_moved;
    uint32_t *blk_dead;
    uint64_t *mask_moved;
    uint64_t *mask_dead;
    uint32_t *key_samp;
    uint32_t *key_tsamp;
};

struct TableParams {
    int n_traps;
    uint32_t dead_key;
};

int  fused_max_traps();
void launch_init_fused(const TrapDev *traps, const int *trap_start, int n_traps,
                       const InitParams &ip, const FusedPtrs &s, int n, hipStream_t st,
                       Profiler *prof);
void launch_index_fused(const TrapDev *traps, const int *trap_start, int n_traps,
                        const FusedPtrs &s, uint32_t *key, const int *n_dev, int n_cap,
                        uint32_t dead_key, bool keep_dead, hipStream_t st, Profiler *prof);
void launch_table_fused(const uint32_t *key, const int *key_base, const TableParams &tp,
                        const int *n_in, int *n_out, int *trap_start, int n_cap, hipStream_t st,
                        Profiler *prof);
void launch_cell_range_fused(const uint32_t *key, const int *n_dev, uint32_t dead_key,
                             uint32_t *cell_start, uint32_t *cell_end, int n_cap, hipStream_t st);
void launch_collide_fused(const TrapDev *traps, const int *key_base, int n_traps,
                          const CollideParams &cp, const uint32_t *key, const uint32_t *perm,
                          const int *n_dev, const FusedPtrs &s, float *sigvrmax,
                          unsigned long long *coll_accum, const long long *step_dev, int n_cap,
                          hipStream_t st, Profiler *prof);
// Per-block observable records of the fused move (see k_move_fused).
constexpr int kMaxFinalizeP = 64;
struct ObsRecords {
    int    *trap;      // first trap of each move block            [stride]
    double *sum;       // its sums, kAccK arrays of length stride   [kAccK * stride]
    int     stride;
    double *acc_over;  // boundary terms per trap                   [T * kAccK]
    double *lvl2;      // finalize partials                         [T * kMaxFinalizeP * kAccK]
    unsigned int *cnt; // finalize counters per trap                [T]
};
int  obs_record_stride(int n_cap);
// Collisions inside the fused step kernel (see launch_move_fused).
struct StepCollide {
    CollideParams cp;
    float *sigvrmax;
    unsigned long long *coll_accum;
    int dry;   // compute but write neither sigma v_r,max nor counts (autotuner trials)
};
// block: threads per block (64 .. 1024); apt: atoms per thread (1 or 2).  A
// block moves block * apt consecutive slots and leaves one observables record
// and one pair of re-sort counts.  With coll (the fused step; needs the
// sorted keys in mp.key_prev), each wave first runs the collisions of the
// cells of its 64 * apt slots; the oversized cells across those spans'
// boundaries must have been run by launch_collide_edges with move_block =
// 64 * apt.
void launch_move_fused(int block, int apt, const TrapDev *traps, const MoveParams &mp,
                       const FusedPtrs &src, const FusedPtrs &dst, const uint32_t *perm,
                       const int *trap_start, const int *n_dev, uint32_t *key_out,
                       float *evap_pos, float *evap_vel, const ObsRecords &rec,
                       const long long *step_dev, int n_cap, hipStream_t st, Profiler *prof,
                       const StepCollide *coll = nullptr);
// The collisions of the oversized cells across the boundaries between spans
// of move_block sorted slots (the edge pass of the fused step).
void launch_collide_edges(const TrapDev *traps, const int *key_base, int n_traps,
                          const CollideParams &cp, const uint32_t *key, const uint32_t *perm,
                          const int *n_dev, const FusedPtrs &s, float *sigvrmax,
                          unsigned long long *coll_accum, const long long *step_dev, int move_block,
                          int n_cap, hipStream_t st, Profiler *prof);
// When n_next is set (the move followed a sort), the last block also writes
// the next step's trap ranges (trap_start) and alive count (*n_next): each
// trap loses its atoms removed in this step, which the next sort puts last.
// trap_next / n_next (when not null): the trap ranges and alive count after
// the next sort, from the move's per-block removal counts and masks.  step
// is read from step_cur; step + 1 is written to step_next.
// prep.blocks > 0: the same launch also runs the next incremental re-sort's
// first phase on the move's masks.
void launch_obs_finalize(int n_traps, int max_trap_atoms, int move_block, double mass,
                         const int *trap_start, int *trap_next, const uint32_t *blk_dead,
                         const uint64_t *mask_dead, const ObsRecords &rec,
                         unsigned long long *coll_accum, const long long *step_cur,
                         long long *step_next, dsmc_observables *obs_log, int obs_cap,
                         int *n_next, const ResortPrepArgs &prep, hipStream_t st,
                         Profiler *prof);
// Observables of the current state per trap into acc (kAccK doubles per
// trap; kAccN = alive atoms), skipping slots with the dead key; a fixed-order
// reduction.  first / last: n_traps ints of scratch each.
void launch_observe_fused(const TrapDev *traps, const int *key_base, int n_traps, double mass,
                          const FusedPtrs &s, const uint32_t *key, const int *n_dev, int n_cap,
                          uint32_t dead_key, int *first, int *last, double *acc, hipStream_t st);
void launch_gather_fused(const FusedPtrs &dst, const FusedPtrs &src, const uint32_t *map,
                         const int *n_dev, int n_cap, int spin, hipStream_t st, Profiler *prof);
double move_fused_bytes_per_atom(int spin);

}  // namespace dsmc
