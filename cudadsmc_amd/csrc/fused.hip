// fused.hip -- FP32 precision form of the DSMC step for any number of
// independent traps, laid out as one flat work list.
//
// Layout.  All traps share one set of structure-of-arrays FP32 buffers; atoms
// are kept trap-major and, within a trap, in cell order.  A composite cell key
// key_base[trap] + local_cell (local overflow cell = n_cells of the trap; one
// global "dead" key after every trap) makes one radix sort order all traps at
// once and pushes evaporated atoms to the end, so the sort is also the stream
// compaction and the alive count never has to visit the host.
//
// Step (one motion step, lpc = 1):
//   sort(keys)           -> sorted keys + permutation: incrementally from the
//                           last move's mover masks (k_ins_locate, k_ins_merge
//                           in primitives.hip) or by a full radix sort, after
//                           which k_table_fused finds the trap ranges and the
//                           alive count
//   k_collide_fused      -> each wave scans 512 sorted slots for cells of two
//                           or more atoms and runs their trials, velocities
//                           read through the permutation (no reorder pass)
//   k_move_fused         -> gathers state through the permutation into the
//                           other buffer (so the state ends up sorted), spin
//                           substeps, Kahan-compensated Verlet, evaporation,
//                           next cell key, FP64 observables per block
//   k_obs_finalize       -> per-trap sums in fixed order and the log record,
//                           the next trap ranges and alive count, and the
//                           incremental re-sort's first phase
// Every kernel reads the alive count and the step number from device memory,
// so the loop can be captured once and replayed as a HIP graph.
#include "fused.hpp"
#include "resort_device.hpp"
#include "sincos_f32.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace dsmc {

constexpr int kMaxTrapsLds = 8192;
int fused_max_traps() { return kMaxTrapsLds; }

constexpr int kFBlock = 256;

static int fgrid(long long n, int bs) { return div_up(n > 0 ? n : 1, bs); }

// Largest t in [0, n) with arr[t] <= v (arr ascending, arr[0] <= v).
__device__ __forceinline__ int last_le(const int *arr, int n, int v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (arr[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Largest t in [lo, hi] with arr[t] <= v, given arr[lo] <= v.
__device__ __forceinline__ int last_le_in(const int *arr, int lo, int hi, int v)
{
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (arr[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void load_table(int *lds, const int *src, int cnt)
{
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) lds[j] = src[j];
    __syncthreads();
}

// s + a b with the running remainder c: y = a b + c in one rounding (fma),
// then an error-free TwoSum of s + y (Knuth), so the remainder stays exact
// whatever the magnitudes.
__device__ __forceinline__ void kahan_fma(float &s, float &c, float a, float b)
{
    const float y = fmaf(a, b, c);
    const float t = s + y;
    const float bp = t - s;
    const float ap = t - bp;
    c = (s - ap) + (y - bp);
    s = t;
}

__device__ __forceinline__ void split_df(double v, float &hi, float &lo)
{
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

// ---------------------------------------------------------------------------
// Initial distribution (setUp.cu:129-165) into the FP32 state: the FP64 sample
// is stored as float + float remainder.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFBlock) void k_init_fused(const TrapDev *__restrict__ traps,
                                                        const int *__restrict__ trap_start,
                                                        int T, InitParams ip, FusedPtrs s, int n)
{
    extern __shared__ int ts[];
    load_table(ts, trap_start, T + 1);
    const int a = blockIdx.x * kFBlock + threadIdx.x;
    if (a >= n) return;
    const int t = last_le(ts, T, a);
    const TrapDev &tr = traps[t];
    double px, py, pz, vx, vy, vz;
    init_atom(tr, ip, (uint32_t)(a - ts[t]), px, py, pz, vx, vy, vz);
    FState st;
    split_df(px, st.px, st.rpx);
    split_df(py, st.py, st.rpy);
    split_df(pz, st.pz, st.rpz);
    split_df(vx, st.vx, st.rvx);
    split_df(vy, st.vy, st.rvy);
    split_df(vz, st.vz, st.rvz);
    rec_store(s.r, (uint32_t)a, st, a);
    if (ip.spin) {
        D3 Bn;
        (void)field_dir(tr, px, py, pz, Bn);
        s.s[0][a] = (float)Bn.x; s.s[1][a] = (float)Bn.y; s.s[2][a] = (float)Bn.z;
    }
}

// indexAtoms for the FP32 state: the FP64 index of p + cp.  With keep_dead, a
// slot that already holds the dead key (an atom removed by the last move,
// which stays in place until the next sort drops it) keeps it.
__global__ __launch_bounds__(kFBlock) void k_index_fused(const TrapDev *__restrict__ traps,
                                                         const int *__restrict__ trap_start,
                                                         int T, FusedPtrs s,
                                                         uint32_t *__restrict__ key,
                                                         const int *__restrict__ n_dev,
                                                         uint32_t dead_key, int keep_dead)
{
    extern __shared__ int ts[];
    load_table(ts, trap_start, T + 1);
    const int i = blockIdx.x * kFBlock + threadIdx.x;
    if (i >= *n_dev) return;
    if (keep_dead && key[i] == dead_key) return;
    const int t = last_le(ts, T, i);
    const TrapDev &tr = traps[t];
    const f4 *q = s.r + rec_index((uint32_t)i);
    const f4 p0 = q[0], p2 = q[2 * kRecTile];
    const double X = (double)p0.x + (double)rem_value(p0.w);
    const double Y = (double)p0.y + (double)rem_value(p2.x);
    const double Z = (double)p0.z + (double)rem_value(p2.y);
    key[i] = (uint32_t)(tr.key_base + cell_of_fast(tr, X, Y, Z));
}

// ---------------------------------------------------------------------------
// Cell table from sorted keys: start/end of every occupied key, first slot of
// every trap, and the alive count (= first dead slot).  n_in + 1 threads: the
// extra one closes the last trap ranges.
// ---------------------------------------------------------------------------
constexpr int kTableItems = 16;  // slots per thread of the trap-boundary scan

__global__ __launch_bounds__(kFBlock) void k_table_fused(const uint32_t *__restrict__ key,
                                                         const int *__restrict__ key_base,
                                                         TableParams tp,
                                                         const int *__restrict__ n_in_dev,
                                                         int *__restrict__ n_out_dev,
                                                         int *__restrict__ trap_start)
{
    extern __shared__ int kb[];
    __shared__ int blk_traps[2];
    const int T = tp.n_traps;
    load_table(kb, key_base, T + 1);
    const int n = *n_in_dev;
    const int b0 = blockIdx.x * (kFBlock * kTableItems);
    if (b0 > n) return;
    const int b1 = min(n + 1, b0 + kFBlock * kTableItems);  // slots b0..b1-1, slot n closes
    // A range whose keys all lie in one trap holds no trap boundary except
    // possibly at its first slot: then only that slot (and slot n) is examined.
    if (threadIdx.x == 0) {
        const int last = min(n, b1) - 1;
        blk_traps[0] = b0 < n ? last_le(kb, T + 1, (int)key[b0]) : T;
        blk_traps[1] = last >= b0 ? last_le(kb, T + 1, (int)key[last]) : T;
    }
    __syncthreads();
    if (blk_traps[0] == blk_traps[1]) {
        // no trap boundary inside: only the range's first slot (a boundary
        // when the previous slot belongs to another trap) and slot n
        if (threadIdx.x == 0) {
            const int i = b0;
            const uint32_t k = (i < n) ? key[i] : tp.dead_key;
            const int cur = (i == n) ? T : last_le(kb, T + 1, (int)k);
            const int prev = (i == 0) ? -1 : last_le(kb, T + 1, (int)key[i - 1]);
            for (int t = prev + 1; t <= cur; ++t) trap_start[t] = i;
            if (cur != prev && prev < T && cur == T) *n_out_dev = i;
        }
        if (threadIdx.x == 1 && n >= b0 + 1 && n < b1) {
            const int prev = last_le(kb, T + 1, (int)key[n - 1]);
            for (int t = prev + 1; t <= T; ++t) trap_start[t] = n;
            if (prev < T) *n_out_dev = n;
        }
        return;
    }
    // Several traps: every slot is examined.  All keys are loaded first
    // (independent loads in flight together), then compared; the trap of a
    // key inside the range is searched among the range's own traps only.
    const int t0 = blk_traps[0], t1 = blk_traps[1];
    uint32_t kc[kTableItems], kp[kTableItems];
#pragma unroll
    for (int it = 0; it < kTableItems; ++it) {
        const int i = b0 + it * kFBlock + threadIdx.x;
        kc[it] = (i < n) ? key[i] : tp.dead_key;
        kp[it] = (i > 0 && i - 1 < n) ? key[i - 1] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int it = 0; it < kTableItems; ++it) {
        const int i = b0 + it * kFBlock + threadIdx.x;
        if (i >= b1) continue;
        const uint32_t k = kc[it];
        if (i > 0 && i < n && kp[it] == k) continue;
        const int cur = (i == n) ? T : last_le_in(kb, t0, t1, (int)k);
        const int prev = (i == 0) ? -1
                         : (i == b0) ? last_le(kb, T + 1, (int)kp[it])
                                     : last_le_in(kb, t0, t1, (int)kp[it]);
        if (cur != prev) {
            for (int t = prev + 1; t <= cur; ++t) trap_start[t] = i;
            if (prev < T && cur == T) *n_out_dev = i;
        }
    }
}

// Start and end slot of every occupied key (on demand: the output surface of
// the cell table; the step itself does not need it).
__global__ __launch_bounds__(kFBlock) void k_cell_range_fused(const uint32_t *__restrict__ key,
                                                              const int *__restrict__ n_dev,
                                                              uint32_t dead_key,
                                                              uint32_t *__restrict__ cell_start,
                                                              uint32_t *__restrict__ cell_end)
{
    const int n = *n_dev;
    const int i = blockIdx.x * kFBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    if (k == dead_key) return;
    if (i == 0 || key[i - 1] != k) cell_start[k] = (uint32_t)i;
    if (i == n - 1 || key[i + 1] != k) cell_end[k] = (uint32_t)(i + 1);
}

// ---------------------------------------------------------------------------
// collide (collisions.cu:291-445) over occupied cells: the thread of the first
// slot of each cell runs that cell's trials in sequence.  Trials (pair choice,
// |v_r|, acceptance) in FP32; an accepted collision updates the pair in FP64
// from value + remainder and stores the result back as float + remainder, so
// momentum and energy are conserved to FP64 rounding.
// ---------------------------------------------------------------------------
// Cell state for the trials of one cell (collisions.cu:306-322).
struct CellJob {
    int      i;       // first sorted slot
    int      n;       // atoms in the cell
    int      t;       // trap
    uint32_t cell;    // local cell (Philox counter word 0)
    uint32_t k;       // composite key (sigma v_r,max index)
    int      ncol;    // trials
    float    pc;      // acceptance probability per unit |v_r|
    double   smax;    // sigma v_r,max at entry
};

// Cell of atoms [i, end) with composite key k in trap t (parameters tr);
// smax = its sigma v_r,max.
__device__ __forceinline__ bool cell_setup(const TrapDev &tr, int t, uint32_t k,
                                           const CollideParams &cp, double smax, int i, int end,
                                           CellJob &c)
{
    c.i = i;
    c.n = end - i;
    c.ncol = 0;
    if (c.n < 2) return false;
    c.k = k;
    c.t = t;
    c.cell = c.k - (uint32_t)tr.key_base;
    if ((int)c.cell >= tr.n_cells) return false;  // overflow cell: no collisions
    c.smax = smax;
    const double Mc = 0.5 * (c.n - 1) * c.n;
    if (cp.pair_mode) {
        // reference pair mode: collisions.cu:318-322 including the truncation
        // of Mc * lambda
        const double lpc = cp.lpc;
        const double X = Mc * tr.alpha * lpc * cp.dt * c.smax / tr.cell_vol;
        const double lambda = ceil(X) / Mc;
        c.ncol = (int)(Mc * lambda);
        c.pc = (float)(tr.alpha * lpc * cp.dt / tr.cell_vol * cp.sigma / lambda);
    } else {
        // ceil(X) trials, acceptance alpha lpc dt sigma v_r / (Vc lambda) with
        // lambda = ceil(X) / Mc; the per-trap factor is precomputed
        const double cX = ceil(Mc * tr.coll_c1 * c.smax);
        c.ncol = (int)cX;
        c.pc = (float)(tr.coll_c1s * Mc / cX);
    }
    return c.ncol > 0;
}

// Pair (slot offsets within the cell) and acceptance uniform of trial tr
// (collisions.cu:400-425).
__device__ __forceinline__ void trial_draw(const CollideParams &cp, uint32_t stream, uint32_t cell,
                                           uint32_t step, int tr, int n, int &ia, int &ib,
                                           float &u_acc)
{
    ia = 0; ib = 1;
    if (cp.pair_mode == 0) {
        const U32x4 o = draw(cp.seed, stream, cell, step, (uint32_t)tr, kPurposeCollide, 0);
        ia = (int)uint_below(o.x, (uint32_t)n);
        ib = (int)uint_below(o.y, (uint32_t)(n - 1));
        ib += (ib >= ia);
        u_acc = u01f(o.z);
    } else {
        if (n > 2) {
            ia = 0; ib = 0;
            for (uint32_t q = 0; q < 64 && ia == ib; ++q) {
                const U32x4 o = draw(cp.seed, stream, cell, step, (uint32_t)tr, kPurposeCollide, 16 + q);
                ia = (int)floor(u01(o.x, o.y) * (n - 1));
                ib = (int)floor(u01(o.z, o.w) * (n - 1));
            }
            if (ia == ib) { ia = 0; ib = 1; }
        }
        const U32x4 o = draw(cp.seed, stream, cell, step, (uint32_t)tr, kPurposeCollide, 1);
        u_acc = u01f(o.x);
    }
}

typedef float f2 __attribute__((ext_vector_type(2)));

// (vx, vy, vz, rvx) of slot j: plane 1; (rvy, rvz): upper half of plane 2.
__device__ __forceinline__ f4 *vel_rec(f4 *r, uint32_t j) { return r + rec_index(j) + kRecTile; }
__device__ __forceinline__ f2 *vel_rem_yz(f4 *r, uint32_t j)
{
    return (f2 *)(r + rec_index(j) + 2 * kRecTile) + 1;
}

__device__ __forceinline__ float rel_speed_v(const f4 va, const f4 vb)
{
    const float dx = va.x - vb.x;
    const float dy = va.y - vb.y;
    const float dz = va.z - vb.z;
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// Isotropic unit vector for the FP32 form: Marsaglia's (1972) disc method in
// FP32 with two attempts per Philox block (acceptance pi/4 each), then
// normalised in FP64 so that |n| = 1 to FP64 rounding and the collision
// conserves energy to FP64 rounding.  Falls back to +z after 32 rejections
// (probability 0.215^32).
__device__ __forceinline__ void sphere_point_f(unsigned long long seed, uint32_t stream,
                                               uint32_t cell, uint32_t step, uint32_t tr,
                                               double &nx, double &ny, double &nz)
{
    float fx = 0.0f, fy = 0.0f, fz = 1.0f;
#pragma unroll 1
    for (uint32_t k = 0; k < 16; ++k) {
        const U32x4 o = draw(seed, stream, cell, step, tr, kPurposeCollide, 2 + k);
        float u1 = 2.0f * u01f(o.x) - 1.0f, u2 = 2.0f * u01f(o.y) - 1.0f;
        float q = u1 * u1 + u2 * u2;
        if (!(q < 1.0f)) {
            u1 = 2.0f * u01f(o.z) - 1.0f;
            u2 = 2.0f * u01f(o.w) - 1.0f;
            q = u1 * u1 + u2 * u2;
        }
        if (q < 1.0f) {
            const float r = 2.0f * sqrtf(1.0f - q);
            fx = u1 * r; fy = u2 * r; fz = 1.0f - 2.0f * q;
            break;
        }
    }
    nx = fx; ny = fy; nz = fz;
    const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= inv; ny *= inv; nz *= inv;
}

// Accepted collision of atoms a and b of trial tr, in FP64 from value +
// remainder, returned as float + remainder (collisions.cu:426-436): va, vb =
// (vx, vy, vz, rvx), wa, wb = (rvy, rvz).
__device__ __forceinline__ void scatter_vals(const CollideParams &cp, uint32_t stream, uint32_t cell,
                                             uint32_t step, int tr, f4 &va, f2 &wa, f4 &vb, f2 &wb)
{
    const double vax = (double)va.x + (double)va.w;
    const double vay = (double)va.y + (double)wa.x;
    const double vaz = (double)va.z + (double)wa.y;
    const double vbx = (double)vb.x + (double)vb.w;
    const double vby = (double)vb.y + (double)wb.x;
    const double vbz = (double)vb.z + (double)wb.y;
    const double ddx = vax - vbx, ddy = vay - vby, ddz = vaz - vbz;
    const double magVrel = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    const double cx = 0.5 * (vax + vbx), cy = 0.5 * (vay + vby), cz = 0.5 * (vaz + vbz);
    double nx, ny, nz;
    sphere_point_f(cp.seed, stream, cell, step, (uint32_t)tr, nx, ny, nz);
    const double wx = magVrel * nx, wy = magVrel * ny, wz = magVrel * nz;
    FState na, nb;
    split_df(cx - 0.5 * wx, na.vx, na.rvx);
    split_df(cy - 0.5 * wy, na.vy, na.rvy);
    split_df(cz - 0.5 * wz, na.vz, na.rvz);
    split_df(cx + 0.5 * wx, nb.vx, nb.rvx);
    split_df(cy + 0.5 * wy, nb.vy, nb.rvy);
    split_df(cz + 0.5 * wz, nb.vz, nb.rvz);
    va = f4{ na.vx, na.vy, na.vz, na.rvx };
    wa = f2{ na.rvy, na.rvz };
    vb = f4{ nb.vx, nb.vy, nb.vz, nb.rvx };
    wb = f2{ nb.rvy, nb.rvz };
}

// Where the trials of a cell find the atoms' velocities.  A cell's atoms are
// its slots c.i .. c.i + c.n - 1 in sorted order; slot() maps a sorted slot
// to the index the velocity accessors take.
//   VelGlobal: the state records in HBM, atom of sorted slot j = perm[j]
//              (the state is in the previous order until the move gathers it);
//   VelLds:    the fused step's block copy in LDS, index = slot - block start.
struct VelGlobal {
    f4 *r;
    const uint32_t *perm;
    __device__ __forceinline__ uint32_t slot(int j) const { return perm[j]; }
    __device__ __forceinline__ f4 vel(uint32_t a) const { return *vel_rec(r, a); }
    __device__ __forceinline__ void scatter(const CollideParams &cp, uint32_t stream, uint32_t cell,
                                            uint32_t step, int tr, uint32_t a, uint32_t b) const
    {
        f4 *const ha = vel_rec(r, a), *const hb = vel_rec(r, b);
        f2 *const za = vel_rem_yz(r, a), *const zb = vel_rem_yz(r, b);
        f4 va = *ha, vb = *hb;
        f2 wa = *za, wb = *zb;
        scatter_vals(cp, stream, cell, step, tr, va, wa, vb, wb);
        *ha = va; *za = wa;
        *hb = vb; *zb = wb;
    }
    // the wave's next round reads what one lane wrote
    __device__ __forceinline__ void publish() const { __threadfence(); }
};

struct VelLds {
    f4 *v;
    f2 *w;
    int b0;
    __device__ __forceinline__ uint32_t slot(int j) const { return (uint32_t)(j - b0); }
    __device__ __forceinline__ f4 vel(uint32_t a) const { return v[a]; }
    __device__ __forceinline__ void scatter(const CollideParams &cp, uint32_t stream, uint32_t cell,
                                            uint32_t step, int tr, uint32_t a, uint32_t b) const
    {
        f4 va = v[a], vb = v[b];
        f2 wa = w[a], wb = w[b];
        scatter_vals(cp, stream, cell, step, tr, va, wa, vb, wb);
        v[a] = va; w[a] = wa;
        v[b] = vb; w[b] = wb;
    }
    __device__ __forceinline__ void publish() const
    {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
};

// The first trial of a cell (its pair, acceptance uniform and relative
// speed), loaded before the cell's trial count is known: the count needs
// sigma v_r,max, and the two loads then overlap.
struct Trial0 {
    uint32_t a, b;
    float u, vr;
};

template <class V>
__device__ __forceinline__ void trial0_load(const CollideParams &cp, uint32_t stream, uint32_t cell,
                                            uint32_t step, int n, int i, const V &vs, Trial0 &t)
{
    int ia, ib;
    trial_draw(cp, stream, cell, step, 0, n, ia, ib, t.u);
    t.a = vs.slot(i + ia);
    t.b = vs.slot(i + ib);
    t.vr = rel_speed_v(vs.vel(t.a), vs.vel(t.b));
}

// The trials of one cell in sequence, by one lane (collisions.cu:397-441).
template <class V>
__device__ __forceinline__ int cell_trials_serial(const CollideParams &cp, const CellJob &c,
                                                  uint32_t stream, uint32_t step, const V &vs,
                                                  double &smax, const Trial0 &t0)
{
    int count = 0;
    for (int tr = 0; tr < c.ncol; ++tr) {
        uint32_t a = t0.a, b = t0.b;
        float u = t0.u, vr = t0.vr;
        if (tr > 0) {
            int ia, ib;
            trial_draw(cp, stream, c.cell, step, tr, c.n, ia, ib, u);
            a = vs.slot(c.i + ia);
            b = vs.slot(c.i + ib);
            vr = rel_speed_v(vs.vel(a), vs.vel(b));
        }
        const double sv = (double)vr * cp.sigma;
        if (sv > smax) smax = sv;
        if (c.pc * vr > u) {
            vs.scatter(cp, stream, c.cell, step, tr, a, b);
            ++count;
        }
    }
    return count;
}

// The trials of one cell by the whole wave, with the result of the sequential
// loop: 64 consecutive trials are evaluated at once; up to the first accepted
// one they are exactly the sequential trials (no velocity changed before
// them); that collision is applied and evaluation resumes after it.  All 64
// lanes call this with the same job.
template <class V>
__device__ __forceinline__ int cell_trials_wave(const CollideParams &cp, const CellJob &c,
                                                uint32_t stream, uint32_t step, const V &vs,
                                                double &smax)
{
    const int lane = threadIdx.x & (kWave - 1);
    int count = 0;
    int tr0 = 0;
    while (tr0 < c.ncol) {
        const int tr = tr0 + lane;
        const bool active = tr < c.ncol;
        uint32_t a = 0, b = 0;
        double sv = 0.0;
        bool acc = false;
        if (active) {
            int ia, ib;
            float u;
            trial_draw(cp, stream, c.cell, step, tr, c.n, ia, ib, u);
            a = vs.slot(c.i + ia);
            b = vs.slot(c.i + ib);
            const float vr = rel_speed_v(vs.vel(a), vs.vel(b));
            sv = (double)vr * cp.sigma;
            acc = c.pc * vr > u;
        }
        const uint64_t mask = __ballot(acc);
        const int f = mask ? __ffsll((unsigned long long)mask) - 1 : kWave;
        double m = (active && lane <= f) ? sv : 0.0;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
        if (m > smax) smax = m;
        if (f < kWave) {
            if (lane == f) vs.scatter(cp, stream, c.cell, step, tr, a, b);
            vs.publish();   // the next round reads the updated velocities
            ++count;
            tr0 += f + 1;
        } else {
            tr0 += kWave;
        }
    }
    return count;
}

constexpr int kCollItems = 8;  // slots per lane scanned for cell heads
constexpr uint32_t kNoKey = 0xFFFFFFFFu;   // no slot there (keys are below 2^31)
constexpr int kWaveTrialsDefault = 16;

// Trials of one cell (serial by one lane, or by the whole wave when heavy);
// sigma v_r,max is stored when it grew (unless `dry`).  Returns the accepted
// collisions.
template <class V>
__device__ __forceinline__ int run_cell(const CollideParams &cp, const CellJob &c, uint32_t stream,
                                        uint32_t step, const V &vs, float *__restrict__ sigvrmax,
                                        bool wave, const Trial0 &t0, bool dry = false)
{
    double smax = c.smax;
    const int cnt = wave ? cell_trials_wave(cp, c, stream, step, vs, smax)
                         : cell_trials_serial(cp, c, stream, step, vs, smax, t0);
    const bool writer = !wave || (threadIdx.x & (kWave - 1)) == 0;
    if (writer && smax != c.smax && !dry) sigvrmax[c.k] = (float)smax;
    return cnt;
}

// Each wave takes kWave * kCollItems consecutive sorted slots and runs the
// cells that start there (a cell ends where the next one starts, possibly in
// the next chunk).  Only cells of two or more atoms are compacted for the
// lanes (ballot + popcount), one per lane; cells with more than
// `wave_trials` trials are then run by the whole wave.  When the chunk lies
// in one trap, its parameters are wave-uniform (scalar loads).  Five waves
// per SIMD: the kernel is bound by the latency of its dependent loads (key,
// sigma v_r,max, permutation, velocities).  CB threads per block (one or
// four waves).
template <int CB>
__global__ __launch_bounds__(CB, 5) void k_collide_fused(
    const TrapDev *__restrict__ traps, const int *__restrict__ key_base, int T, CollideParams cp,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ perm,
    const int *__restrict__ n_dev, int n_cap, FusedPtrs s, float *__restrict__ sigvrmax,
    unsigned long long *__restrict__ coll_accum, const long long *__restrict__ step_dev,
    int wave_trials)
{
    constexpr int kChunk = kWave * kCollItems;
    extern __shared__ int kb[];
    __shared__ int heads[CB / kWave][kChunk + 1];
    __shared__ uint32_t hkey[CB / kWave][kChunk];   // key of each head
    __shared__ short multi[CB / kWave][kChunk];   // heads index of cells with >= 2 atoms
    __shared__ short heavy[CB / kWave][kChunk];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int chunk0 = (blockIdx.x * (CB / kWave) + w) * kChunk;
    const VelGlobal vg = { s.r, perm };
    const uint64_t lt = lanemask_lt();
    // the chunk's keys, the one before it and the 64 after it (for the end
    // of its last cell) are loaded with the alive count, the step and the
    // trap table's first 2 CB entries; kNoKey marks slots past the alive
    // atoms
    const int tb_j = (int)threadIdx.x;
    const int tb0 = tb_j < T + 1 ? key_base[tb_j] : 0;
    const int tb1 = tb_j + CB < T + 1 ? key_base[tb_j + CB] : 0;
    uint32_t kk[kCollItems];
#pragma unroll
    for (int it = 0; it < kCollItems; ++it) {
        const int i = chunk0 + it * kWave + lane;
        kk[it] = i < n_cap ? key[i] : kNoKey;
    }
    const int jp = chunk0 + kChunk + lane;
    uint32_t pk = jp < n_cap ? key[jp] : kNoKey;
    uint32_t k_before = chunk0 > 0 && chunk0 <= n_cap ? key[chunk0 - 1] : kNoKey;
    const int n = *n_dev;
    const uint32_t step = (uint32_t)*step_dev;
    const int chunk1 = min(n, chunk0 + kChunk);
#pragma unroll
    for (int it = 0; it < kCollItems; ++it)
        if (chunk0 + it * kWave + lane >= n) kk[it] = kNoKey;
    if (jp >= n) pk = kNoKey;
    if (chunk0 > n) k_before = kNoKey;
    const uint32_t k_after = __shfl(pk, 0, kWave);
    if (tb_j < T + 1) kb[tb_j] = tb0;
    if (tb_j + CB < T + 1) kb[tb_j + CB] = tb1;
    for (int j = tb_j + 2 * CB; j < T + 1; j += CB) kb[j] = key_base[j];
    __syncthreads();
    int nh = 0, nm = 0;
    uint32_t klast = 0;   // key of the chunk's last cell head
#pragma unroll
    for (int it = 0; it < kCollItems; ++it) {
        const int i = chunk0 + it * kWave + lane;
        const uint32_t k = kk[it];
        // neighbours from the adjacent lanes (and items)
        uint32_t kp = __shfl_up(k, 1, kWave), kn = __shfl_down(k, 1, kWave);
        const uint32_t prev_last = it > 0 ? __shfl(kk[it > 0 ? it - 1 : 0], kWave - 1, kWave) : k_before;
        const uint32_t next_first = it + 1 < kCollItems ? __shfl(kk[it + 1 < kCollItems ? it + 1 : it], 0, kWave)
                                                        : k_after;
        if (lane == 0) kp = prev_last;
        if (lane == kWave - 1) kn = next_first;
        bool head = false, mult = false;
        if (i < n) {
            head = i == 0 || kp != k;
            mult = head && i + 1 < n && kn == k;
        }
        const uint64_t bh = __ballot(head), bm = __ballot(mult);
        const int pos = nh + __popcll(bh & lt);
        if (head) { heads[w][pos] = i; hkey[w][pos] = k; }
        if (mult) multi[w][nm + __popcll(bm & lt)] = (short)pos;
        if (bh) klast = __shfl(k, 63 - __clzll((long long)bh), kWave);
        nh += __popcll(bh);
        nm += __popcll(bm);
    }
    // sigma v_r,max of this lane's first cell, loaded now (the load overlaps
    // the end-of-chunk probe and the barrier); each round loads the next
    // round's while it runs
    float sm_next = lane < nm ? sigvrmax[hkey[w][multi[w][lane]]] : 0.0f;
    if (nh > 0) {
        // end of the last cell: the first later slot with another key.  The
        // wave probes the next 64 slots at once; a longer cell (e.g. the
        // overflow cell of atoms outside the grid) is searched by galloping.
        const uint64_t bd = __ballot(chunk1 == chunk0 + kChunk && pk != kNoKey && pk != klast);
        int end = chunk1 + kWave;
        if (bd) end = chunk1 + __ffsll((unsigned long long)bd) - 1;
        if (chunk1 + kWave >= n) end = min(end, n);
        if (lane == 0) {
            if (!bd && end < n) {
                int lo = end, stp = 1;
                while (lo + stp - 1 < n && key[lo + stp - 1] == klast) { lo += stp; stp <<= 1; }
                int hi = min(n, lo + stp - 1);  // the first other key is in [lo, hi]
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (key[mid] == klast) lo = mid + 1; else hi = mid;
                }
                end = lo;
            }
            heads[w][nh] = end;
        }
    }
    // traps of the chunk's first and last cells: usually the same
    int t_lo = 0, t_hi = 0;
    if (nh > 0) {
        t_lo = __builtin_amdgcn_readfirstlane(last_le(kb, T + 1, (int)__shfl(kk[0], 0, kWave)));
        t_hi = __builtin_amdgcn_readfirstlane(last_le(kb, T + 1, (int)klast));
    }
    __syncthreads();
    int nheavy = 0;
    unsigned long long wave_cnt = 0;   // accepted collisions of trap t_lo (uniform chunks)
    for (int m0 = 0; m0 < nm; m0 += kWave) {
        const int mi = m0 + lane;
        const float sm = sm_next;
        if (mi + kWave < nm) sm_next = sigvrmax[hkey[w][multi[w][mi + kWave]]];
        CellJob c;
        Trial0 t0;
        bool ok = false;
        int h = 0;
        uint32_t stream = 0;
        if (mi < nm) {
            h = multi[w][mi];
            const int i = heads[w][h], end = heads[w][h + 1];
            const uint32_t k = hkey[w][h];
            auto setup = [&](const TrapDev &tr, int t) {
                stream = tr.stream;
                trial0_load(cp, stream, k - (uint32_t)tr.key_base, step, end - i, i, vg, t0);
                ok = cell_setup(tr, t, k, cp, (double)sm, i, end, c);
            };
            if (t_lo == t_hi) {
                setup(traps[t_lo], t_lo);
            } else {
                const int t = last_le(kb, T + 1, (int)k);
                setup(traps[t], t);
            }
        }
        const bool is_heavy = ok && c.ncol > wave_trials;
        const uint64_t hb = __ballot(is_heavy);
        if (is_heavy) heavy[w][nheavy + __popcll(hb & lt)] = (short)mi;
        nheavy += __popcll(hb);
        if (ok && !is_heavy) {
            const int cnt = run_cell(cp, c, stream, step, vg, sigvrmax, false, t0);
            if (t_lo == t_hi) wave_cnt += (unsigned long long)cnt;
            else if (cnt) atomicAdd(&coll_accum[c.t], (unsigned long long)cnt);
        }
    }
    __syncthreads();
    for (int q = 0; q < nheavy; ++q) {
        const int h = multi[w][heavy[w][q]];
        const int i = heads[w][h];
        const uint32_t k = hkey[w][h];
        const int t = t_lo == t_hi ? t_lo : last_le(kb, T + 1, (int)k);
        CellJob c;
        (void)cell_setup(traps[t], t, k, cp, (double)sigvrmax[k], i, heads[w][h + 1], c);
        const Trial0 none = { 0u, 0u, 0.0f, 0.0f };
        const int cnt = run_cell(cp, c, traps[t].stream, step, vg, sigvrmax, true, none);
        if (lane == 0) {
            if (t_lo == t_hi) wave_cnt += (unsigned long long)cnt;
            else if (cnt) atomicAdd(&coll_accum[t], (unsigned long long)cnt);
        }
    }
    if (t_lo == t_hi) {
        wave_cnt = wave_sum(wave_cnt);
        if (lane == 0 && wave_cnt) atomicAdd(&coll_accum[t_lo], wave_cnt);
    }
}

// ---------------------------------------------------------------------------
// Fused motion step: gather through the permutation, spin substeps,
// Kahan-compensated velocity Verlet (moveAtoms.cu:51-91, acceleration
// recomputed instead of stored), periodic wrap, evaporation (spin projection
// of evaporation.cu:90-94 and/or energy cut), next cell key, and FP64
// observables accumulated per block and trap; k_obs_finalize then writes one
// log record per trap and advances the device step counter.
// ---------------------------------------------------------------------------
// One atom of the fused move, in registers between its load and its update.
struct MoveAtom {
    FState f;
    float sx, sy, sz;
    int id;
    int t;          // trap
    uint32_t kold;  // key of the last sort (before the move), when counting
};

template <bool SPIN>
__device__ __forceinline__ void move_load(const MoveParams &mp, const FusedPtrs &A,
                                          const TrapDev &tr, uint32_t j, MoveAtom &m)
{
    rec_load_stream(A.r, j, m.f, m.id);
    m.sx = m.sy = m.sz = 0.0f;
    if (SPIN) { m.sx = ld_at(A.s[0], j); m.sy = ld_at(A.s[1], j); m.sz = ld_at(A.s[2], j); }
    m.kold = 0;
}

// The Bloch vector s rotated by the angle phi about the unit axis n
// (Rodrigues): s cos phi + (n x s) sin phi + n (n . s)(1 - cos phi); the
// precession of one spin substep, phi = gs muB |B| dt / hbar.
__device__ __forceinline__ void rotate_about(const F3 &n, float phi, float &sx, float &sy, float &sz)
{
    float sn, cs;
    sincos_f32(phi, &sn, &cs);
    const float nd = fmaf(n.x, sx, fmaf(n.y, sy, n.z * sz));
    const float kx = fmaf(n.y, sz, -(n.z * sy));
    const float ky = fmaf(n.z, sx, -(n.x * sz));
    const float kz = fmaf(n.x, sy, -(n.y * sx));
    const float om = nd * (1.0f - cs);
    sx = fmaf(sx, cs, fmaf(kx, sn, n.x * om));
    sy = fmaf(sy, cs, fmaf(ky, sn, n.y * om));
    sz = fmaf(sz, cs, fmaf(kz, sn, n.z * om));
}

// Half a velocity kick, v += a dt / 2 (Kahan), with the trap's acceleration
// a at the position; in the quadrupole a = g (x, y, 4 z) with g = -mu B' /
// (m rho), so each component is one fma into the remainder.
__device__ __forceinline__ void half_kick(const TrapDev &tr, FState &f, float hdt)
{
    if (tr.kind == kTrapQuad) {
        const float r2 = quad_r2(f.px, f.py, f.pz);
        const float g = r2 > 0.0f ? (tr.f_neg_mu_over_m * tr.f_dBdz * hdt) * rsqrt_f(r2) : 0.0f;
        kahan_fma(f.vx, f.rvx, g, f.px);
        kahan_fma(f.vy, f.rvy, g, f.py);
        kahan_fma(f.vz, f.rvz, g, 4.0f * f.pz);
    } else {
        const F3 a = accel_f(tr, f.px, f.py, f.pz);
        kahan_fma(f.vx, f.rvx, a.x, hdt);
        kahan_fma(f.vy, f.rvy, a.y, hdt);
        kahan_fma(f.vz, f.rvz, a.z, hdt);
    }
}

// Spin substeps, Kahan velocity Verlet, wrap, evaporation test, next key;
// stores the atom into slot i of B.  vals = (vx^2, vy^2, vz^2, U) of a
// survivor; lost_e = energy of a removed atom.
template <bool SPIN>
__device__ __forceinline__ void move_update(const MoveParams &mp, const TrapDev &tr, MoveAtom &m,
                                            long long step, int i, const FusedPtrs &B,
                                            uint32_t *__restrict__ key_out,
                                            float *__restrict__ evap_pos,
                                            float *__restrict__ evap_vel, double (&vals)[4],
                                            double &lost_e, bool &lost, bool &moved)
{
    const float dt = mp.f_dt;
    const float hdt = 0.5f * dt;
    half_kick(tr, m.f, hdt);

    if (SPIN) {
        const double dts = mp.dt / mp.substeps;
        const float w = (float)(mp.gs * kMuB * dts / kHbar);
        const float fdts = (float)dts;
        float sx = m.sx, sy = m.sy, sz = m.sz;
        if (tr.kind == kTrapQuad) {
            // the quadrupole field inline: B ~ (x, y, -2 z), |B| = B' rho
            const float dBdz = tr.f_dBdz;
            for (int k = 0; k < mp.substeps; ++k) {
                const float tau = mp.interp ? ((float)k + 0.5f) * fdts : 0.0f;
                const float x = fmaf(m.f.vx, tau, m.f.px), y = fmaf(m.f.vy, tau, m.f.py),
                            z = fmaf(m.f.vz, tau, m.f.pz);
                const float r2 = fmaf(x, x, fmaf(y, y, 4.0f * (z * z)));
                // at the field's zero (or a NaN position) no rotation, as field_dir_f
                const bool zero = !(r2 > 0.0f);
                const float inv = zero ? 0.0f : rsqrt_f(r2);
                const F3 Bn = { x * inv, y * inv, zero ? 1.0f : (-2.0f * z) * inv };
                rotate_about(Bn, zero ? 0.0f : w * (dBdz * (r2 * inv)), sx, sy, sz);
            }
        } else {
            for (int k = 0; k < mp.substeps; ++k) {
                const float tau = mp.interp ? ((float)k + 0.5f) * fdts : 0.0f;
                F3 Bn;
                const float Bm = field_dir_f(tr, fmaf(m.f.vx, tau, m.f.px), fmaf(m.f.vy, tau, m.f.py),
                                             fmaf(m.f.vz, tau, m.f.pz), Bn);
                rotate_about(Bn, w * Bm, sx, sy, sz);
            }
        }
        const double L = sqrt((double)sx * sx + (double)sy * sy + (double)sz * sz);
        m.sx = (float)((double)sx / L); m.sy = (float)((double)sy / L); m.sz = (float)((double)sz / L);
    }

    kahan_fma(m.f.px, m.f.rpx, m.f.vx, dt);
    kahan_fma(m.f.py, m.f.rpy, m.f.vy, dt);
    kahan_fma(m.f.pz, m.f.rpz, m.f.vz, dt);
    if (tr.periodic) {  // wrap the FP64 value, then split it again
        double X = (double)m.f.px + (double)m.f.rpx, Y = (double)m.f.py + (double)m.f.rpy,
               Z = (double)m.f.pz + (double)m.f.rpz;
        if (X >= tr.hw[0] || X < -tr.hw[0]) { X = wrap_periodic(X, tr.hw[0]); split_df(X, m.f.px, m.f.rpx); }
        if (Y >= tr.hw[1] || Y < -tr.hw[1]) { Y = wrap_periodic(Y, tr.hw[1]); split_df(Y, m.f.py, m.f.rpy); }
        if (Z >= tr.hw[2] || Z < -tr.hw[2]) { Z = wrap_periodic(Z, tr.hw[2]); split_df(Z, m.f.pz, m.f.rpz); }
    }
    // the position remainders as they are stored (rounded below the label's
    // bits), so that the force, the energies and the key are those of the
    // stored position
    m.f.rpx = rem_round(m.f.rpx); m.f.rpy = rem_round(m.f.rpy); m.f.rpz = rem_round(m.f.rpz);
    half_kick(tr, m.f, hdt);

    // Per-atom energies, cut test and cell key in FP32 (the FP64 sums and the
    // exact index near a cell face aside): the squared speeds and the
    // quadrupole potential from the FP32 values; the Ioffe-Pritchard
    // potential mu (|B| - B0) in FP64, where |B| - B0 cancels.
    const float v2x = m.f.vx * m.f.vx, v2y = m.f.vy * m.f.vy, v2z = m.f.vz * m.f.vz;
    const float ekf = mp.f_half_mass * (v2x + v2y + v2z);
    double ep;
    if (tr.kind == kTrapQuad) {
        // rho = r2 / sqrt(r2) with the reciprocal square root of the force
        // (the same r2, so one rsq serves both)
        const float r2 = quad_r2(m.f.px, m.f.py, m.f.pz);
        ep = r2 > 0.0f ? (double)(tr.f_mu_dBdz * (r2 * rsqrt_f(r2))) : 0.0;
    } else if (tr.kind == kTrapIP) {
        ep = potential_energy(tr, (double)m.f.px + (double)m.f.rpx, (double)m.f.py + (double)m.f.rpy,
                              (double)m.f.pz + (double)m.f.rpz);
    } else {
        ep = 0.0;
    }
    lost = false;
    if (mp.spin_loss) {
        F3 Bn;
        (void)field_dir_f(tr, m.f.px, m.f.py, m.f.pz, Bn);
        lost = Bn.x * m.sx + Bn.y * m.sy + Bn.z * m.sz < 0.0f;
    }
    if (mp.energy_cut && tr.kT_cut > 0.0) {
        const double f = tr.eta_steps > 0 ? fmin((double)step / (double)tr.eta_steps, 1.0) : 1.0;
        const double eta = tr.eta_start + (tr.eta_end - tr.eta_start) * f;
        lost = lost || ((double)ekf + ep > eta * tr.kT_cut);
    }
    uint32_t kout;
    moved = false;
    if (lost) {
        // a removed atom's record in FP64 from value + remainder (rare)
        const double X = (double)m.f.px + (double)m.f.rpx, Y = (double)m.f.py + (double)m.f.rpy,
                     Z = (double)m.f.pz + (double)m.f.rpz;
        const double VX = (double)m.f.vx + (double)m.f.rvx, VY = (double)m.f.vy + (double)m.f.rvy,
                     VZ = (double)m.f.vz + (double)m.f.rvz;
        kout = mp.dead_key;
        const size_t id = (size_t)m.id;
        evap_pos[3 * id + 0] = (float)X; evap_pos[3 * id + 1] = (float)Y;
        evap_pos[3 * id + 2] = (float)Z;
        evap_vel[3 * id + 0] = (float)VX; evap_vel[3 * id + 1] = (float)VY;
        evap_vel[3 * id + 2] = (float)VZ;
        lost_e = 0.5 * mp.mass * (VX * VX + VY * VY + VZ * VZ) + potential_energy(tr, X, Y, Z);
    } else {
        int c;
        if (tr.f32_index) {
            const float p[3] = { m.f.px, m.f.py, m.f.pz }, rp[3] = { m.f.rpx, m.f.rpy, m.f.rpz };
            c = cell_of_f32(tr, p, rp);
        } else {
            c = cell_of_fast(tr, (double)m.f.px + (double)m.f.rpx, (double)m.f.py + (double)m.f.rpy,
                             (double)m.f.pz + (double)m.f.rpz);
        }
        kout = (uint32_t)(tr.key_base + c);
        vals[0] = (double)v2x; vals[1] = (double)v2y; vals[2] = (double)v2z;
        vals[3] = ep;
        if (mp.count_moved) moved = kout != m.kold;
    }
    const uint32_t u = (uint32_t)i;
    rec_store_stream(B.r, u, m.f, m.id);
    if (SPIN) {
        st_stream_at(B.s[0], u, m.sx); st_stream_at(B.s[1], u, m.sy); st_stream_at(B.s[2], u, m.sz);
    }
    key_out[u] = kout;   // cached: the re-sort reads the keys next
}

// Arguments of the collisions inside the fused step kernel (COLL).
// `dry`: sigma v_r,max and the collision counts are not written (autotuner
// trial launches).
struct CollArgs {
    CollideParams cp;
    float *sigvrmax;
    unsigned long long *coll_accum;
    int wave_trials;
    int dry;
};

// The fused step runs the collisions wave by wave: a wave holds W = 64 APT
// consecutive sorted slots (its "span") and, next to them, kExt slots of the
// neighbouring spans on each side.  A cell across a span boundary with at
// most kExt slots on either side of it is run by both waves (the same
// trials, so the same velocities; the wave holding its first slot stores
// sigma v_r,max and counts its collisions).  Larger cells across a boundary
// ("oversized") are run before the step kernel by k_collide_edges.
constexpr int kExt = 16;

// Oversized: the cell across boundary B (slots B - 1 and B hold key k) has
// more than kExt slots on one side of it.  kb = key of slot B - kExt - 1
// (kNoKey below 0), ka = key of slot B + kExt (kNoKey from n on).
__device__ __forceinline__ bool oversized(uint32_t k, uint32_t kb, uint32_t ka)
{
    return kb == k || ka == k;
}

// LDS of one wave of the fused step: keys of the slots ws - kExt - 1 ..
// ws + W + kExt (index slot - (ws - kExt - 1); kNoKey outside 0 .. n - 1),
// velocities of the slots ws - kExt .. ws + W + kExt - 1 (index slot - (ws -
// kExt)), sigma v_r,max of the span's slots, the cell list.
template <int APT>
struct WaveColl {
    static constexpr int W = kWave * APT;
    f4 v[W + 2 * kExt];
    f2 w[W + 2 * kExt];
    uint32_t k[W + 2 * kExt + 2];
    float sm[W];
    short mh[W / 2 + 1], me[W / 2 + 1], hv[W / 2 + 1];
};

// The collisions of one wave's span [ws, ws + len) (len <= W).  pre_h /
// pre_e, post_h / post_e: the cells across its first and last boundary that
// it runs (-1: none).  Each lane runs one cell (those with more than
// wave_trials trials are run by the whole wave); the same trial sequence per
// cell as k_collide_fused, so the velocities are bitwise the same.
template <int APT>
__device__ __forceinline__ void wave_collide(const TrapDev *__restrict__ traps, const int *ts, int T,
                                             bool one, int t_first, const CollArgs &ca, uint32_t step,
                                             int ws, int len, WaveColl<APT> &L, int pre_h, int pre_e,
                                             int post_h)
{
    constexpr int W = WaveColl<APT>::W;
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t lt = lanemask_lt();
    const int xb = ws - kExt - 1, vb = ws - kExt;
    const int le = ws + len;
    const VelLds vl = { L.v, L.w, vb };
    // heads of the cells of >= 2 atoms in the span that this wave runs: those
    // that end inside it, and the one across its last boundary when it is
    // not oversized (at most 32 per 64 slots)
    int nm = 0;
#pragma unroll
    for (int k = 0; k < APT; ++k) {
        const int l = k * kWave + lane;
        bool mult = false;
        int e = 0;
        if (l < len) {
            const int sl = ws + l;
            const uint32_t kk = L.k[sl - xb];
            if (L.k[sl - 1 - xb] != kk && L.k[sl + 1 - xb] == kk) {
                int lo = sl + 1, hi = le + kExt;   // first later slot with another key
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (L.k[mid - xb] == kk) lo = mid + 1; else hi = mid;
                }
                e = lo;
                mult = e <= le || sl == post_h;
            }
        }
        const uint64_t bm = __ballot(mult);
        if (mult) {
            const int p = nm + __popcll(bm & lt);
            L.mh[p] = (short)(ws + l - vb);
            L.me[p] = (short)(e - vb);
        }
        nm += __popcll(bm);
    }
    if (pre_h >= 0) {   // the cell across the first boundary
        if (lane == 0) {
            L.mh[nm] = (short)(pre_h - vb);
            L.me[nm] = (short)(pre_e - vb);
        }
        ++nm;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int nheavy = 0;
    unsigned long long wave_cnt = 0;   // accepted collisions of owned cells (one-trap blocks)
    for (int m0 = 0; m0 < nm; m0 += kWave) {
        const int mi = m0 + lane;
        CellJob c;
        Trial0 t0;
        bool ok = false, own = false;
        uint32_t stream = 0;
        if (mi < nm) {
            const int h = vb + L.mh[mi], e = vb + L.me[mi];
            own = h >= ws;
            const uint32_t k = L.k[h - xb];
            const int t = one ? t_first : last_le(ts, T, max(h, ws));
            const TrapDev &tr = traps[t];
            stream = tr.stream;
            const float sv = own ? L.sm[h - ws] : ca.sigvrmax[k];
            trial0_load(ca.cp, stream, k - (uint32_t)tr.key_base, step, e - h, h, vl, t0);
            ok = cell_setup(tr, t, k, ca.cp, (double)sv, h, e, c);
        }
        const bool is_heavy = ok && c.ncol > ca.wave_trials;
        const uint64_t hb = __ballot(is_heavy);
        if (is_heavy) L.hv[nheavy + __popcll(hb & lt)] = (short)mi;
        nheavy += __popcll(hb);
        if (ok && !is_heavy) {
            const int cnt = run_cell(ca.cp, c, stream, step, vl, ca.sigvrmax, false, t0, ca.dry != 0 || !own);
            if (own) {
                if (one) wave_cnt += (unsigned long long)cnt;
                else if (cnt && !ca.dry) atomicAdd(&ca.coll_accum[c.t], (unsigned long long)cnt);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < nheavy; ++q) {
        const int mi = L.hv[q];
        const int h = vb + L.mh[mi], e = vb + L.me[mi];
        const bool own = h >= ws;
        const uint32_t k = L.k[h - xb];
        const int t = one ? t_first : last_le(ts, T, max(h, ws));
        CellJob c;
        (void)cell_setup(traps[t], t, k, ca.cp, (double)(own ? L.sm[h - ws] : ca.sigvrmax[k]), h, e, c);
        const Trial0 none = { 0u, 0u, 0.0f, 0.0f };
        const int cnt = run_cell(ca.cp, c, traps[t].stream, step, vl, ca.sigvrmax, true, none,
                                 ca.dry != 0 || !own);
        if (lane == 0 && own) {
            if (one) wave_cnt += (unsigned long long)cnt;
            else if (cnt && !ca.dry) atomicAdd(&ca.coll_accum[t], (unsigned long long)cnt);
        }
    }
    if (one) {
        wave_cnt = wave_sum(wave_cnt);
        if (lane == 0 && wave_cnt && !ca.dry) atomicAdd(&ca.coll_accum[t_first], wave_cnt);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    (void)W;
}

// A block of BS threads moves APT * BS consecutive slots, APT per thread: all
// of a thread's atoms are loaded before the first is updated, so their loads
// are in flight together.  COLL (the fused step): between the loads and the
// updates, the block runs the collisions of its cells on an LDS copy of the
// velocities (block_collide).
// Without spin, two atoms per thread fit 80 registers: six waves per SIMD.
template <int BS, int APT, bool SPIN, bool COLL>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(APT == 2 && !SPIN && !COLL ? 6 : 1)))
void k_move_fused(
    const TrapDev *__restrict__ traps, MoveParams mp, FusedPtrs A, FusedPtrs B,
    const uint32_t *__restrict__ perm, const int *__restrict__ trap_start,
    const int *__restrict__ n_dev, uint32_t *__restrict__ key_out, float *__restrict__ evap_pos,
    float *__restrict__ evap_vel, int *__restrict__ rec_trap, double *__restrict__ rec_sum,
    int rec_stride, double *__restrict__ acc_over, const long long *__restrict__ step_dev,
    CollArgs ca)
{
    extern __shared__ int ts[];
    __shared__ double red[(BS / kWave) * 4];
    __shared__ uint32_t wcnt[BS / kWave][2];
    // (COLL) LDS of each wave's collisions
    __shared__ WaveColl<APT> wcl[COLL ? BS / kWave : 1];
    const int T = mp.n_traps;
    load_table(ts, trap_start, T + 1);
    const int n = *n_dev;
    const long long step = *step_dev;
    // logical block: consecutive on one XCD, so an out-of-place atom's cache
    // lines are mostly in the L2 its neighbours filled
    const int blk = (int)blockIdx.x;
    const int b0 = blk * (BS * APT);
    // traps covered by this block (the same in every lane): usually one
    const int t_first = __builtin_amdgcn_readfirstlane(b0 < n ? last_le(ts, T, b0) : 0);
    const int t_last =
        __builtin_amdgcn_readfirstlane(b0 < n ? last_le(ts, T, min(n, b0 + BS * APT) - 1) : 0);
    MoveAtom at[APT];
    // per atom: sum v_x^2, v_y^2, v_z^2 and U of survivors (kinetic energy and
    // the alive count follow from them and the trap ranges), and the energy of
    // evaporated atoms
    double vals[APT][4], lost_e[APT];
    bool lost[APT], moved[APT];
#pragma unroll
    for (int k = 0; k < APT; ++k) {
        vals[k][0] = vals[k][1] = vals[k][2] = vals[k][3] = 0.0;
        lost_e[k] = 0.0;
        lost[k] = moved[k] = false;
    }
    if constexpr (COLL) {
        // The fused step, wave by wave (no block barrier until the sums): the
        // wave's span is slots ws .. ws + W - 1, a lane's atoms ws + 64 k +
        // lane.  All loads are issued at once: each atom's slot in the
        // previous order, its sorted key and sigma v_r,max of that cell, its
        // three record planes, label and spin; and the keys and velocities of
        // kExt slots on each side.  The velocity planes go to the wave's LDS,
        // it runs the collisions of its cells there, and the atoms take their
        // velocities back.  Positions, remainders and labels stay in
        // registers.
        constexpr int W = kWave * APT;
        WaveColl<APT> &L = wcl[COLL ? threadIdx.x / kWave : 0];
        const int lane = threadIdx.x & (kWave - 1);
        const int ws = b0 + (int)(threadIdx.x / kWave) * W;
        const int len = min(n - ws, W);   // <= 0: no atoms
        const int xb = ws - kExt - 1, vb = ws - kExt;
        const bool one = t_first == t_last;
        uint32_t jj[APT];
        f4 p0[APT];
        float rpy[APT], rpz[APT];
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int i = ws + k * kWave + lane;
            at[k].t = -1;
            jj[k] = 0;
            if (i < n) {
                int t = t_first;
                if (!one) while (t < t_last && ts[t + 1] <= i) ++t;
                at[k].t = t;
                jj[k] = perm ? ld_stream_at(perm, (uint32_t)i) : (uint32_t)i;
                at[k].kold = ld_at(mp.key_prev, (uint32_t)i);
            }
        }
        // neighbours: lanes 0 .. kExt load the kExt + 1 keys before the span
        // (and the kExt velocities), lanes 32 .. 32 + kExt those after it
        {
            const bool pre = lane <= kExt, post = lane >= 32 && lane <= 32 + kExt;
            const int sl = pre ? xb + lane : ws + W + (lane - 32);
            uint32_t kn = kNoKey;
            f4 v1 = f4{ 0.0f, 0.0f, 0.0f, 0.0f }, v2 = v1;
            const bool in = (pre || post) && sl >= 0 && sl < n;
            if (in) {
                kn = ld_at(mp.key_prev, (uint32_t)sl);
                const bool vel = pre ? lane >= 1 : lane < 32 + kExt;   // slots with a velocity entry
                if (vel) {
                    const uint32_t j = perm ? ld_stream_at(perm, (uint32_t)sl) : (uint32_t)sl;
                    const f4 *q = A.r + rec_index(j);
                    v1 = q[kRecTile];
                    v2 = q[2 * kRecTile];
                }
            }
            if (pre) {
                L.k[lane] = kn;
                if (lane >= 1) { L.v[lane - 1] = v1; L.w[lane - 1] = f2{ v2.z, v2.w }; }
            } else if (post) {
                L.k[W + kExt + 1 + (lane - 32)] = kn;
                if (lane < 32 + kExt) {
                    L.v[W + kExt + (lane - 32)] = v1;
                    L.w[W + kExt + (lane - 32)] = f2{ v2.z, v2.w };
                }
            }
        }
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int l = k * kWave + lane;
            L.k[kExt + 1 + l] = at[k].t >= 0 ? at[k].kold : kNoKey;
            if (at[k].t >= 0) {
                const f4 *q = A.r + rec_index(jj[k]);
                L.sm[l] = ca.sigvrmax[at[k].kold];
                const f4 v1 = __builtin_nontemporal_load(q + kRecTile);
                const f4 v2 = __builtin_nontemporal_load(q + 2 * kRecTile);
                p0[k] = __builtin_nontemporal_load(q);
                at[k].id = rec_label(p0[k].w, v2.x, v2.y);
                at[k].sx = at[k].sy = at[k].sz = 0.0f;
                if (SPIN) {
                    at[k].sx = ld_at(A.s[0], jj[k]);
                    at[k].sy = ld_at(A.s[1], jj[k]);
                    at[k].sz = ld_at(A.s[2], jj[k]);
                }
                L.v[kExt + l] = v1;
                L.w[kExt + l] = f2{ v2.z, v2.w };
                rpy[k] = v2.x;
                rpz[k] = v2.y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // the cells across the span's first and last boundary that it runs
        int pre_h = -1, pre_e = -1, post_h = -1;
        if (len > 0 && ws > 0) {
            const uint32_t k0 = L.k[ws - xb];
            if (L.k[ws - 1 - xb] == k0 && !oversized(k0, L.k[ws - kExt - 1 - xb], L.k[ws + kExt - xb])) {
                int lo = ws - kExt, hi = ws - 1;   // first slot with k0
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (L.k[mid - xb] == k0) hi = mid; else lo = mid + 1;
                }
                pre_h = lo;
                lo = ws + 1; hi = ws + kExt;       // first later slot with another key
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (L.k[mid - xb] == k0) lo = mid + 1; else hi = mid;
                }
                pre_e = lo;
            }
        }
        const int le = ws + len;
        if (len == W && le < n) {
            const uint32_t k1 = L.k[le - 1 - xb];
            if (L.k[le - xb] == k1 && !oversized(k1, L.k[le - kExt - 1 - xb], L.k[le + kExt - xb])) {
                int lo = le - kExt, hi = le - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (L.k[mid - xb] == k1) hi = mid; else lo = mid + 1;
                }
                post_h = lo;
            }
        }
        if (len > 0)
            wave_collide<APT>(traps, ts, T, one, t_first, ca, (uint32_t)step, ws, len, L, pre_h, pre_e,
                              post_h);
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int l = k * kWave + lane;
            if (at[k].t >= 0) {
                const f4 v = L.v[kExt + l];
                const f2 w = L.w[kExt + l];
                at[k].f.px = p0[k].x; at[k].f.py = p0[k].y; at[k].f.pz = p0[k].z;
                at[k].f.rpx = rem_value(p0[k].w);
                at[k].f.vx = v.x; at[k].f.vy = v.y; at[k].f.vz = v.z; at[k].f.rvx = v.w;
                at[k].f.rpy = rem_value(rpy[k]); at[k].f.rpz = rem_value(rpz[k]);
                at[k].f.rvy = w.x; at[k].f.rvz = w.y;
            }
        }
        (void)vb;
        if (one) {
            const TrapDev &tr = traps[t_first];
#pragma unroll
            for (int k = 0; k < APT; ++k) {
                const int i = ws + k * kWave + lane;
                if (at[k].t >= 0)
                    move_update<SPIN>(mp, tr, at[k], step, i, B, key_out, evap_pos, evap_vel, vals[k],
                                      lost_e[k], lost[k], moved[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < APT; ++k) {
                const int i = ws + k * kWave + lane;
                if (at[k].t >= 0)
                    move_update<SPIN>(mp, traps[at[k].t], at[k], step, i, B, key_out, evap_pos, evap_vel,
                                      vals[k], lost_e[k], lost[k], moved[k]);
            }
        }
    } else if (t_first == t_last) {
        // one trap: its parameters are wave-uniform (scalar loads and branches)
        const TrapDev &tr = traps[t_first];
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int i = b0 + k * BS + (int)threadIdx.x;
            at[k].t = i < n ? t_first : -1;
            if (i < n) {
                move_load<SPIN>(mp, A, tr, perm ? ld_stream_at(perm, (uint32_t)i) : (uint32_t)i, at[k]);
                if (mp.count_moved) at[k].kold = ld_at(mp.key_prev, (uint32_t)i);
            }
        }
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int i = b0 + k * BS + (int)threadIdx.x;
            if (at[k].t >= 0)
                move_update<SPIN>(mp, tr, at[k], step, i, B, key_out, evap_pos, evap_vel, vals[k],
                            lost_e[k], lost[k], moved[k]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int i = b0 + k * BS + (int)threadIdx.x;
            at[k].t = -1;
            if (i < n) {
                int t = t_first;
                while (t < t_last && ts[t + 1] <= i) ++t;
                at[k].t = t;
                move_load<SPIN>(mp, A, traps[t], perm ? ld_stream_at(perm, (uint32_t)i) : (uint32_t)i, at[k]);
                if (mp.count_moved) at[k].kold = ld_at(mp.key_prev, (uint32_t)i);
            }
        }
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const int i = b0 + k * BS + (int)threadIdx.x;
            if (at[k].t >= 0)
                move_update<SPIN>(mp, traps[at[k].t], at[k], step, i, B, key_out, evap_pos, evap_vel,
                            vals[k], lost_e[k], lost[k], moved[k]);
        }
    }

    // Per-trap block sums.  Atoms are trap-major, so a block covers traps
    // t_first..t_last, usually one.  The sums of t_first go to this block's
    // record (structure of arrays, read coalesced by k_obs_finalize); those of
    // later traps go to acc_over by atomics.  Only the block holding a trap's
    // first slot can do that, so each accumulator receives at most one term
    // and the totals do not depend on the order of execution.
    bool any = false;
    if (mp.count_moved) {   // per-wave masks and counts (published by the barrier below)
        uint32_t cm = 0, cd = 0;
#pragma unroll
        for (int k = 0; k < APT; ++k) {
            const uint64_t bm = __ballot(moved[k]), bd = __ballot(lost[k]);
            cm += (uint32_t)__popcll(bm);
            cd += (uint32_t)__popcll(bd);
            if ((threadIdx.x & (kWave - 1)) == 0) {
                // the 64-slot word of this wave's k-th atoms
                const int wd = COLL ? b0 / kWave + (int)(threadIdx.x / kWave) * APT + k
                                    : (b0 + k * BS) / kWave + (int)threadIdx.x / kWave;
                mp.mask_moved[wd] = bm;
                mp.mask_dead[wd] = bd;
                if (at[k].t >= 0) {   // slot 64 wd is below n
                    mp.key_samp[wd] = at[k].kold;
                    if (wd % (kResortTile / kWave) == 0) mp.key_tsamp[wd / (kResortTile / kWave)] = at[k].kold;
                }
            }
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            wcnt[threadIdx.x / kWave][0] = cm;
            wcnt[threadIdx.x / kWave][1] = cd;
        }
    }
#pragma unroll
    for (int k = 0; k < APT; ++k) any = any || lost[k];
    const bool any_lost = __syncthreads_or(any);
    if (mp.count_moved && threadIdx.x == 0) {
        uint32_t nm = 0, nd = 0;
        for (int w = 0; w < BS / kWave; ++w) { nm += wcnt[w][0]; nd += wcnt[w][1]; }
        mp.blk_moved[blk] = nm;
        mp.blk_dead[blk] = nd;
    }
    if (b0 < n) {
        const int nseg = t_last - t_first + 1;
        for (int sg = 0; sg < nseg; ++sg) {
            const int tt = t_first + sg;
            double v[4] = { 0.0, 0.0, 0.0, 0.0 }, e[4] = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
            for (int k = 0; k < APT; ++k)
                if (at[k].t == tt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) v[q] += vals[k][q];
            block_sum4<BS>(v, red);
            if (any_lost) {
#pragma unroll
                for (int k = 0; k < APT; ++k)
                    if (at[k].t == tt && lost[k]) { e[0] += lost_e[k]; e[1] += 1.0; }
                block_sum4<BS>(e, red);
            }
            if (threadIdx.x == 0) {
                const double r[kAccK] = { 0.0, v[3], v[0], v[1], v[2], e[0], e[1], 0.0 };
                if (sg == 0) {
#pragma unroll
                    for (int q = 0; q < kAccK; ++q) rec_sum[(size_t)q * rec_stride + blk] = r[q];
                    rec_trap[blk] = t_first;
                } else {
                    for (int q = 0; q < kAccK; ++q)
                        if (r[q] != 0.0) atomicAdd(&acc_over[tt * kAccK + q], r[q]);
                }
            }
        }
    }
}

// Grid (P, T): block (p, t) adds the records of every P-th move block of trap
// t (fixed order) and, when the next sort's trap ranges are asked for, the
// removal counts of every P-th move block before the trap; the last of the P
// blocks combines the partial sums in order, adds the boundary term, appends
// the trap's observables to the log and clears the per-step accumulators.
// The next sort puts the atoms removed now at the end, so trap t then starts
// at its first slot minus the atoms removed before it, and the alive count is
// the last trap's end minus all removed atoms.  P is about one block per 256
// move blocks of the largest trap; with P > 1 the last of a trap's blocks to
// finish (an atomic count) combines the others' partial sums, so no block
// waits for another.  Block
// (0, 0) writes the next step number to the other counter.  Blocks from P T
// on run the next incremental re-sort's first phase (resort_prep_block).
__global__ __launch_bounds__(256) void k_obs_finalize(
    int T, int P, int move_block, double half_mass, const int *__restrict__ trap_start,
    int *__restrict__ trap_next, const uint32_t *__restrict__ blk_dead,
    const uint64_t *__restrict__ mask_dead, const int *__restrict__ rec_trap,
    const double *__restrict__ rec_sum, int rec_stride, double *__restrict__ lvl2,
    unsigned int *__restrict__ cnt, double *__restrict__ acc_over,
    unsigned long long *__restrict__ coll_accum, const long long *__restrict__ step_cur,
    long long *__restrict__ step_next, dsmc_observables *__restrict__ obs_log, int obs_cap,
    int *__restrict__ n_next, ResortPrepArgs prep)
{
    static_assert(kPrepWords == 256, "prep blocks are finalize blocks");
    __shared__ double red[(256 / kWave) * kAccK];
    __shared__ uint32_t ptmp[256 / kWave + 1];
    __shared__ unsigned long long ptmp64[256 / kWave + 1];
    __shared__ bool fin;
    __shared__ double part[kMaxFinalizeP * kAccK];   // partial sums of a trap's P blocks
    if ((int)blockIdx.x >= P * T) {
        resort_prep_block((int)blockIdx.x - P * T, *prep.n_dev, prep.blk_moved, prep.blk_dead,
                          prep.per_tile, prep.mask_m, prep.mask_d, prep.pre_m, prep.pre_d,
                          prep.counts, prep.mov_v, prep.bcnt, ptmp, ptmp64);
        return;
    }
    const int p = (int)blockIdx.x % P, t = (int)blockIdx.x / P;
    const int tid = (int)threadIdx.x, stride = P * 256;
    const long long step = *step_cur;
    if (p == 0 && t == 0 && tid == 0) *step_next = step + 1;
    const int s0 = trap_start[t], s1 = trap_start[t + 1];
    // the per-step accumulators of the trap, loaded early (used by the block
    // that finalizes it)
    double over[kAccK];
    unsigned long long coll = 0ull;
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < kAccK; ++q) over[q] = acc_over[t * kAccK + q];
        coll = coll_accum[t];
    }
    double a[kAccK];
#pragma unroll
    for (int q = 0; q < kAccK; ++q) a[q] = 0.0;
    uint32_t dead = 0;   // removed atoms before s0 (this thread's share)
    if (trap_next) {
        // whole move blocks below s0, four counts per 16-byte load, eight
        // loads per thread in flight
        const int bd = s0 / move_block, nv = bd >> 2;
        const uint4 *v4 = (const uint4 *)blk_dead;
        for (int v0 = p * 256 + tid; v0 < nv; v0 += 8 * stride) {
            uint4 c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int v = v0 + u * stride;
                c[u] = v < nv ? v4[v] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) dead += c[u].x + c[u].y + c[u].z + c[u].w;
        }
        if (p == 0 && tid < (bd & 3)) dead += blk_dead[4 * nv + tid];
        // then the 64-slot words of the block holding s0, up to s0
        const int w0 = bd * (move_block / kWave), w1 = s0 / kWave, rem = s0 % kWave;
        for (int w = w0 + p * 256 + tid; w < w1 + (rem != 0); w += stride) {
            uint64_t m = mask_dead[w];
            if (w == w1) m &= ~0ull >> (kWave - rem);
            dead += (uint32_t)__popcll(m);
        }
    }
    if (s1 > s0) {
        const int bA = s0 / move_block, bB = (s1 - 1) / move_block;
        for (int b = bA + p * 256 + tid; b <= bB; b += stride) {
            double r[kAccK];   // loaded together with the record's trap
#pragma unroll
            for (int q = 0; q < kAccK; ++q) r[q] = rec_sum[(size_t)q * rec_stride + b];
            if (rec_trap[b] == t) {
#pragma unroll
                for (int q = 0; q < kAccK; ++q) a[q] += r[q];
            }
        }
    }
    a[kAccDead] += (double)dead;   // the records leave this sum zero
    static_assert(kAccK == 8, "eight accumulators");
    block_sum8<256>(a, red);
    // one block per trap, or the last of the trap's P blocks, finalizes it
    double f[kAccK];
#pragma unroll
    for (int q = 0; q < kAccK; ++q) f[q] = a[q];
    if (P == 1) {
        if (tid == 0) fin = true;
    } else if (tid == 0) {
#pragma unroll
        for (int q = 0; q < kAccK; ++q)
            __hip_atomic_store(&lvl2[((size_t)t * P + p) * kAccK + q], a[q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        fin = atomicAdd(&cnt[t], 1u) == (unsigned)(P - 1);
    }
    __syncthreads();
    if (!fin) return;   // whole block leaves together
    if (P > 1) {
        // the last of the trap's P blocks to finish: every partial loaded at
        // once, then thread q sums quantity q in block order (the same order
        // whichever block finishes last)
        __threadfence();
        for (int e = tid; e < P * kAccK; e += 256)
            part[e] = __hip_atomic_load(&lvl2[(size_t)t * P * kAccK + e], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (tid < kAccK) {
            double s = 0.0;
            for (int pp = 0; pp < P; ++pp) s += part[pp * kAccK + tid];
            part[tid] = s;   // no other thread reads entry tid
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll
            for (int q = 0; q < kAccK; ++q) f[q] = part[q];
            cnt[t] = 0u;
        }
    }
    if (tid != 0) return;
    for (int q = 0; q < kAccK; ++q) {
        f[q] += over[q];
        acc_over[t * kAccK + q] = 0.0;
    }
    coll_accum[t] = 0ull;
    dsmc_observables &o = obs_log[(step % obs_cap) * T + t];
    o.ek = half_mass * (f[kAccV2x] + f[kAccV2y] + f[kAccV2z]);
    o.ep = f[kAccEp];
    o.v2[0] = f[kAccV2x]; o.v2[1] = f[kAccV2y]; o.v2[2] = f[kAccV2z];
    o.e_evap = f[kAccEevap];
    // survivors of this step (the removed atoms are still in the slot range
    // the move covered; the next sort drops them)
    o.n_alive = (long long)(s1 - s0) - (long long)f[kAccNevap];
    o.n_evap = (long long)f[kAccNevap];
    o.collisions = (long long)coll;
    o.step = step;
    if (trap_next) {
        const int d0 = (int)f[kAccDead];
        trap_next[t] = s0 - d0;
        if (t == T - 1) {
            const int nn = s1 - d0 - (int)f[kAccNevap];
            trap_next[T] = nn;
            *n_next = nn;
        }
    }
}

// ---------------------------------------------------------------------------
// On-demand observables of the current FP32 state (dsmc_observe,
// dsmc_calculate_temp).  The slots hold the traps in order: each trap's alive
// atoms are contiguous except for removed atoms (dead key) that a move left
// in place.  The trap of a slot comes from its key.  First pass: the first
// and last alive slot of every trap (integer min / max, the same in any
// order of execution).  Second pass: one block per trap sums its alive slots
// in a fixed order, so the sums are reproducible run to run.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFBlock) void k_obs_bounds(const int *__restrict__ key_base, int T,
                                                        const uint32_t *__restrict__ key,
                                                        const int *__restrict__ n_dev,
                                                        uint32_t dead_key, int *__restrict__ first,
                                                        int *__restrict__ last)
{
    extern __shared__ int kb[];
    load_table(kb, key_base, T + 1);
    const int i = blockIdx.x * kFBlock + threadIdx.x;
    const int n = *n_dev;
    const uint32_t k = i < n ? key[i] : dead_key;
    const bool live = k != dead_key;
    const int t = live ? last_le(kb, T, (int)k) : -1;
    const uint64_t lm = __ballot(live);
    if (!lm) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int l0 = __ffsll((unsigned long long)lm) - 1, l1 = 63 - __clzll((long long)lm);
    const int t0 = __shfl(t, l0, kWave), t1 = __shfl(t, l1, kWave);
    if (t0 == t1) {   // the wave's alive slots lie in one trap
        if (lane == 0) {
            const int i0 = i - lane;
            atomicMin(&first[t0], i0 + l0);
            atomicMax(&last[t0], i0 + l1);
        }
    } else if (live) {
        atomicMin(&first[t], i);
        atomicMax(&last[t], i);
    }
}

constexpr int kObsBlock = 1024;

__global__ __launch_bounds__(kObsBlock) void k_obs_sums(const TrapDev *__restrict__ traps,
                                                        double mass, FusedPtrs s,
                                                        const uint32_t *__restrict__ key,
                                                        uint32_t dead_key,
                                                        const int *__restrict__ first,
                                                        const int *__restrict__ last,
                                                        double *__restrict__ acc)
{
    __shared__ double red[(kObsBlock / kWave) * 6];
    const int t = blockIdx.x;
    const TrapDev &tr = traps[t];
    // a trap without alive atoms keeps first = INT_MAX, last = -1: no slots
    // (a + threadIdx.x would overflow there, so the range is tested first)
    const int a = first[t], b = last[t];
    const int i0 = a <= b ? a + (int)threadIdx.x : b + 1;
    double v[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };   // v2x, v2y, v2z, U, n, (unused)
    for (int i = i0; i <= b; i += kObsBlock) {
        if (key[i] == dead_key) continue;
        FState f;
        rec_load(s.r, (uint32_t)i, f);
        const double X = (double)f.px + (double)f.rpx;
        const double Y = (double)f.py + (double)f.rpy;
        const double Z = (double)f.pz + (double)f.rpz;
        const double VX = (double)f.vx + (double)f.rvx;
        const double VY = (double)f.vy + (double)f.rvy;
        const double VZ = (double)f.vz + (double)f.rvz;
        v[0] += VX * VX;
        v[1] += VY * VY;
        v[2] += VZ * VZ;
        v[3] += potential_energy(tr, X, Y, Z);
        v[4] += 1.0;
    }
    block_sum<kObsBlock, 6>(v, red);
    if (threadIdx.x == 0) {
        double *o = acc + (size_t)t * kAccK;
        o[kAccEk] = 0.5 * mass * (v[0] + v[1] + v[2]);
        o[kAccEp] = v[3];
        o[kAccV2x] = v[0]; o[kAccV2y] = v[1]; o[kAccV2z] = v[2];
        o[kAccEevap] = 0.0;
        o[kAccNevap] = 0.0;
        o[kAccN] = v[4];
    }
}

void launch_observe_fused(const TrapDev *traps, const int *key_base, int n_traps, double mass,
                          const FusedPtrs &s, const uint32_t *key, const int *n_dev, int n_cap,
                          uint32_t dead_key, int *first, int *last, double *acc, hipStream_t st)
{
    fill_u32((uint32_t *)first, 0x7FFFFFFFu, n_traps, st, nullptr);
    fill_u32((uint32_t *)last, 0xFFFFFFFFu, n_traps, st, nullptr);
    hipLaunchKernelGGL(k_obs_bounds, dim3(fgrid(n_cap, kFBlock)), dim3(kFBlock),
                       sizeof(int) * (size_t)(n_traps + 1), st, key_base, n_traps, key, n_dev,
                       dead_key, first, last);
    DSMC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_obs_sums, dim3(n_traps), dim3(kObsBlock), 0, st, traps, mass, s, key,
                       dead_key, first, last, acc);
    DSMC_HIP(hipGetLastError());
}

// dst[i] = src[map[i]] for every FP32 state array.
__global__ __launch_bounds__(kFBlock) void k_gather_fused(FusedPtrs dst, FusedPtrs src,
                                                          const uint32_t *__restrict__ map,
                                                          const int *__restrict__ n_dev, int spin)
{
    const int i = blockIdx.x * kFBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t j = map[i];
    const f4 *q = src.r + rec_index(j);
    f4 *d = dst.r + rec_index((uint32_t)i);
#pragma unroll
    for (int c = 0; c < kRecPlanes; ++c) d[c * kRecTile] = q[c * kRecTile];
    if (spin)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst.s[c][i] = src.s[c][j];
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------

void launch_init_fused(const TrapDev *traps, const int *trap_start, int n_traps,
                       const InitParams &ip, const FusedPtrs &s, int n, hipStream_t st,
                       Profiler *prof)
{
    const size_t lds = sizeof(int) * (size_t)(n_traps + 1);
    Profiler *p_ = prof;
    if (p_ && p_->enabled) p_->begin(st, "init_fused", (double)n * 60.0);
    hipLaunchKernelGGL(k_init_fused, dim3(fgrid(n, kFBlock)), dim3(kFBlock), lds, st, traps,
                       trap_start, n_traps, ip, s, n);
    DSMC_HIP(hipGetLastError());
    if (p_ && p_->enabled) p_->end(st);
}

void launch_index_fused(const TrapDev *traps, const int *trap_start, int n_traps,
                        const FusedPtrs &s, uint32_t *key, const int *n_dev, int n_cap,
                        uint32_t dead_key, bool keep_dead, hipStream_t st, Profiler *prof)
{
    const size_t lds = sizeof(int) * (size_t)(n_traps + 1);
    if (prof && prof->enabled) prof->begin(st, "index_fused", (double)n_cap * 28.0);
    hipLaunchKernelGGL(k_index_fused, dim3(fgrid(n_cap, kFBlock)), dim3(kFBlock), lds, st, traps,
                       trap_start, n_traps, s, key, n_dev, dead_key, keep_dead ? 1 : 0);
    DSMC_HIP(hipGetLastError());
    if (prof && prof->enabled) prof->end(st);
}

void launch_table_fused(const uint32_t *key, const int *key_base, const TableParams &tp,
                        const int *n_in, int *n_out, int *trap_start, int n_cap, hipStream_t st,
                        Profiler *prof)
{
    const size_t lds = sizeof(int) * (size_t)(tp.n_traps + 1);
    if (prof && prof->enabled) prof->begin(st, "table_fused", (double)n_cap * 4.0);
    hipLaunchKernelGGL(k_table_fused, dim3(fgrid((long long)n_cap + 1, kFBlock * kTableItems)),
                       dim3(kFBlock),
                       lds, st, key, key_base, tp, n_in, n_out, trap_start);
    DSMC_HIP(hipGetLastError());
    if (prof && prof->enabled) prof->end(st);
}

void launch_cell_range_fused(const uint32_t *key, const int *n_dev, uint32_t dead_key,
                             uint32_t *cell_start, uint32_t *cell_end, int n_cap, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_range_fused, dim3(fgrid(n_cap, kFBlock)), dim3(kFBlock), 0, st, key,
                       n_dev, dead_key, cell_start, cell_end);
    DSMC_HIP(hipGetLastError());
}

void launch_collide_fused(const TrapDev *traps, const int *key_base, int n_traps,
                          const CollideParams &cp, const uint32_t *key, const uint32_t *perm,
                          const int *n_dev, const FusedPtrs &s, float *sigvrmax,
                          unsigned long long *coll_accum, const long long *step_dev, int n_cap,
                          hipStream_t st, Profiler *prof)
{
    const size_t lds = sizeof(int) * (size_t)(n_traps + 1);
    // DSMC_COLLIDE_WAVE_TRIALS: trial count above which a cell is run by a
    // whole wave (results are the same either way)
    const char *env = std::getenv("DSMC_COLLIDE_WAVE_TRIALS");
    const int wave_trials = env ? std::atoi(env) : kWaveTrialsDefault;
    if (prof && prof->enabled) prof->begin(st, "collide_fused", (double)n_cap * 4.0);
    // One wave per block when the context holds several traps: with trap
    // groups stepped on two streams, one-wave blocks find room beside the
    // other group's move blocks (batched step of two groups 157 -> 154 us on
    // MI355X, bitwise the same results); a single trap's collide runs about
    // 1 us faster in blocks of four waves.  Every block holds the trap table
    // in LDS, so beyond 511 traps (a 2 KB table) one-wave blocks would put
    // fewer than five waves on a SIMD, and four-wave blocks are used.
    if (n_traps > 1 && n_traps + 1 <= 512)
        hipLaunchKernelGGL(k_collide_fused<kWave>, dim3(fgrid(n_cap, kWave * kCollItems)),
                           dim3(kWave), lds, st, traps, key_base, n_traps, cp, key, perm, n_dev,
                           n_cap, s, sigvrmax, coll_accum, step_dev, wave_trials);
    else
        hipLaunchKernelGGL(k_collide_fused<kFBlock>, dim3(fgrid(n_cap, kFBlock * kCollItems)),
                           dim3(kFBlock), lds, st, traps, key_base, n_traps, cp, key, perm, n_dev,
                           n_cap, s, sigvrmax, coll_accum, step_dev, wave_trials);
    DSMC_HIP(hipGetLastError());
    if (prof && prof->enabled) prof->end(st);
}

double move_fused_bytes_per_atom(int spin)
{
    // read: perm 4 + the 12 floats of the state records, which carry the label
    // (+ 3 spin floats) + the last sort's key 4; write: the same state + key 4
    const double state = 12 * 4 + (spin ? 12 : 0);
    return 4.0 + state + 4.0 + state + 4.0;
}

template <int BS, int APT, bool COLL>
static void move_fused_bs(const TrapDev *traps, const MoveParams &mp, const FusedPtrs &src,
                          const FusedPtrs &dst, const uint32_t *perm, const int *trap_start,
                          const int *n_dev, uint32_t *key_out, float *evap_pos, float *evap_vel,
                          const ObsRecords &rec, const long long *step_dev, const CollArgs &ca,
                          int n_cap, hipStream_t st)
{
    const size_t lds = sizeof(int) * (size_t)(mp.n_traps + 1);
    if (mp.spin)
        hipLaunchKernelGGL((k_move_fused<BS, APT, true, COLL>), dim3(fgrid(n_cap, BS * APT)), dim3(BS),
                           lds, st, traps, mp, src, dst, perm, trap_start, n_dev, key_out,
                           evap_pos, evap_vel, rec.trap, rec.sum, rec.stride, rec.acc_over,
                           step_dev, ca);
    else
        hipLaunchKernelGGL((k_move_fused<BS, APT, false, COLL>), dim3(fgrid(n_cap, BS * APT)), dim3(BS),
                           lds, st, traps, mp, src, dst, perm, trap_start, n_dev, key_out,
                           evap_pos, evap_vel, rec.trap, rec.sum, rec.stride, rec.acc_over,
                           step_dev, ca);
}

int obs_record_stride(int n_cap) { return fgrid(n_cap, 64) + 1; }

static int wave_trials_setting()
{
    // DSMC_COLLIDE_WAVE_TRIALS: trial count above which a cell is run by a
    // whole wave (results are the same either way)
    const char *env = std::getenv("DSMC_COLLIDE_WAVE_TRIALS");
    return env ? std::atoi(env) : kWaveTrialsDefault;
}

void launch_move_fused(int block, int apt, const TrapDev *traps, const MoveParams &mp,
                       const FusedPtrs &src, const FusedPtrs &dst, const uint32_t *perm,
                       const int *trap_start, const int *n_dev, uint32_t *key_out,
                       float *evap_pos, float *evap_vel, const ObsRecords &rec,
                       const long long *step_dev, int n_cap, hipStream_t st, Profiler *prof,
                       const StepCollide *coll)
{
    CollArgs ca;
    std::memset(&ca, 0, sizeof ca);
    if (coll) {
        ca.cp = coll->cp;
        ca.sigvrmax = coll->sigvrmax;
        ca.coll_accum = coll->coll_accum;
        ca.wave_trials = wave_trials_setting();
        ca.dry = coll->dry;
    }
    if (prof && prof->enabled)
        prof->begin(st, coll ? "step_fused" : "move_fused", (double)n_cap * move_fused_bytes_per_atom(mp.spin));
#define DSMC_MOVE_CASE(BSV, APTV)                                                               \
    if (block == BSV && apt == APTV) {                                                          \
        if (coll)                                                                               \
            move_fused_bs<BSV, APTV, true>(traps, mp, src, dst, perm, trap_start, n_dev, key_out, \
                                           evap_pos, evap_vel, rec, step_dev, ca, n_cap, st);   \
        else                                                                                    \
            move_fused_bs<BSV, APTV, false>(traps, mp, src, dst, perm, trap_start, n_dev, key_out,\
                                            evap_pos, evap_vel, rec, step_dev, ca, n_cap, st);  \
    } else
    DSMC_MOVE_CASE(64, 1)
    DSMC_MOVE_CASE(128, 1)
    DSMC_MOVE_CASE(512, 1)
    DSMC_MOVE_CASE(1024, 1)
    DSMC_MOVE_CASE(64, 2)
    DSMC_MOVE_CASE(128, 2)
    DSMC_MOVE_CASE(256, 2)
    DSMC_MOVE_CASE(512, 2)
    {
        if (coll)
            move_fused_bs<256, 1, true>(traps, mp, src, dst, perm, trap_start, n_dev, key_out, evap_pos,
                                        evap_vel, rec, step_dev, ca, n_cap, st);
        else
            move_fused_bs<256, 1, false>(traps, mp, src, dst, perm, trap_start, n_dev, key_out, evap_pos,
                                         evap_vel, rec, step_dev, ca, n_cap, st);
    }
#undef DSMC_MOVE_CASE
    DSMC_HIP(hipGetLastError());
    if (prof && prof->enabled) prof->end(st);
}

// ---------------------------------------------------------------------------
// Edge pass of the fused step: the oversized cells across the boundaries
// between blocks of the step kernel (block = move_block sorted slots), i.e.
// cells that hold slots B - 1 and B for some B = q * move_block < n and have
// more than kExt slots on one side of B.  One lane per boundary checks it
// (four keys); each oversized cell is then run by the whole wave at the
// first boundary it crosses (its first slot in [B - move_block, B)), on the
// state in HBM through the permutation, as k_collide_fused does, before the
// step kernel gathers it.  Same trials as k_collide_fused for the same cell.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFBlock) void k_collide_edges(
    const TrapDev *__restrict__ traps, const int *__restrict__ key_base, int T, CollideParams cp,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ perm,
    const int *__restrict__ n_dev, FusedPtrs s, float *__restrict__ sigvrmax,
    unsigned long long *__restrict__ coll_accum, const long long *__restrict__ step_dev,
    int move_block, int wave_trials)
{
    extern __shared__ int kb[];
    load_table(kb, key_base, T + 1);
    const int lane = threadIdx.x & (kWave - 1);
    const int n = *n_dev;
    const long long q = (long long)blockIdx.x * kFBlock + threadIdx.x + 1;
    const long long B = q * move_block;
    bool over = false;
    uint32_t k = 0;
    if (B < n) {
        const int b = (int)B;
        k = key[b];
        const uint32_t kp = key[b - 1];
        const uint32_t kbf = b - kExt - 1 >= 0 ? key[b - kExt - 1] : kNoKey;
        const uint32_t kaf = b + kExt < n ? key[b + kExt] : kNoKey;
        over = kp == k && oversized(k, kbf, kaf);
        // run at the first boundary the cell crosses
        const int lo = b - move_block;
        if (over && lo > 0 && key[lo - 1] == k) over = false;
    }
    uint64_t todo = __ballot(over);
    if (!todo) return;   // whole wave
    const VelGlobal vg = { s.r, perm };
    const uint32_t step = (uint32_t)*step_dev;
    while (todo) {
        const int src = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1;
        const uint32_t kc = __shfl(k, src, kWave);
        const int b = (int)(((long long)blockIdx.x * kFBlock + (threadIdx.x & ~(kWave - 1)) + src + 1) * move_block);
        const int t = last_le(kb, T, (int)kc);
        const TrapDev &tr = traps[t];
        if ((int)(kc - (uint32_t)tr.key_base) >= tr.n_cells) continue;   // overflow cell: no collisions
        // first slot: probed 64 at a time below b (it lies in [b - move_block, b))
        int head = b - move_block;
        for (int j0 = b - kWave; j0 > b - move_block - kWave; j0 -= kWave) {
            const int j = j0 + lane;
            const uint64_t d = __ballot(j >= b - move_block && j < b && key[j] != kc);
            if (d) { head = j0 + 64 - __clzll((long long)d); break; }
        }
        // end: the first slot after b with another key, probed 64 at a time
        int end = n;
        for (int j0 = b + 1; j0 < n; j0 += kWave) {
            const int j = j0 + lane;
            const uint64_t d = __ballot(j < n && key[j] != kc);
            if (d) { end = j0 + __ffsll((unsigned long long)d) - 1; break; }
        }
        CellJob c;
        if (!cell_setup(tr, t, kc, cp, (double)sigvrmax[kc], head, end, c)) continue;
        int cnt;
        if (c.ncol > wave_trials) {
            const Trial0 none = { 0u, 0u, 0.0f, 0.0f };
            cnt = run_cell(cp, c, tr.stream, step, vg, sigvrmax, true, none);
        } else {
            cnt = 0;
            if (lane == 0) {
                Trial0 t0;
                trial0_load(cp, tr.stream, c.cell, step, c.n, c.i, vg, t0);
                cnt = run_cell(cp, c, tr.stream, step, vg, sigvrmax, false, t0);
            }
            vg.publish();
        }
        if (lane == 0 && cnt) atomicAdd(&coll_accum[t], (unsigned long long)cnt);
    }
}

void launch_collide_edges(const TrapDev *traps, const int *key_base, int n_traps,
                          const CollideParams &cp, const uint32_t *key, const uint32_t *perm,
                          const int *n_dev, const FusedPtrs &s, float *sigvrmax,
                          unsigned long long *coll_accum, const long long *step_dev, int move_block,
                          int n_cap, hipStream_t st, Profiler *prof)
{
    const int boundaries = std::max(1, fgrid(n_cap, move_block) - 1);
    const size_t lds = sizeof(int) * (size_t)(n_traps + 1);
    if (prof && prof->enabled) prof->begin(st, "collide_edges", 0.0);
    hipLaunchKernelGGL(k_collide_edges, dim3(div_up(boundaries, kFBlock)), dim3(kFBlock), lds,
                       st, traps, key_base, n_traps, cp, key, perm, n_dev, s, sigvrmax, coll_accum,
                       step_dev, move_block, wave_trials_setting());
    DSMC_HIP(hipGetLastError());
    if (prof && prof->enabled) prof->end(st);
}

void launch_obs_finalize(int n_traps, int max_trap_atoms, int move_block, double mass,
                         const int *trap_start, int *trap_next, const uint32_t *blk_dead,
                         const uint64_t *mask_dead, const ObsRecords &rec,
                         unsigned long long *coll_accum, const long long *step_cur,
                         long long *step_next, dsmc_observables *obs_log, int obs_cap,
                         int *n_next, const ResortPrepArgs &prep, hipStream_t st,
                         Profiler *prof)
{
    // blocks per trap: about 256 move-block records each (a single block
    // summing thousands of records is a chain of dependent load rounds)
    const int blocks = div_up(max_trap_atoms > 0 ? max_trap_atoms : 1, move_block) + 1;
    const int P = std::min(kMaxFinalizeP, div_up(blocks, 256));
    DSMC_LAUNCH(prof, "obs_finalize", (double)n_traps * 64.0, k_obs_finalize,
                P * n_traps + prep.blocks, 256, st, n_traps, P, move_block, 0.5 * mass,
                trap_start, trap_next, blk_dead, mask_dead, (const int *)rec.trap,
                (const double *)rec.sum, rec.stride, rec.lvl2, rec.cnt, rec.acc_over, coll_accum,
                step_cur, step_next, obs_log, obs_cap, n_next, prep);
}

void launch_gather_fused(const FusedPtrs &dst, const FusedPtrs &src, const uint32_t *map,
                         const int *n_dev, int n_cap, int spin, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "gather_fused", (double)n_cap * (4.0 + 2.0 * (48 + (spin ? 12 : 0))),
                k_gather_fused, fgrid(n_cap, kFBlock), kFBlock, st, dst, src, map, n_dev, spin);
}

}  // namespace dsmc
