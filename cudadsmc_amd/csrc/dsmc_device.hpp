// dsmc_device.hpp -- device-side building blocks shared by every kernel:
// Philox4x32-10, uniform/integer draws, the trap field and force in FP64 and
// FP32, the cell index, and wave64 helpers.  Compiled for gfx950 only.
//
// The FP64 expressions keep the reference's operand order (cited per
// function) and the library is built with -ffp-contract=off, so FP64 results
// are reproducible bit for bit by a CPU build with the same rule.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsmc {

// Physical constants, include/deviceSystemParameters.cuh:14-19.
constexpr double kMuB  = 9.27400915e-24;
constexpr double kKB   = 1.3806503e-23;
constexpr double kHbar = 1.05457148e-34;
constexpr double kPiRef = 3.14159265;   // the reference's d_pi (deviceSystemParameters.cuh:16)

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Per-trap parameters, precomputed on the host.  Replaces the __constant__ /
// __device__ globals of deviceSystemParameters.cuh:12-31 (which every
// translation unit of the reference defines its own copy of).
// ---------------------------------------------------------------------------
enum TrapKind : int { kTrapIP = 0, kTrapQuad = 1, kTrapBox = 2 };

struct TrapDev {
    int    kind;
    int    nc[3];
    int    n_cells;        // nc product; local key n_cells is the overflow cell
    int    key_base;       // composite key of local cell 0
    int    atom_begin;     // initial first slot of this trap
    int    n_atoms0;       // initial atom count
    double hw[3];          // grid half width
    double gmin[3];        // -1.0 * hw          (collisions.cu:221-226)
    double len[3];         // (2.0 * hw) / nc    (collisions.cu:187-192)
    double inv_len[3];     // 1.0 / len, for the fast index path
    double cell_vol;       // len.x * len.y * len.z (collisions.cu:319)
    double B0, dBdx, d2Bdx2, dBdz;
    double mu;             // J/T
    double neg_mu_over_m;  // (-mu) / m          (moveAtoms.cu:102)
    double alpha;          // real atoms per test particle
    double coll_c1;        // alpha lpc dt / cell_vol (collision count per pair and sigma v_r)
    double coll_c1s;       // coll_c1 * sigma
    double sigvrmax0;      // initSigvrmax value (setUp.cu:289)
    double kT_cut;         // kB * T_cut (energy cut); <= 0 disables
    double eta_start, eta_end;
    long long eta_steps;
    double T_init;         // initial temperature (position sampler)
    double T_axis[3];      // initial temperature per velocity axis
    int    periodic;
    uint32_t stream;       // Philox stream of this trap
    // FP32 copies for the precision form
    float  f_hw[3];
    float  f_inv_len[3];   // 1 / len, for the FP32 index path
    float  f_B0, f_dBdx, f_d2Bdx2, f_dBdz;
    float  f_neg_mu_over_m;
    float  f_mu_dBdz;      // quadrupole: U = mu dBdz rho
    int    f32_index;      // every nc <= 256: the FP32 index path is exact (cell_of_f32)
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
struct U32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U32x4 philox(U32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32->64 multiply each (v_mad_u64_u32) instead of separate
        // high and low halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        U32x4 n = { hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0 };
        c = n;
    }
    return c;
}

enum : uint32_t { kPurposeCollide = 1u, kPurposeInit = 2u, kPurposeFlip = 3u };

// Key = (seed low word, seed high word + stream); every trap has its own
// stream, so its random numbers do not depend on which other traps share the
// launch.  Counter = (a, b, c, (purpose << 24) | draw).
__device__ __forceinline__ U32x4 draw(unsigned long long seed, uint32_t stream, uint32_t a,
                                      uint32_t b, uint32_t c, uint32_t purpose, uint32_t d)
{
    U32x4 ctr = { a, b, c, (purpose << 24) | d };
    return philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32) + stream);
}

// (k + 1/2) 2^-52 with 52 random bits: strictly inside (0,1); exact.
__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo)
{
    const uint64_t u = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
    return ((double)u + 0.5) * 0x1p-52;
}

// (k + 1/2) 2^-23 with 23 random bits: strictly inside (0,1); exact in FP32.
__device__ __forceinline__ float u01f(uint32_t r)
{
    return ((float)(r >> 9) + 0.5f) * 0x1p-23f;
}

__device__ __forceinline__ uint32_t uint_below(uint32_t r, uint32_t n)
{
    return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// Marsaglia (1972) point on the unit sphere; draws first..first+31, then +z.
__device__ inline void sphere_point(unsigned long long seed, uint32_t stream, uint32_t a,
                                    uint32_t b, uint32_t c, uint32_t purpose, uint32_t first,
                                    double &nx, double &ny, double &nz)
{
    for (uint32_t k = 0; k < 32; ++k) {
        U32x4 o = draw(seed, stream, a, b, c, purpose, first + k);
        double u1 = 2.0 * u01(o.x, o.y) - 1.0;
        double u2 = 2.0 * u01(o.z, o.w) - 1.0;
        double s = u1 * u1 + u2 * u2;
        if (s < 1.0) {
            double q = 2.0 * sqrt(1.0 - s);
            nx = u1 * q; ny = u2 * q; nz = 1.0 - 2.0 * s;
            return;
        }
    }
    nx = 0.0; ny = 0.0; nz = 1.0;
}

// ---------------------------------------------------------------------------
// Trap field, FP64 (operand order of moveAtoms.cu:93-157).
// ---------------------------------------------------------------------------
struct D3 { double x, y, z; };

__device__ __forceinline__ D3 ip_field(const TrapDev &t, double x, double y, double z)
{
    const double c = 0.5 * t.d2Bdx2;
    D3 B;
    B.x = (t.B0 * 0.0 + t.dBdx * x) + c * ((-x) * z);
    B.y = (t.B0 * 0.0 + t.dBdx * (-y)) + c * ((-y) * z);
    B.z = (t.B0 * 1.0 + t.dBdx * 0.0) + c * (z * z - 0.5 * (x * x + y * y));
    return B;
}

__device__ __forceinline__ double quad_rho(double x, double y, double z)
{
    return sqrt(x * x + y * y + 4.0 * (z * z));
}

// |B| and B/|B| (moveAtoms.cu:119-133).
__device__ __forceinline__ double field_dir(const TrapDev &t, double x, double y, double z, D3 &Bn)
{
    if (t.kind == kTrapIP) {
        D3 B = ip_field(t, x, y, z);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        Bn.x = B.x / len; Bn.y = B.y / len; Bn.z = B.z / len;
        return len;
    }
    if (t.kind == kTrapQuad) {
        double rho = quad_rho(x, y, z);
        if (rho > 0.0) { Bn.x = x / rho; Bn.y = y / rho; Bn.z = (-2.0 * z) / rho; }
        else { Bn.x = 0.0; Bn.y = 0.0; Bn.z = 1.0; }
        return t.dBdz * rho;
    }
    Bn.x = 0.0; Bn.y = 0.0; Bn.z = 1.0;
    return 0.0;
}

// updateAcc (moveAtoms.cu:93-116) and the quadrupole closed form.
__device__ __forceinline__ D3 accel(const TrapDev &t, double x, double y, double z, bool up)
{
    D3 a = { 0.0, 0.0, 0.0 };
    const double potential = t.neg_mu_over_m;
    if (t.kind == kTrapIP) {
        D3 B = ip_field(t, x, y, z);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        D3 Bn = { B.x / len, B.y / len, B.z / len };
        const double h = 0.5 * t.d2Bdx2;
        const double mh = (-0.5) * t.d2Bdx2;
        D3 dx = { t.dBdx - h * z, 0.0, mh * x };
        D3 dy = { 0.0, (-t.dBdx) - h * z, mh * y };
        D3 dz = { mh * x, mh * y, t.d2Bdx2 * z };
        double gx = dx.x * Bn.x + dx.y * Bn.y + dx.z * Bn.z;
        double gy = dy.x * Bn.x + dy.y * Bn.y + dy.z * Bn.z;
        double gz = dz.x * Bn.x + dz.y * Bn.y + dz.z * Bn.z;
        if (up) { a.x = potential * gx; a.y = potential * gy; a.z = potential * gz; }
        else {
            a.x = (-1.0 * potential) * gx; a.y = (-1.0 * potential) * gy; a.z = (-1.0 * potential) * gz;
        }
    } else if (t.kind == kTrapQuad) {
        double rho = quad_rho(x, y, z);
        if (rho > 0.0) {
            double g = (potential * t.dBdz) / rho;
            if (!up) g = -g;
            a.x = g * x; a.y = g * y; a.z = g * (4.0 * z);
        }
    }
    return a;
}

// Potential energy (energyChecker.py:121, energyComparison.py:53).
__device__ __forceinline__ double potential_energy(const TrapDev &t, double x, double y, double z)
{
    if (t.kind == kTrapIP) {
        D3 B = ip_field(t, x, y, z);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        return t.mu * (len - t.B0);
    }
    if (t.kind == kTrapQuad) return t.mu * (t.dBdz * quad_rho(x, y, z));
    return 0.0;
}

// findAtomIndex chain (collisions.cu:163-226; vectorMath.cuh:56-59): local cell
// index, n_cells for out-of-grid and NaN.
__device__ __forceinline__ int cell_of(const TrapDev &t, double x, double y, double z)
{
    const double p[3] = { x, y, z };
    int idx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double q = floor((p[d] - t.gmin[d]) / t.len[d]);
        if (!(q > -1.0 && q < (double)t.nc[d])) return t.n_cells;
        idx[d] = (int)q;
    }
    return idx[2] * t.nc[0] * t.nc[1] + idx[1] * t.nc[0] + idx[0];
}

// Same result as cell_of, cheaper: the quotient is formed with the reciprocal
// (|q' - q| <= 2 ulp(q)), and only when q' lies within 1e-9 of an integer -- far
// more than 2 ulp for any grid below 2^20 cells a side -- is the exact
// division of the reference formula evaluated.  NaN and infinities fail the
// range test in both branches.
__device__ __forceinline__ int cell_of_fast(const TrapDev &t, double x, double y, double z)
{
    const double p[3] = { x, y, z };
    int idx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double num = p[d] - t.gmin[d];
        double q = floor(num * t.inv_len[d]);
        const double frac = num * t.inv_len[d] - q;
        if (frac < 1e-9 || frac > 1.0 - 1e-9) q = floor(num / t.len[d]);
        if (!(q > -1.0 && q < (double)t.nc[d])) return t.n_cells;
        idx[d] = (int)q;
    }
    return idx[2] * t.nc[0] * t.nc[1] + idx[1] * t.nc[0] + idx[0];
}

// Same result as cell_of for the position p + rp (value + Kahan remainder),
// mostly in FP32: q' = (p + hw) / len from the FP32 value differs from the
// exact FP64 quotient of p + rp by less than 4 2^-24 nc (below 6.2e-5 of a
// cell for nc <= 256, which t.f32_index guarantees); only when q' lies
// within 1e-4 of an integer is the reference formula evaluated in FP64 on
// p + rp.  NaN fails the range test in both branches.
__device__ __forceinline__ int cell_of_f32(const TrapDev &t, const float (&p)[3], const float (&rp)[3])
{
    int idx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float qf = (p[d] + t.f_hw[d]) * t.f_inv_len[d];
        const float fl = floorf(qf);
        const float fr = qf - fl;
        if (fr < 1e-4f || fr > 1.0f - 1e-4f) {
            const double q = floor(((double)p[d] + (double)rp[d] - t.gmin[d]) / t.len[d]);
            if (!(q > -1.0 && q < (double)t.nc[d])) return t.n_cells;
            idx[d] = (int)q;
        } else {
            if (!(fl > -1.0f && fl < (float)t.nc[d])) return t.n_cells;
            idx[d] = (int)fl;
        }
    }
    return idx[2] * t.nc[0] * t.nc[1] + idx[1] * t.nc[0] + idx[0];
}

__device__ __forceinline__ double wrap_periodic(double x, double hw)
{
    if (x >= hw) x -= 2.0 * hw;
    else if (x < -hw) x += 2.0 * hw;
    return x;
}

// ---------------------------------------------------------------------------
// Trap field, FP32 (precision form).
// ---------------------------------------------------------------------------
struct F3 { float x, y, z; };

__device__ __forceinline__ F3 ip_field_f(const TrapDev &t, float x, float y, float z)
{
    const float c = 0.5f * t.f_d2Bdx2;
    F3 B;
    B.x = t.f_dBdx * x + c * ((-x) * z);
    B.y = t.f_dBdx * (-y) + c * ((-y) * z);
    B.z = t.f_B0 + c * (z * z - 0.5f * (x * x + y * y));
    return B;
}

// The FP32 form normalises with the hardware reciprocal square root (about
// 1 ulp): the force feeds a second-order integrator whose step error is many
// orders larger.
__device__ __forceinline__ float rsqrt_f(float v) { return __builtin_amdgcn_rsqf(v); }

__device__ __forceinline__ float field_dir_f(const TrapDev &t, float x, float y, float z, F3 &Bn)
{
    if (t.kind == kTrapIP) {
        F3 B = ip_field_f(t, x, y, z);
        const float b2 = B.x * B.x + B.y * B.y + B.z * B.z;
        const float inv = rsqrt_f(b2);
        Bn.x = B.x * inv; Bn.y = B.y * inv; Bn.z = B.z * inv;
        return b2 * inv;
    }
    if (t.kind == kTrapQuad) {
        const float r2 = x * x + y * y + 4.0f * (z * z);
        if (r2 > 0.0f) {
            const float inv = rsqrt_f(r2);
            Bn.x = x * inv; Bn.y = y * inv; Bn.z = (-2.0f * z) * inv;
            return t.f_dBdz * (r2 * inv);
        }
        Bn.x = 0.0f; Bn.y = 0.0f; Bn.z = 1.0f;
        return 0.0f;
    }
    Bn.x = 0.0f; Bn.y = 0.0f; Bn.z = 1.0f;
    return 0.0f;
}

// x^2 + y^2 + 4 z^2 of the quadrupole field's magnitude (B' sqrt of it)
__device__ __forceinline__ float quad_r2(float x, float y, float z)
{
    return fmaf(x, x, fmaf(y, y, 4.0f * (z * z)));
}

__device__ __forceinline__ F3 accel_f(const TrapDev &t, float x, float y, float z)
{
    F3 a = { 0.0f, 0.0f, 0.0f };
    if (t.kind == kTrapIP) {
        F3 B = ip_field_f(t, x, y, z);
        const float inv = rsqrt_f(B.x * B.x + B.y * B.y + B.z * B.z);
        F3 Bn = { B.x * inv, B.y * inv, B.z * inv };
        const float h = 0.5f * t.f_d2Bdx2;
        float gx = (t.f_dBdx - h * z) * Bn.x - h * x * Bn.z;
        float gy = (-t.f_dBdx - h * z) * Bn.y - h * y * Bn.z;
        float gz = -h * x * Bn.x - h * y * Bn.y + t.f_d2Bdx2 * z * Bn.z;
        float p = t.f_neg_mu_over_m;
        a.x = p * gx; a.y = p * gy; a.z = p * gz;
    } else if (t.kind == kTrapQuad) {
        const float r2 = quad_r2(x, y, z);
        if (r2 > 0.0f) {
            const float g = t.f_neg_mu_over_m * t.f_dBdz * rsqrt_f(r2);
            a.x = g * x; a.y = g * y; a.z = g * (4.0f * z);
        }
    }
    return a;
}

// Streaming stores for large outputs: marked non-temporal so that they do not
// stay dirty in the (per-XCD, non-coherent) L2, whose write-back at the end
// of a kernel would otherwise be paid before the next kernel can start.
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) { __builtin_nontemporal_store(v, p); }

// Element i of an array of at most 2^30 4-byte elements, addressed as the
// array base plus a 32-bit byte offset: the load or store then takes the base
// from scalar registers and one offset register (no 64-bit address per array).
template <typename T>
__device__ __forceinline__ T ld_at(const T *base, uint32_t i)
{
    static_assert(sizeof(T) == 4, "4-byte elements");
    return *(const T *)((const char *)base + (size_t)(i * 4u));
}
template <typename T>
__device__ __forceinline__ T ld_stream_at(const T *base, uint32_t i)
{
    static_assert(sizeof(T) == 4, "4-byte elements");
    return __builtin_nontemporal_load((const T *)((const char *)base + (size_t)(i * 4u)));
}
template <typename T>
__device__ __forceinline__ void st_stream_at(T *base, uint32_t i, T v)
{
    static_assert(sizeof(T) == 4, "4-byte elements");
    __builtin_nontemporal_store(v, (T *)((char *)base + (size_t)(i * 4u)));
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t lanemask_lt()
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Inclusive scan across the wave.
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan over a block of BLOCK threads (one value per thread).  Uses
// `tmp` (>= BLOCK/64 + 1 entries of LDS).  Returns the exclusive prefix and
// writes the block total to *total.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *tmp, T *total)
{
    constexpr int W = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    T inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) tmp[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int i = 0; i < W; ++i) { T c = tmp[i]; tmp[i] = run; run += c; }
        tmp[W] = run;
    }
    __syncthreads();
    T res = inc - v + tmp[w];
    *total = tmp[W];
    __syncthreads();
    return res;
}

// Sums K doubles over a block; the totals are valid in thread 0 only.
// `tmp` holds (BLOCK / 64) * K doubles of LDS.  The order of additions is
// fixed (butterfly inside the wave, then waves in order), so the result is
// reproducible run to run.
template <int BLOCK, int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *tmp)
{
    constexpr int W = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) tmp[w * K + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int ww = 0; ww < W; ++ww) s += tmp[ww * K + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Sums 4 doubles over a block; the totals are valid in thread 0 only.  In a
// wave the four are reduced together: at the first two exchanges (lanes l and
// l ^ 32, then l ^ 16) each lane keeps half of its values and sends the other
// half, so 7 exchanges do what 24 would for one value at a time; lane l ends
// with the wave total of quantity 2 * bit5(l) + bit4(l).  `tmp` holds
// (BLOCK / 64) * 4 doubles.  The order of additions is fixed.
template <int BLOCK>
__device__ __forceinline__ void block_sum4(double (&v)[4], double *tmp)
{
    constexpr int W = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0;
    double k0 = b5 ? v[2] : v[0], k1 = b5 ? v[3] : v[1];
    const double s0 = b5 ? v[0] : v[2], s1 = b5 ? v[1] : v[3];
    k0 += __shfl_xor(s0, 32, kWave);
    k1 += __shfl_xor(s1, 32, kWave);
    double k = b4 ? k1 : k0;
    k += __shfl_xor(b4 ? k0 : k1, 16, kWave);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) k += __shfl_xor(k, o, kWave);
    if ((lane & 15) == 0) tmp[w * 4 + (b5 ? 2 : 0) + (b4 ? 1 : 0)] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double t = 0.0;
            for (int ww = 0; ww < W; ++ww) t += tmp[ww * 4 + q];
            v[q] = t;
        }
    }
    __syncthreads();
}

// Sums 8 doubles over a block the same way (exchanges l ^ 32, 16, 8 halve
// the values each lane keeps; then l ^ 4, 2, 1): 10 exchanges per wave instead
// of 48.  Lane l ends with quantity 4 bit5(l) + 2 bit4(l) + bit3(l).  The
// totals are valid in thread 0; `tmp` holds (BLOCK / 64) * 8 doubles.
template <int BLOCK>
__device__ __forceinline__ void block_sum8(double (&v)[8], double *tmp)
{
    constexpr int W = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0;
    double a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double keep = b5 ? v[4 + q] : v[q], send = b5 ? v[q] : v[4 + q];
        a[q] = keep + __shfl_xor(send, 32, kWave);
    }
    double c[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double keep = b4 ? a[2 + q] : a[q], send = b4 ? a[q] : a[2 + q];
        c[q] = keep + __shfl_xor(send, 16, kWave);
    }
    double k = (b3 ? c[1] : c[0]) + __shfl_xor(b3 ? c[0] : c[1], 8, kWave);
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) k += __shfl_xor(k, o, kWave);
    if ((lane & 7) == 0) tmp[w * 8 + (b5 ? 4 : 0) + (b4 ? 2 : 0) + (b3 ? 1 : 0)] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double t = 0.0;
            for (int ww = 0; ww < W; ++ww) t += tmp[ww * 8 + q];
            v[q] = t;
        }
    }
    __syncthreads();
}

// Mask of the lanes whose 8-bit digit equals this lane's, among `active`.
// Lanes of `active` whose low B bits of v equal this lane's.
template <int B>
__device__ __forceinline__ uint64_t match_bits(uint32_t v, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const bool bit = (v >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

__device__ __forceinline__ uint64_t match_digit8(uint32_t d, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

}  // namespace dsmc
