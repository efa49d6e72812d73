// label_bits.hpp -- the atom label inside the FP32 state records.
//
// The label (0 .. 2^31 - 1) of a slot travels in its state record: the low
// kLabelBits mantissa bits of each position remainder hold one part of it
// (rpx bits 0-10 of the label, rpy bits 11-21, rpz bits 22-30).  A position
// remainder is stored rounded to nearest-even at 23 - kLabelBits explicit
// mantissa bits, so a position is held to 24 + 13 = 37 significant bits (the
// velocity keeps its 48); the label then costs the move no bytes of its own.
// Infinity stays infinity and a NaN stays a NaN.  Plain C++ so that the same
// code is tested on the host.
#pragma once

#include <cstdint>

#ifndef DSMC_HOST_DEVICE
#if defined(__HIPCC__)
#define DSMC_HOST_DEVICE __host__ __device__
#else
#define DSMC_HOST_DEVICE
#endif
#endif

namespace dsmc {

constexpr int kLabelBits = 11;
constexpr uint32_t kLabelMask = (1u << kLabelBits) - 1u;

DSMC_HOST_DEVICE inline uint32_t f_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
DSMC_HOST_DEVICE inline float f_of_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

// r rounded to nearest-even with the low kLabelBits mantissa bits zero: the
// carry of the rounding increment runs into the exponent when the mantissa
// overflows, as a binary rounding does (the largest floats round to infinity)
// (without branches: a select of three results)
DSMC_HOST_DEVICE inline float rem_round(float r)
{
    const uint32_t u = f_bits(r);
    const uint32_t rounded = (u + (kLabelMask >> 1) + ((u >> kLabelBits) & 1u)) & ~kLabelMask;
    const uint32_t nan = (u & 0x80000000u) | 0x7FC00000u;   // quiet, payload cleared
    return f_of_bits(__builtin_isfinite(r) ? rounded : __builtin_isnan(r) ? nan : u);
}

// a position remainder as stored, from a remainder that rem_round has
// already rounded: one part of the label in its low bits
DSMC_HOST_DEVICE inline float rem_put(float rounded, uint32_t part)
{
    return f_of_bits(f_bits(rounded) | (part & kLabelMask));
}

// a position remainder as stored: rounded, with one part of the label
DSMC_HOST_DEVICE inline float rem_tag(float r, uint32_t part) { return rem_put(rem_round(r), part); }

// the value of a stored position remainder
DSMC_HOST_DEVICE inline float rem_value(float stored) { return f_of_bits(f_bits(stored) & ~kLabelMask); }

// the label of a slot from its three stored position remainders
DSMC_HOST_DEVICE inline int rec_label(float rpx, float rpy, float rpz)
{
    return (int)((f_bits(rpx) & kLabelMask) | ((f_bits(rpy) & kLabelMask) << kLabelBits) |
                 ((f_bits(rpz) & kLabelMask) << (2 * kLabelBits)));
}

// part k (0, 1, 2: for rpx, rpy, rpz) of a label
DSMC_HOST_DEVICE inline uint32_t label_part(int label, int k)
{
    return ((uint32_t)label >> (k * kLabelBits)) & kLabelMask;
}

}  // namespace dsmc
