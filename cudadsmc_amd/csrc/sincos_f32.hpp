// sincos_f32.hpp -- sine and cosine of an FP32 angle, for the spin substeps
// of the FP32 form (the rotation angle of each substep).
//
// Cody and Waite's argument reduction (1980): x = k pi/2 + r with pi/2 split
// into three floats (hi + mid + lo, their sum pi/2 to about 1e-22) and the
// products taken with fma, so r is accurate for |k| up to thousands; then the
// Taylor series of sin and cos on |r| <= pi/4 (coefficients 1/n!, truncated
// after r^9 and r^10: remainders below 2e-9 and 2e-10, far under FP32
// rounding), and the quadrant k mod 4.  Beyond kSinCosReduceMax (and for NaN
// or infinity) it falls back to the library's sincosf.  Plain C++ so that the
// same code is tested on the host.
#pragma once

#include <cmath>

#ifndef DSMC_HOST_DEVICE
#if defined(__HIPCC__)
#define DSMC_HOST_DEVICE __host__ __device__
#else
#define DSMC_HOST_DEVICE
#endif
#endif

namespace dsmc {

constexpr float kSinCosReduceMax = 8192.0f;

DSMC_HOST_DEVICE inline void sincos_f32(float x, float *sp, float *cp)
{
    if (!(fabsf(x) <= kSinCosReduceMax)) {
        sincosf(x, sp, cp);
        return;
    }
    const float kf = rintf(x * 0.636619772f);            // nearest multiple of pi/2
    float r = fmaf(-kf, 1.57079637e+00f, x);              // pi/2 = hi
    r = fmaf(-kf, -4.37113883e-08f, r);                   //      + mid
    r = fmaf(-kf, -1.71512451e-15f, r);                   //      + lo
    const float r2 = r * r;
    // sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9!
    const float ps = fmaf(r2, fmaf(r2, fmaf(r2, 2.75573192e-06f, -1.98412698e-04f), 8.33333333e-03f),
                          -1.66666667e-01f);
    const float s = fmaf(r * r2, ps, r);
    // cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
    const float pc = fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -2.75573192e-07f, 2.48015873e-05f), -1.38888889e-03f),
                                   4.16666667e-02f),
                          -0.5f);
    const float c = fmaf(r2, pc, 1.0f);
    // quadrant k mod 4: odd k swaps sine and cosine; the sine changes sign
    // for k = 2, 3 and the cosine for k = 1, 2 (mod 4)
    const int q = (int)kf;
    const bool swap = (q & 1) != 0;
    const float sv = swap ? c : s, cv = swap ? s : c;
    *sp = (q & 2) ? -sv : sv;
    *cp = ((q + 1) & 2) ? -cv : cv;
}

}  // namespace dsmc
