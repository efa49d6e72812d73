// init_device.hpp -- per-atom initial-condition sampler shared by both
// numerical forms (setUp.cu:129-216, restated with Philox draws).
//
// Draw layout (Philox counter (atom, 0, attempt, (2 << 24) | draw), key
// stream = the trap's stream):
//   draws 0, 1   velocities, Box-Muller;
//   draws 2, 3   quadrupole radius (three uniforms) or box coordinates;
//   draws 4..35  quadrupole direction (Marsaglia 1972);
//   IP rejection attempt k: counter word 2 = k, draws 8, 9 (proposal), 10
//   (acceptance).
#pragma once

#include "dsmc_device.hpp"

namespace dsmc {

struct InitParams {
    unsigned long long seed;
    double mass;
    double gs;
    int    spin;
    int    pad;
};

__device__ inline void box_muller(U32x4 o, double &z0, double &z1)
{
    const double two_pi = 6.283185307179586;
    const double u1 = u01(o.x, o.y);
    const double u2 = u01(o.z, o.w);
    const double r = sqrt(-2.0 * log(u1));
    z0 = r * cos(two_pi * u2);
    z1 = r * sin(two_pi * u2);
}

__device__ inline void init_atom(const TrapDev &t, const InitParams &ip, uint32_t a,
                                 double &px, double &py, double &pz, double &vx, double &vy,
                                 double &vz)
{
    const unsigned long long seed = ip.seed;
    const uint32_t st = t.stream;
    px = 0.0; py = 0.0; pz = 0.0;
    if (t.kind == kTrapQuad) {
        // density exp(-mu B' rho / kT): radius Gamma(3, r_th) in (x, y, 2z)
        const double rth = kKB * t.T_init / (t.mu * t.dBdz);
        U32x4 o = draw(seed, st, a, 0u, 0, kPurposeInit, 2);
        const double u1 = u01(o.x, o.y), u2 = u01(o.z, o.w);
        o = draw(seed, st, a, 0u, 0, kPurposeInit, 3);
        const double u3 = u01(o.x, o.y);
        const double rad = -rth * (log(u1) + log(u2) + log(u3));
        double nx, ny, nz;
        sphere_point(seed, st, a, 0u, 0, kPurposeInit, 4, nx, ny, nz);
        px = rad * nx; py = rad * ny; pz = 0.5 * (rad * nz);
    } else if (t.kind == kTrapIP) {
        // selectAtomInThermalDistribution (setUp.cu:178-205)
        for (uint32_t k = 0; k < 100000u; ++k) {
            double n1, n2, n3, spare;
            box_muller(draw(seed, st, a, 0u, k, kPurposeInit, 8), n1, n2);
            box_muller(draw(seed, st, a, 0u, k, kPurposeInit, 9), n3, spare);
            const double cx = n1 * t.hw[0] / 3., cy = n2 * t.hw[1] / 3., cz = n3 * t.hw[2] / 3.;
            const D3 B = ip_field(t, cx, cy, cz);
            const double magB = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
            const double U = 0.5 * (magB - t.B0) * ip.gs * kMuB;
            const double Pr = exp(-U / kKB / t.T_init);
            const U32x4 o = draw(seed, st, a, 0u, k, kPurposeInit, 10);
            if (u01(o.x, o.y) < Pr) { px = cx; py = cy; pz = cz; break; }
        }
    } else {
        U32x4 o = draw(seed, st, a, 0u, 0, kPurposeInit, 2);
        const double ux = u01(o.x, o.y), uy = u01(o.z, o.w);
        o = draw(seed, st, a, 0u, 0, kPurposeInit, 3);
        const double uz = u01(o.x, o.y);
        px = -t.hw[0] + (2.0 * t.hw[0]) * ux;
        py = -t.hw[1] + (2.0 * t.hw[1]) * uy;
        pz = -t.hw[2] + (2.0 * t.hw[2]) * uz;
    }
    double z0, z1, z2, spare;
    box_muller(draw(seed, st, a, 0u, 0, kPurposeInit, 0), z0, z1);
    box_muller(draw(seed, st, a, 0u, 0, kPurposeInit, 1), z2, spare);
    vx = z0 * sqrt(kKB * t.T_axis[0] / ip.mass);
    vy = z1 * sqrt(kKB * t.T_axis[1] / ip.mass);
    vz = z2 * sqrt(kKB * t.T_axis[2] / ip.mass);
}

}  // namespace dsmc
