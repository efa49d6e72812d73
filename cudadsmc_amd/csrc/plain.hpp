// plain.hpp -- kernel parameter structs and host launchers of the FP64 form.
#pragma once

#include "../../include/dsmc.h"
#include "dsmc_device.hpp"
#include "dsmc_internal.hpp"
#include "init_device.hpp"

namespace dsmc {

// Device pointers of one FP64 state buffer (structure of arrays).
struct PlainPtrs {
    double  *d[9];      // x, y, z, vx, vy, vz, ax, ay, az
    int     *id;        // atom label of each slot
    uint8_t *up;        // atomIsSpinUp
    double  *psi[4];    // psiUp re, im, psiDn re, im (spin runs only)
};

struct CollideParams {
    unsigned long long seed;
    double dt;
    double lpc;
    double sigma;
    int    pair_mode;
    int    pad;
};

struct SpinParams {
    double gs;
    double dt;          // motion step
    double dt_spin;     // dt / substeps
    int    substeps;
    int    interp;
};

struct FlipParams {
    unsigned long long seed;
    double mass;
    double gap;         // potential step of a flip per |B|: 2 mu (J/T)
    uint32_t step;      // Philox counter word 1
    int    pad;
};

struct DecayParams {
    double dt;          // lpc * dt
    double sigma;
    double mass;
};

struct EvapParams {
    double mass;
    double e_cut;       // J, <= 0: off
    int    spin_loss;
    int    pad;
};

void launch_index_plain(const TrapDev &t, const PlainPtrs &s, uint32_t *key, const int *n_dev,
                        int n_cap, hipStream_t st, Profiler *prof);
void launch_gather_plain(const PlainPtrs &dst, const PlainPtrs &src, const uint32_t *map,
                         const int *n_dev, int n_cap, int spin, hipStream_t st, Profiler *prof);
void launch_cell_table_plain(const uint32_t *key, const int *n_dev, int n_cap, int2 *se,
                             int *count, int *prefix, int *partial, int ncells,
                             hipStream_t st, Profiler *prof);
void launch_collide_plain(const TrapDev &t, const CollideParams &cp, const PlainPtrs &s,
                          double *sigvrmax, const int *prefix, int *coll_count,
                          unsigned long long *coll_total, int ncells, uint32_t step,
                          hipStream_t st, Profiler *prof);
void launch_move_plain(const TrapDev &t, double dt, const PlainPtrs &s, const int *n_dev,
                       int n_cap, hipStream_t st, Profiler *prof);
void launch_spin_plain(const TrapDev &t, const SpinParams &sp, const PlainPtrs &s,
                       const int *n_dev, int n_cap, hipStream_t st, Profiler *prof);
void launch_normalise_plain(const PlainPtrs &s, const int *n_dev, int n_cap, hipStream_t st,
                            Profiler *prof);
void launch_local_pops_plain(const TrapDev &t, const PlainPtrs &s, double *pops, const int *n_dev,
                             int n_cap, hipStream_t st, Profiler *prof);
void launch_flip_plain(const TrapDev &t, const FlipParams &fp, const PlainPtrs &s,
                       const double *pops, unsigned long long *flips, const int *n_dev, int n_cap,
                       hipStream_t st, Profiler *prof);
void launch_decay_plain(const TrapDev &t, const DecayParams &dp, const PlainPtrs &s,
                        const int *n_dev, int n_cap, hipStream_t st, Profiler *prof);
void launch_evap_tag_plain(const TrapDev &t, const EvapParams &ep, const PlainPtrs &s, int *tag,
                           double *evap_pos, double *evap_vel, double *evap_accum,
                           const int *n_dev, int n_cap, hipStream_t st, Profiler *prof);
int  observe_plain_blocks(int n_cap);
void launch_observe_plain(const TrapDev &t, double mass, const PlainPtrs &s, const int *n_dev,
                          int n_cap, double *partials, dsmc_observables *out,
                          unsigned long long *coll_accum, double *evap_accum, long long step,
                          hipStream_t st, Profiler *prof);
void launch_init_plain(const TrapDev &t, const InitParams &ip, const PlainPtrs &s, int n,
                       hipStream_t st, Profiler *prof);
void launch_accel_plain(const TrapDev &t, const PlainPtrs &s, int n, hipStream_t st,
                        Profiler *prof);

}  // namespace dsmc
