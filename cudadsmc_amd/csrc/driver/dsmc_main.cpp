// dsmc_main.cpp -- C++ host driver over libdsmc (include/dsmc.h) with the
// reference's command line and loop structure (CUDADSMC/src/main.cu:36-618):
//
//   cudadsmc [Nc N]                     the reference run: Ioffe-Pritchard trap,
//                                       Nc^3 cells (default 50), N atoms (default
//                                       1e4), dt = 1e-7 s, 10 motion steps per
//                                       collision step, 100 output frames over 5 ms
//   cudadsmc --quadrupole N Nc STEPS [fp32|fp64]
//                                       spherical quadrupole, B' = 2.5 T/m, 20 uK,
//                                       grid +-5 thermal radii, dt = 1e-6 s
//   options: --out FILE, --prints K, --nreal X (real atoms of the quadrupole
//            run, default 3e6 as h_N of initialSystemParameters.cuh:26)
//
// Output: with libdsmc_h5 (built where libhdf5 is installed) the HDF5 file of
// main.cu:202-577, default outputData.h5 (main.cu:33-34), one frame of the
// thirteen /atomData datasets per print.  A FILE not ending in .h5, or a build
// without libhdf5, gets a flat binary file instead:
//   "DSMCOUT1", int32 n_capacity, int32 n_cells_plus_1, then per frame
//   double time, int32 n_alive, double pos[n_alive][3], double vel[n_alive][3],
//   int32 atomID[n_alive], double E_kinetic, double E_potential, int64 collisions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dsmc.h"
#ifdef DSMC_WITH_H5
#include "dsmc_h5.h"
#endif

namespace {

constexpr double kMuB = 9.27400915e-24;
constexpr double kMRb = 1.443160648e-25;
constexpr double kPi = 3.14159265;
constexpr double kA = 5.3e-9;
constexpr double kKB = 1.3806503e-23;

void die(dsmc_ctx *ctx, dsmc_status st, const char *what)
{
    std::fprintf(stderr, "%s failed (%d): %s\n", what, (int)st, dsmc_last_error(ctx));
    std::exit(1);
}

#define CHECK(ctx, call)                          \
    do {                                          \
        dsmc_status st_ = (call);                 \
        if (st_ != DSMC_OK) die((ctx), st_, #call); \
    } while (0)

struct Frame {
    FILE *f = nullptr;
};

void write_frame(FILE *f, dsmc_ctx *ctx, double time)
{
    dsmc_info info;
    CHECK(ctx, dsmc_get_info(ctx, &info));
    const int n = info.n_alive;
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n);
    std::vector<int> id((size_t)n);
    CHECK(ctx, dsmc_copy_out(ctx, DSMC_FIELD_POS, pos.data(), pos.size() * 8));
    CHECK(ctx, dsmc_copy_out(ctx, DSMC_FIELD_VEL, vel.data(), vel.size() * 8));
    CHECK(ctx, dsmc_copy_out(ctx, DSMC_FIELD_ATOM_ID, id.data(), id.size() * 4));
    std::vector<dsmc_observables> obs((size_t)info.n_traps);
    CHECK(ctx, dsmc_observe(ctx, obs.data(), info.n_traps));
    long long coll = 0;
    std::fwrite(&time, 8, 1, f);
    std::fwrite(&n, 4, 1, f);
    std::fwrite(pos.data(), 8, pos.size(), f);
    std::fwrite(vel.data(), 8, vel.size(), f);
    std::fwrite(id.data(), 4, id.size(), f);
    std::fwrite(&obs[0].ek, 8, 1, f);
    std::fwrite(&obs[0].ep, 8, 1, f);
    std::fwrite(&coll, 8, 1, f);
}

}  // namespace

int main(int argc, const char *argv[])
{
    dsmc_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    dsmc_trap trap;
    std::memset(&trap, 0, sizeof trap);
#ifdef DSMC_WITH_H5
    std::string out = "outputData.h5";
#else
    std::string out = "outputData.bin";
#endif
    int prints = 100;
    long long total_steps = 0;
    double final_time = 5e-3;

    std::vector<std::string> pos_args;
    bool quad = false;
    double n_real = 3e6;   // h_N (initialSystemParameters.cuh:26)
    std::string precision = "fp64";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--prints" && i + 1 < argc) prints = std::atoi(argv[++i]);
        else if (a == "--nreal" && i + 1 < argc) n_real = std::atof(argv[++i]);
        else if (a == "--quadrupole") quad = true;
        else pos_args.push_back(a);
    }

    cfg.mass = kMRb;
    cfg.sigma = 8. * kPi * kA * kA;
    cfg.gs = 0.5;
    cfg.use_graph = 1;
    if (!quad) {
        // main.cu:45-69: CUDADSMC [Nc N]
        int Nc = 50, N = 10000;
        if (pos_args.size() == 2) {
            Nc = std::atoi(pos_args[0].c_str());
            N = std::atoi(pos_args[1].c_str());
        } else if (pos_args.size() > 2) {
            std::printf("Too many arguments supplied.\n");
            return 0;
        }
        trap.kind = DSMC_TRAP_IP;
        trap.n_atoms = N;
        trap.nc[0] = trap.nc[1] = trap.nc[2] = Nc;
        trap.half_width[0] = trap.half_width[1] = trap.half_width[2] = 1e-3;
        trap.B0 = 0.01; trap.dBdx = 20.; trap.d2Bdx2 = 40000.;
        trap.mu = 0.5 * 0.5 * kMuB;
        trap.T = 20e-6;
        trap.n_real = 3e6;
        cfg.precision = DSMC_FP64_PLAIN;
        cfg.pair_mode = DSMC_PAIR_REFERENCE;
        cfg.evaporation = DSMC_EVAP_SPIN;
        cfg.spin = 1;
        cfg.spin_substeps = 1;
        cfg.mcwf = 1;   // getLocalPopulations / flipAtoms / exponentialDecay (main.cu:457-516)
        cfg.lpc = 10;
        cfg.dt = 1e-7;
        cfg.seed = 1234;
        const int cpp = (int)std::ceil(final_time / (cfg.lpc * prints * cfg.dt));
        total_steps = (long long)cpp * cfg.lpc * prints;
        std::printf("lpc = %i, cpp = %i\n", cfg.lpc, cpp);
    } else {
        if (pos_args.size() < 3) {
            std::fprintf(stderr, "usage: cudadsmc --quadrupole N Nc STEPS [fp32|fp64]\n");
            return 2;
        }
        const int N = std::atoi(pos_args[0].c_str());
        const int Nc = std::atoi(pos_args[1].c_str());
        total_steps = std::atoll(pos_args[2].c_str());
        if (pos_args.size() > 3) precision = pos_args[3];
        const double T = 20e-6, dBdz = 2.5, mu = 0.5 * kMuB;
        const double rth = kKB * T / (mu * dBdz);
        trap.kind = DSMC_TRAP_QUADRUPOLE;
        trap.n_atoms = N;
        trap.nc[0] = trap.nc[1] = trap.nc[2] = Nc;
        trap.half_width[0] = trap.half_width[1] = trap.half_width[2] = 5.0 * rth;
        trap.dBdz = dBdz;
        trap.mu = mu;
        trap.T = T;
        trap.n_real = n_real;
        cfg.precision = precision == "fp32" ? DSMC_FP32_KAHAN : DSMC_FP64_PLAIN;
        cfg.pair_mode = DSMC_PAIR_UNIFORM;
        cfg.lpc = 1;
        cfg.dt = 1e-6;
        cfg.seed = 1;
        if (prints > total_steps) prints = (int)total_steps > 0 ? (int)total_steps : 1;
    }

    dsmc_ctx *ctx = nullptr;
    dsmc_status st = dsmc_create(&cfg, &trap, 1, &ctx);
    if (st != DSMC_OK) die(nullptr, st, "dsmc_create");
    CHECK(ctx, dsmc_init_distribution(ctx));

    const bool h5 = out.size() > 3 && out.compare(out.size() - 3, 3, ".h5") == 0;
    FILE *f = nullptr;
#ifdef DSMC_WITH_H5
    dsmc_h5 *hf = nullptr;
    auto frame = [&](double t) {
        if (!h5) return write_frame(f, ctx, t);
        if (dsmc_h5_write_frame(hf, ctx, t) != DSMC_OK) {
            std::fprintf(stderr, "HDF5 write failed: %s\n", dsmc_h5_last_error());
            std::exit(1);
        }
    };
    if (h5 && dsmc_h5_create(out.c_str(), "/atomData", &hf) != DSMC_OK) {
        std::fprintf(stderr, "%s\n", dsmc_h5_last_error());
        dsmc_destroy(ctx);
        return 1;
    }
#else
    if (h5) {
        std::fprintf(stderr, "this build has no HDF5 output (libhdf5 was not found)\n");
        dsmc_destroy(ctx);
        return 1;
    }
    auto frame = [&](double t) { write_frame(f, ctx, t); };
#endif
    if (!h5) {
        f = std::fopen(out.c_str(), "wb");
        if (!f) {
            std::fprintf(stderr, "cannot open %s for writing\n", out.c_str());
            dsmc_destroy(ctx);
            return 1;
        }
        const int cells1 = trap.nc[0] * trap.nc[1] * trap.nc[2] + 1;
        std::fwrite("DSMCOUT1", 1, 8, f);
        std::fwrite(&trap.n_atoms, 4, 1, f);
        std::fwrite(&cells1, 4, 1, f);
    }
    double time = 0.0;
    frame(time);

    const long long per_print = total_steps / prints;
    for (int i = 0; i < prints; ++i) {
        CHECK(ctx, dsmc_step(ctx, (int)per_print));
        time += per_print * cfg.dt;
        frame(time);
        std::printf("%%%i complete, t = %g s\n", i * 100 / prints, time);
    }
    std::printf("\n");
    if (f) std::fclose(f);
#ifdef DSMC_WITH_H5
    if (hf) dsmc_h5_close(hf);
#endif
    dsmc_destroy(ctx);
    return 0;
}
