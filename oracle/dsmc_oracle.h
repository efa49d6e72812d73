/*
 * dsmc_oracle.h -- CPU restatement of CUDADSMC's per-timestep DSMC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product (cudadsmc_amd/libdsmc.so) never links or calls it.
 *
 * Layout follows the reference exactly: AoS double3 state indexed by physical
 * atom id, plus an atomID indirection list that sorting permutes and
 * evaporation compacts (CUDADSMC/src/main.cu:88-140).  Every function names the
 * reference file:line it restates.  Arithmetic is FP64 with the reference's
 * operand order; build with -ffp-contract=off so that no multiply-add is fused.
 *
 * Deliberate, documented deviations from the reference (the reference's own
 * behaviour is not reproducible without CUDA):
 *   - RNG: cuRAND XORWOW (setUp.cu:71) is replaced by stateless
 *     Philox4x32-10 (Salmon et al., SC'11) keyed by the seed, with counters
 *     (id, step, trial, draw).  The GPU library uses the same counters, so
 *     the two agree draw for draw.
 *   - the isotropic post-collision direction (collisions.cu:537-546, three
 *     normalised Gaussians) is drawn with Marsaglia's 1972 disc method, which
 *     has the same (uniform-on-sphere) distribution and needs only +,*,sqrt.
 *   - cell length, grid extent, dt, loops-per-collision, magnetic moment and
 *     trap kind are run-time parameters instead of compile-time constants.
 *
 * Parity pins for each function are in tests/test_oracle.py.
 */
#ifndef DSMC_ORACLE_H
#define DSMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } orc_d3;
typedef struct { double re, im; } orc_c;

enum { ORC_TRAP_IP = 0, ORC_TRAP_QUAD = 1, ORC_TRAP_BOX = 2 };
enum { ORC_PAIR_UNIFORM = 0, ORC_PAIR_REFERENCE = 1 };

/* Physical constants (include/deviceSystemParameters.cuh:14-19). */
#define ORC_GS   0.5
#define ORC_MUB  9.27400915e-24
#define ORC_MRB  1.443160648e-25
#define ORC_PI   3.14159265
#define ORC_A    5.3e-9
#define ORC_KB   1.3806503e-23
#define ORC_HBAR 1.05457148e-34

typedef struct {
    int    kind;            /* ORC_TRAP_*                                     */
    double B0, dBdx, d2Bdx2;/* Ioffe-Pritchard (deviceSystemParameters.cuh:25-27) */
    double dBdz;            /* quadrupole radial gradient B' (T/m)            */
    double mu;              /* moment in the force a = -mu grad|B| / m (J/T)  */
    double mass;            /* kg                                             */
    double gs;              /* spin precession: theta = 0.5 gs muB |B| dt / hbar */
    int    nc[3];           /* cells per dimension                            */
    double half_width[3];   /* grid spans [-hw, +hw) per axis                 */
    double alpha;           /* real atoms per test particle                   */
    double sigma;           /* cross-section 8 pi a^2                         */
    double dt;              /* s                                              */
    int    lpc;             /* loops (motion substeps) per collision step     */
    int    pair_mode;       /* ORC_PAIR_*                                     */
    int    periodic;        /* wrap positions into the grid box (ORC_TRAP_BOX) */
    uint64_t seed;          /* Philox key                                     */
    uint32_t stream;        /* Philox stream of the trap (added to key word 1) */
} orc_params;

/* ---- Philox4x32-10 and the draws built on it ---------------------------- */
void     orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double   orc_u01(uint32_t hi, uint32_t lo);           /* (0,1), 52-bit grid    */
uint32_t orc_uint_below(uint32_t r, uint32_t n);     /* floor(r*n/2^32)       */

/* ---- trap field -------------------------------------------------------- */
orc_d3 orc_field(const orc_params *p, orc_d3 pos);          /* B(r)            */
orc_d3 orc_accel(const orc_params *p, orc_d3 pos, int up);  /* updateAcc       */
double orc_potential(const orc_params *p, orc_d3 pos);      /* U(r), J         */

/* ---- hot path, mirroring the reference host calls --------------------- */
void orc_index_atoms(const orc_params *p, const orc_d3 *pos, int *cellID,
                     const int *atomID, int n, int d1_quirk);
void orc_sort_arrays(int *cellID, int *atomID, int n, int nkeys);
void orc_cell_table(const int *cellID, int n, int ncells,
                    int *start_end /* 2*(ncells+1) */, int *count, int *prefix);
long long orc_collide(const orc_params *p, orc_d3 *vel, double *sigvrmax,
                      const int *prefix, int *collisionCount, const int *atomID,
                      int ncells, uint32_t step /* motion step of this collision step */);
void orc_move_atoms(const orc_params *p, orc_d3 *pos, orc_d3 *vel, orc_d3 *acc,
                    const uint8_t *spinUp, const int *atomID, int n);
void orc_evolve_wavefunction(const orc_params *p, const orc_d3 *pos, orc_c *psiUp,
                             orc_c *psiDn, const int *atomID, int n, double dt_spin);
void orc_normalise(orc_c *psiUp, orc_c *psiDn, const int *atomID, int n);
/* Majorana spin-flip model (wavefunction.cu:75-237); pops: 2 per label. */
void orc_local_populations(const orc_params *p, const orc_d3 *pos, const orc_c *psiUp,
                           const orc_c *psiDn, const int *atomID, int n, double *pops);
long long orc_flip_atoms(const orc_params *p, const orc_d3 *pos, orc_d3 *vel, const orc_c *psiUp,
                         const orc_c *psiDn, const double *pops, uint8_t *spinUp,
                         const int *atomID, int n, uint32_t step);
double orc_decoherence_tau(const orc_params *p, orc_d3 pos);
void orc_exponential_decay(const orc_params *p, const orc_d3 *pos, orc_c *psiUp, orc_c *psiDn,
                           const int *atomID, int n, double dt);
void orc_rotate_bloch(const orc_params *p, const orc_d3 *pos, orc_d3 *s,
                      const int *atomID, int n, double dt_spin);
void orc_evaporation_tag(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                         orc_d3 *evapPos, orc_d3 *evapVel, const orc_c *psiUp,
                         const orc_c *psiDn, const int *atomID, int *evapTag, int n,
                         double e_cut);
int  orc_remove_if(int *atomID, const int *evapTag, int n);
double orc_calculate_temp(const orc_params *p, const orc_d3 *vel, const int *atomID, int n);
void orc_energy(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                const int *atomID, int n, double out[3] /* Ek, Ep, sum v^2 */);
void orc_energy_spin(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                     const uint8_t *spinUp, const int *atomID, int n, double out[3]);

/* ---- initial conditions (setUp.cu:129-252, 287-292) ------------------- */
void orc_init_distribution(const orc_params *p, double T, const double Taniso[3],
                           orc_d3 *pos, orc_d3 *vel, orc_d3 *acc, orc_c *psiUp,
                           orc_c *psiDn, int *atomID, uint8_t *spinUp, int n);
double orc_sigvrmax_init(const orc_params *p, double T);

/* ---- whole loop (main.cu:410-524 without median, output and MCWF) ------ */
typedef struct {
    double seconds;         /* wall time of the step loop                     */
    long long atom_steps;   /* sum over motion substeps of N alive            */
    long long collisions;   /* total accepted collisions                      */
    int threads;            /* OpenMP threads used                            */
} orc_run_stats;

void orc_run(const orc_params *p, orc_d3 *pos, orc_d3 *vel, orc_d3 *acc,
             int *atomID, int *cellID, double *sigvrmax, int n, int n_collision_steps,
             uint32_t step0, orc_run_stats *stats);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
