/*
 * dsmc.h -- C ABI of the MI355X-native DSMC step library (libdsmc.so).
 *
 * The reference (CUDADSMC) has no library interface: its "operator API" is the
 * set of host wrappers and kernel launches that src/main.cu calls in a fixed
 * order (main.cu:410-524).  Each entry point below replaces one of those calls
 * and cites it.  All entry points are extern "C", take plain pointers, sizes
 * and POD structs, never throw, and return a dsmc_status; dsmc_last_error()
 * gives the message.  A context owns every device buffer and one HIP stream;
 * calls on a context are ordered on that stream and are synchronous with
 * respect to the host only where a result is returned to host memory.
 * Contexts are not thread-safe; different contexts may be used from
 * different host threads.
 *
 * Two numerical forms share this interface:
 *   DSMC_FP64_PLAIN  -- the drop-in: FP64 structure-of-arrays state, one kernel
 *                       per reference call, single trap.
 *   DSMC_FP32_KAHAN  -- FP32 state with Kahan-compensated position/velocity
 *                       updates, FP64 reductions, fused move/index/evaporate
 *                       kernel, any number of independent traps, optional
 *                       HIP-graph replay of the step loop.
 */
#ifndef DSMC_H
#define DSMC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSMC_ABI_VERSION 1

typedef enum dsmc_status {
    DSMC_OK      = 0,
    DSMC_EINVAL  = 1,   /* bad argument or configuration                     */
    DSMC_EHIP    = 2,   /* a HIP runtime call failed                          */
    DSMC_ENOMEM  = 3,   /* device or host allocation failed                   */
    DSMC_ESTATE  = 4,   /* call not valid in the context's current state      */
    DSMC_ENOTSUP = 5    /* not supported by this numerical form               */
} dsmc_status;

enum { DSMC_TRAP_IP = 0, DSMC_TRAP_QUADRUPOLE = 1, DSMC_TRAP_BOX = 2 };
enum { DSMC_FP64_PLAIN = 0, DSMC_FP32_KAHAN = 1 };
/* DSMC_PAIR_REFERENCE reproduces collisions.cu:320-322 and :400-413 exactly:
 * pairs floor(u (n-1)) (the last atom of a cell is never chosen) and the
 * trial count (int)(Mc * (ceil(X) / Mc)), which rounds to ceil(X) - 1 for some
 * Mc.  DSMC_PAIR_UNIFORM draws unordered pairs uniformly and runs ceil(X)
 * trials. */
enum { DSMC_PAIR_UNIFORM = 0, DSMC_PAIR_REFERENCE = 1 };
enum { DSMC_EVAP_SPIN = 1, DSMC_EVAP_ENERGY = 2 };      /* bit flags          */
/* FP32 form: how each step's sort is done.  Both give the same permutation
 * (the stable sort by cell key).  DSMC_SORT_INCREMENTAL re-sorts only the
 * atoms whose cell changed since the previous sort and merges them into the
 * rest; DSMC_SORT_FULL radix-sorts every key. */
enum { DSMC_SORT_INCREMENTAL = 0, DSMC_SORT_FULL = 1 };

/* One trap (one independent simulation).  Replaces the compile-time globals of
 * include/initialSystemParameters.cuh:14-33 and deviceSystemParameters.cuh:12-31
 * plus the CLI arguments "Nc N" of main.cu:45-69. */
typedef struct dsmc_trap {
    int    kind;            /* DSMC_TRAP_*                                        */
    int    n_atoms;         /* test particles (0: an empty trap in a batch)       */
    int    nc[3];           /* cells per dimension (cellsPerDimension)            */
    int    philox_stream;   /* added to the trap's index to give its Philox       */
                            /* stream: a trap's random numbers depend only on the */
                            /* seed and this stream, not on the other traps       */
    double half_width[3];   /* grid spans [-hw, +hw) (d_maxGridWidth, m)          */
    double B0, dBdx, d2Bdx2;/* Ioffe-Pritchard field (T, T/m, T/m^2)             */
    double dBdz;            /* quadrupole: |B| = dBdz sqrt(x^2 + y^2 + 4 z^2)     */
    double mu;              /* moment in a = -mu grad|B| / m (J/T)                */
    double T;               /* initial temperature (K)                            */
    double T_axis[3];       /* per-axis initial temperature, 0 = use T            */
    double n_real;          /* real atoms; alpha = n_real / n_atoms (main.cu:53)  */
    double eta_start;       /* energy cut E_cut = eta(step) kB T_cut; <= 0: off   */
    double eta_end;         /* eta ramps linearly eta_start -> eta_end ...        */
    double T_cut;           /* ... over eta_steps steps                           */
    long long eta_steps;
} dsmc_trap;

typedef struct dsmc_config {
    int    precision;       /* DSMC_FP64_PLAIN or DSMC_FP32_KAHAN                 */
    int    pair_mode;       /* DSMC_PAIR_*                                        */
    int    evaporation;     /* DSMC_EVAP_* flags, 0 = none                        */
    int    spin;            /* evolve spins (wavefunction.cu:11-73)               */
    int    spin_substeps;   /* spin rotations per motion step (>= 1)              */
    int    spin_interp;     /* 0: rotate at the pre-move position (reference);   */
                            /* 1: at substep midpoints along the drift            */
    int    lpc;             /* motion substeps per collision step (main.cu:168)   */
    int    use_graph;       /* FP32 form: replay dsmc_step through a HIP graph    */
    double dt;              /* motion time step (s)                               */
    double mass;            /* kg (d_mRb)                                         */
    double sigma;           /* cross-section (8 pi a^2, collisions.cu:312)        */
    double gs;              /* spin precession factor (d_gs)                      */
    unsigned long long seed;/* Philox key                                         */
    int    block_size;      /* per-atom kernels; 0 = default                      */
    int    obs_capacity;    /* steps kept in the observables ring; 0 = default    */
    int    sort_mode;       /* DSMC_SORT_* (FP32 form)                            */
    int    mcwf;            /* FP64 form with spin: Majorana spin-flip model of   */
                            /* wavefunction.cu:75-237 (populations, flips,        */
                            /* decoherence) in every collision step               */
    int    reserved[4];
} dsmc_config;

/* FP64 observables of one trap after one motion step. */
typedef struct dsmc_observables {
    double    ek;           /* sum 1/2 m v^2 over alive test particles (J)        */
    double    ep;           /* sum U(r) (J)                                       */
    double    v2[3];        /* sum vx^2, vy^2, vz^2 (m^2/s^2)                     */
    double    e_evap;       /* energy of the particles removed in this step (J)   */
    long long n_alive;      /* particles alive after this step (moved and kept);  */
                            /* n_alive + n_evap were moved in this step           */
    long long n_evap;       /* particles removed in this step                     */
    long long collisions;   /* accepted collisions in this step                   */
    long long step;         /* motion step index                                  */
} dsmc_observables;

typedef struct dsmc_info {
    int       precision;
    int       n_traps;
    int       n_capacity;   /* total test particles allocated                     */
    int       n_alive;      /* after the last completed call (host-synchronised)  */
    int       n_keys;       /* composite cell keys incl. overflow and dead keys   */
    int       key_bits;
    int       radix_passes;
    int       block_size;
    long long step;         /* motion steps taken                                 */
    long long collision_steps;
    long long obs_capacity; /* steps held by the observables log                  */
} dsmc_info;

/* Fields for dsmc_copy_out / dsmc_copy_in.  Per-atom fields are in slot order
 * of the current state buffer, double precision, AoS (x, y, z) triples, the
 * layout the reference copies to the host (main.cu:526-536).  They hold the
 * alive atoms only: in the FP32 form an atom removed by a move keeps its slot
 * (with the dead key) until the next sort drops it, and reads skip it. */
enum {
    DSMC_FIELD_POS = 0,        /* n x 3 double                                     */
    DSMC_FIELD_VEL,            /* n x 3 double                                     */
    DSMC_FIELD_ACC,            /* n x 3 double (FP64 form only)                    */
    DSMC_FIELD_ATOM_ID,        /* n int: label of the atom held in each slot       */
    DSMC_FIELD_CELL_ID,        /* n int: cell key of each slot (sorted after sort) */
    DSMC_FIELD_PERM,           /* n int: slot each sorted element came from        */
    DSMC_FIELD_SPIN,           /* n x 3 double Bloch vector (FP32 form)            */
                               /* or n x 4 double (psiUp re,im, psiDn re,im)       */
    DSMC_FIELD_EVAP_POS,       /* n_capacity x 3 double, indexed by atom label     */
    DSMC_FIELD_EVAP_VEL,       /* n_capacity x 3 double, indexed by atom label     */
    DSMC_FIELD_EVAP_TAG,       /* n int (FP64 form)                                */
    DSMC_FIELD_CELL_START_END, /* (n_keys) x 2 int  (FP64 form: (C+1) x 2)         */
    DSMC_FIELD_CELL_COUNT,     /* C+1 int (FP64 form)                              */
    DSMC_FIELD_CELL_PREFIX,    /* C+1 int (FP64 form)                              */
    DSMC_FIELD_SIGVRMAX,       /* n_keys double                                    */
    DSMC_FIELD_COLLISION_COUNT,/* C+1 int (FP64 form)                              */
    DSMC_FIELD_TRAP_START,     /* n_traps + 1 int                                  */
    DSMC_FIELD_POS_COMP,       /* n x 3 double: Kahan remainders (FP32 form), */
                               /* kept to 13 significant bits                      */
    DSMC_FIELD_VEL_COMP,       /* n x 3 double: Kahan remainders (FP32 form)       */
    DSMC_FIELD_OBS,            /* steps x n_traps dsmc_observables (whole log)     */
    DSMC_FIELD_SPIN_UP,        /* n int: 1 spin up, 0 spin down (atomIsSpinUp)     */
    DSMC_FIELD_LOCAL_POPS,     /* n_capacity x 2 double, by label (FP64 form, mcwf) */
    DSMC_FIELD_COUNT_
};

typedef struct dsmc_ctx dsmc_ctx;

/* ---- lifetime --------------------------------------------------------- */

/* cudaCalloc block of main.cu:117-140 and findRNGArrayLength (setUp.cu:18-31):
 * allocates every device buffer for n_traps traps.  FP64 form: n_traps == 1. */
dsmc_status dsmc_create(const dsmc_config *cfg, const dsmc_trap *traps, int n_traps,
                        dsmc_ctx **out);
/* cudaFree block of main.cu:596-615. */
dsmc_status dsmc_destroy(dsmc_ctx *ctx);
/* Message of the last failure on ctx (or of the last failed dsmc_create when
 * ctx is NULL).  Never NULL. */
const char *dsmc_last_error(const dsmc_ctx *ctx);
int         dsmc_abi_version(void);

/* ---- initial conditions ----------------------------------------------- */

/* h_initRNG + h_generateInitialDist + initSigvrmax (main.cu:176-189,
 * setUp.cu:32-165, 287-292): samples positions, velocities, accelerations,
 * field-aligned spins, atom labels, sigma*v_r,max of every cell, then indexes
 * the atoms (main.cu:191-195). */
dsmc_status dsmc_init_distribution(dsmc_ctx *ctx);
/* Replace the state of every atom (n = n_capacity, trap-major order) with host
 * data (AoS n x 3 doubles); resets labels to 0..n-1, sigma*v_r,max to its
 * initial value, the step counter to 0, and indexes the atoms. */
dsmc_status dsmc_set_state(dsmc_ctx *ctx, const double *pos, const double *vel, int n);
/* Replace spins: n x 3 Bloch vectors (FP32 form) or n x 4 spinors (FP64). */
dsmc_status dsmc_set_spin(dsmc_ctx *ctx, const double *spin, int n, int components);

/* ---- the step, one call per reference call ---------------------------- */

/* indexAtoms (collisions.cu:21-46, 163-226) without the unused median:
 * key[slot] = cell of the atom in that slot, out-of-grid -> overflow cell. */
dsmc_status dsmc_index_atoms(dsmc_ctx *ctx);
/* sortArrays (collisions.cu:275-287): stable sort of (key, slot).  The FP64
 * form then gathers the whole state into sorted order; the FP32 form keeps the
 * permutation and gathers inside the next dsmc_move_atoms. */
dsmc_status dsmc_sort_arrays(dsmc_ctx *ctx);
/* deviceMemset + cellStartandEndKernel + findNumberOfAtomsInCell +
 * thrust::exclusive_scan (main.cu:426-442). */
dsmc_status dsmc_build_cell_table(dsmc_ctx *ctx);
/* collide (collisions.cu:291-445). */
dsmc_status dsmc_collide(dsmc_ctx *ctx);
/* evolveWavefunction + normalise (main.cu:465-469, 518-521). */
dsmc_status dsmc_evolve_spin(dsmc_ctx *ctx);
/* Majorana spin-flip model (FP64 form, cfg.mcwf): getLocalPopulations
 * (wavefunction.cu:75-97, main.cu:457-462) before the motion substeps of a
 * collision step; flipAtoms (wavefunction.cu:99-154, main.cu:501-509), with
 * *n_flipped (may be NULL) the atoms flipped, and exponentialDecay over
 * lpc * dt (wavefunction.cu:175-219, main.cu:511-516) after them. */
dsmc_status dsmc_local_populations(dsmc_ctx *ctx);
dsmc_status dsmc_flip_atoms(dsmc_ctx *ctx, long long *n_flipped);
dsmc_status dsmc_exponential_decay(dsmc_ctx *ctx);
/* moveAtoms (moveAtoms.cu:23-49).  FP32 form: also evolves spins, computes the
 * next cell keys, applies evaporation and accumulates the FP64 observables. */
dsmc_status dsmc_move_atoms(dsmc_ctx *ctx);
/* calculateTemp (evaporation.cu:108-128): FP64 temperature of trap 0 (K). */
dsmc_status dsmc_calculate_temp(dsmc_ctx *ctx, double *T);
/* h_evaporationTag (evaporation.cu:19-106), FP64 form. */
dsmc_status dsmc_evaporation_tag(dsmc_ctx *ctx);
/* thrust::remove_if + numberOfAtoms update (main.cu:493-498): stable removal of
 * tagged atoms; *n_alive (may be NULL) receives the new count. */
dsmc_status dsmc_compact(dsmc_ctx *ctx, int *n_alive);
/* The body of main.cu:412-523 repeated: n_steps motion steps, with a collision
 * step every lpc motion steps (and, with cfg.mcwf, populations before and
 * flips and decoherence after its motion substeps).  Nothing is copied to the
 * host. */
dsmc_status dsmc_step(dsmc_ctx *ctx, int n_steps);
/* Block until all work on the context's stream is done. */
dsmc_status dsmc_sync(dsmc_ctx *ctx);
/* dsmc_step on n_ctx contexts of one device, n_steps each, submitted in
 * turns of `chunk` steps per context; each context's work stays on its own
 * stream, so independent groups of traps (one context each) overlap on the
 * GPU.  Asynchronous like dsmc_step; errors are reported on ctxs[0].  No
 * counterpart in the reference (it steps one trap); for callers that shard a
 * batch of independent traps over several contexts on one GPU. */
dsmc_status dsmc_step_group(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk);

/* ---- output surface (main.cu:526-579) --------------------------------- */

dsmc_status dsmc_get_info(dsmc_ctx *ctx, dsmc_info *info);
/* FP64 observables of every trap after the last motion step. */
dsmc_status dsmc_observe(dsmc_ctx *ctx, dsmc_observables *per_trap, int n_traps);
/* Copy one field to host memory; bytes must be at least the field size. */
dsmc_status dsmc_copy_out(dsmc_ctx *ctx, int field, void *host, size_t bytes);
/* Size in bytes of a field right now. */
dsmc_status dsmc_field_bytes(dsmc_ctx *ctx, int field, size_t *bytes);

/* ---- measurement ------------------------------------------------------ */

/* Run n_steps motion steps between two hipEvents recorded on the context's
 * stream; *ms receives the elapsed time. */
dsmc_status dsmc_time_steps(dsmc_ctx *ctx, int n_steps, double *ms);
/* dsmc_step_group between a hipEvent recorded on ctxs[0]'s stream (every
 * other stream waits for it) and the last stream's end; *ms receives the
 * elapsed time.  Synchronises. */
dsmc_status dsmc_time_steps_group(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk,
                                  double *ms);
/* Per-kernel timing: when enabled, every kernel launch is bracketed by
 * hipEvents on the context's stream and launched eagerly (no graph). */
dsmc_status dsmc_profile_enable(dsmc_ctx *ctx, int enable);
typedef struct dsmc_kernel_stat {
    char      name[48];
    long long launches;
    double    total_ms;
    double    bytes_per_launch;  /* algorithmic bytes of the last launch       */
} dsmc_kernel_stat;
dsmc_status dsmc_profile_read(dsmc_ctx *ctx, dsmc_kernel_stat *stats, int max_stats,
                              int *n_stats);
/* FP32 form: choose the fused move's block size and atoms per thread (shapes
 * of at least 512 atoms per block) by timing trial moves on this context's
 * current state (their results are discarded); 256 threads x 2 atoms is kept
 * unless a trial beats it by more than 5 %.  The next sort after a tune on sorted state is a full one.
 * block_size_out (may be NULL) receives the block size; the FP64 form keeps
 * 256.  No reference counterpart (the reference's block sizes are fixed). */
dsmc_status dsmc_autotune(dsmc_ctx *ctx, int *block_size_out);

/* ---- device primitives on caller-owned device memory ------------------ */
/* stream may be NULL (default stream).  Temporary storage is allocated and
 * freed inside the call. */

/* thrust::sort_by_key(int keys, int values) (collisions.cu:282-284): stable
 * LSD radix sort of n (key, value) pairs, keys in [0, 2^key_bits). */
dsmc_status dsmc_sort_by_key(unsigned int *d_keys, unsigned int *d_values, int n,
                             int key_bits, void *stream);
/* Allocation-free form of dsmc_sort_by_key: no counterpart in the reference.
 * It is for callers that cannot allocate device memory, such as code under
 * HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc.  Same result as dsmc_sort_by_key (stable, keys in
 * [0, 2^key_bits)), for n <= 1048576 pairs; larger n returns DSMC_EINVAL.
 * Its scratch is a fixed 8.5 MB device array inside the library, one copy
 * per device, held until the process ends; no call allocates or frees device
 * memory.  Asynchronous on `stream` (NULL: the default stream): the caller
 * synchronises before reading the result.  Every call on one device shares
 * that device's scratch, so calls that overlap in time on one device (other
 * streams, other host threads) race; calls ordered on one stream do not.
 * The first call on a device looks up the scratch's address there, so make
 * it before any graph capture that contains the call. */
dsmc_status dsmc_sort_by_key_noalloc(unsigned int *d_keys, unsigned int *d_values, int n,
                                     int key_bits, void *stream);
/* thrust::exclusive_scan<int> (main.cu:440-442). */
dsmc_status dsmc_exclusive_scan(const int *d_in, int *d_out, int n, void *stream);
/* thrust::remove_if(values, values + n, stencil, identity) (main.cu:493-496):
 * stable, in place; *n_out receives the new length. */
dsmc_status dsmc_remove_if(int *d_values, const int *d_stencil, int n, int *n_out,
                           void *stream);
/* Philox4x32-10 blocks out[4 i .. 4 i + 3] for counters (c0 + i, c1, c2, c3). */
dsmc_status dsmc_philox(unsigned int *d_out, int n, const unsigned int ctr[4],
                        const unsigned int key[2], void *stream);

/* Streaming device-to-device copy of `bytes` (a multiple of 16), launched
 * `repeats` times on the default stream; *ms = mean time per copy.  No
 * reference counterpart: a kernel of known bytes on which the HBM-traffic
 * counters are calibrated. */
dsmc_status dsmc_stream_copy(void *d_dst, const void *d_src, size_t bytes, int repeats, double *ms);
/* Gather of n_slots slots of FP32-form state records (48 bytes a slot, laid
 * out in tiles of 64 slots: three planes of 64 16-byte records) from d_src
 * into d_dst, slot i taking slot d_map[i]; both buffers hold whole tiles.
 * Launched `repeats` times on the default stream; *ms = mean time per launch.
 * No reference counterpart: a kernel of known bytes (52 read, 48 written a
 * slot) with the access pattern of the fused move's state gather, on which
 * the HBM-traffic counters are calibrated. */
dsmc_status dsmc_gather_copy(void *d_dst, const void *d_src, const unsigned int *d_map, int n_slots,
                             int repeats, double *ms);

/* Device memory helpers for hosts without their own allocator. */
dsmc_status dsmc_device_alloc(void **d_ptr, size_t bytes);
dsmc_status dsmc_device_free(void *d_ptr);
dsmc_status dsmc_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
dsmc_status dsmc_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* DSMC_H */
