// context.hip -- the C ABI of include/dsmc.h: context lifetime, buffer
// ownership, the step loop of both numerical forms, HIP-graph replay, per-kernel
// profiling and the output surface.
//
// Host-side call sequence per motion step (main.cu:412-523 without the dead
// median, the MCWF spin-flip model and the HDF5 output):
//   FP64 form:  [index, sort + gather, cell table, collide]  every lpc steps
//               spin, move, evaporation tag + compaction, FP64 observables
//   FP32 form:  sort, table, [collide], fused move           every step
// No call in the loop synchronises with the host.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dsmc.h"
#include "dsmc_internal.hpp"
#include "fused.hpp"
#include "plain.hpp"

using namespace dsmc;

// ---------------------------------------------------------------------------
// Runtime helpers
// ---------------------------------------------------------------------------
namespace dsmc {

void hip_check(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess) return;
    const int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? DSMC_ENOMEM
                                                                                : DSMC_EHIP;
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    throw Error(st, buf);
}

hipEvent_t Profiler::take()
{
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    DSMC_HIP(hipEventCreate(&e));
    return e;
}

void Profiler::begin(hipStream_t s, const char *name, double bytes)
{
    Rec r;
    r.name = name;
    r.bytes = bytes;
    r.b = take();
    r.e = take();
    DSMC_HIP(hipEventRecord(r.b, s));
    pending_.push_back(r);
}

void Profiler::end(hipStream_t s)
{
    if (pending_.empty()) return;
    DSMC_HIP(hipEventRecord(pending_.back().e, s));
}

void Profiler::collect()
{
    for (Rec &r : pending_) {
        DSMC_HIP(hipEventSynchronize(r.e));
        float ms = 0.f;
        DSMC_HIP(hipEventElapsedTime(&ms, r.b, r.e));
        Stat &s = stats_[r.name];
        s.launches += 1;
        s.total_ms += ms;
        s.bytes = r.bytes;
        pool_.push_back(r.b);
        pool_.push_back(r.e);
    }
    pending_.clear();
}

void Profiler::reset()
{
    collect();
    stats_.clear();
}

Profiler::~Profiler()
{
    for (Rec &r : pending_) { (void)hipEventDestroy(r.b); (void)hipEventDestroy(r.e); }
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

}  // namespace dsmc

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

struct dsmc_ctx {
    dsmc_config cfg{};
    std::vector<dsmc_trap> traps;
    std::vector<TrapDev> tdev;
    std::vector<int> key_base;      // n_traps + 1; last = dead key
    std::vector<int> atom_begin;    // n_traps + 1
    int T = 0, n_cap = 0, n_keys = 0, key_bits = 0, passes = 0, ncells = 0;
    int device = 0;
    uint32_t dead_key = 0;
    bool fp32 = false;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::string err;
    long long step = 0;             // motion steps taken
    long long coll_steps = 0;
    int cur = 0;                    // state buffer holding the current state
    int obs_cap = 0;
    int block = 256;                // FP32 move: threads per block
    int apt = 1;                    // FP32 move: atoms per thread
    bool have_sort = false;         // FP32: perm valid for the next move
    bool tbl_valid = false;         // FP32: trap_start and the alive count describe the
                                    // next sort's output (set by a sorted move's finalize)
    bool srt_valid = false;         // FP32: srt_k[srt_cur][i] is the key of slot i of the
                                    // current state (slots in sorted order of those keys)
    bool dead_in_state = false;     // FP32: the current state may hold atoms a move removed
                                    // (dead key in key_a, left in place until a sort)
    std::vector<void *> allocs;

    // common device buffers
    TrapDev *d_traps = nullptr;
    int *d_key_base = nullptr;
    uint32_t *key_a = nullptr, *key_b = nullptr, *val_a = nullptr, *val_b = nullptr;
    uint32_t *srt_k[2] = {nullptr, nullptr}, *srt_v[2] = {nullptr, nullptr};  // FP32 sort outputs
    int srt_cur = 0;                // FP32: srt_*[srt_cur] holds the last sort's result
    int graph_srt = 0;              // srt_cur at graph capture
    int graph_ssel = 0, graph_tsel = 0;   // step and trap-range buffers at graph capture
    bool prepped = false;           // FP32: the last finalize ran the re-sort's first phase
    bool prep_in_finalize = true;   // FP32: run it there (else in the re-sort itself)
    bool split_step = true;         // FP32: dsmc_step runs collide and move as two kernels
    const uint32_t *sorted_key = nullptr, *perm = nullptr;
    SortScratch sort;
    ResortScratch res;              // FP32: incremental re-sort
    double *sigvrmax = nullptr;     // FP64 form
    float *sigvrmax_f = nullptr;    // FP32 form: sigma v_r,max in float (half the bytes)
    dsmc_observables *obs_log = nullptr;  // (obs_cap + 1) * T records
    // device step counter, double-buffered: the kernels of a step read
    // step_buf[ssel] and its finalize writes the next value to the other one
    long long *step_buf = nullptr;
    int ssel = 0;
    long long *step_save = nullptr;
    long long *step_dev() const { return step_buf + ssel; }
    unsigned long long *coll_accum = nullptr;
    int *n_dev = nullptr;           // [0], [1]: alive counts; [2]: scratch
    // FP64 form
    PlainPtrs ps[2];
    int2 *start_end = nullptr;
    int *count = nullptr, *prefix = nullptr, *scan_partial = nullptr, *coll_count = nullptr;
    int *tag = nullptr;
    uint32_t *cmap = nullptr;
    double *evap_pos_d = nullptr, *evap_vel_d = nullptr;
    double *partials = nullptr, *evap_accum = nullptr;
    double *pops = nullptr;                  // spin-flip model: populations by label
    unsigned long long *flips = nullptr;     // spin-flip model: flips of the last call
    // FP32 form
    FusedPtrs fs[2];
    uint32_t *cell_start = nullptr, *cell_end = nullptr;
    uint32_t *blk_moved = nullptr, *blk_dead = nullptr;   // per move block, for the re-sort
    uint64_t *mask_moved = nullptr, *mask_dead = nullptr; // per 64 slots, for the re-sort
    uint32_t *key_samp = nullptr, *key_tsamp = nullptr;   // samples of the last sort's keys
    // trap ranges of the slot order, double-buffered: a sorted move's
    // finalize writes the next sort's ranges to ts_buf[tsel ^ 1]
    int *ts_buf[2] = { nullptr, nullptr };
    int tsel = 0;
    int *trap_start() const { return ts_buf[tsel]; }
    double *acc = nullptr, *acc2 = nullptr;
    ObsRecords rec{};
    int *obs_first = nullptr, *obs_last = nullptr;   // on-demand observables scratch
    int max_trap_atoms = 0;
    float *evap_pos_f = nullptr, *evap_vel_f = nullptr;
    // graph of two FP32 steps (parity 0 then 1)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    template <typename T_>
    T_ *alloc(size_t count_, bool zero = true)
    {
        void *p = nullptr;
        const size_t bytes = std::max<size_t>(count_ * sizeof(T_), 16);
        DSMC_HIP(hipMalloc(&p, bytes));
        allocs.push_back(p);
        if (zero) DSMC_HIP(hipMemsetAsync(p, 0, bytes, stream));
        return (T_ *)p;
    }

    ~dsmc_ctx()
    {
        // drain the stream before destroying a graph it may still be running
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static thread_local std::string g_create_error;

#define DSMC_API_BEGIN(ctx) \
    if (!(ctx)) return DSMC_EINVAL; \
    try {
#define DSMC_API_END(ctx)                                                   \
    return DSMC_OK;                                                         \
    }                                                                       \
    catch (const Error &e) { (ctx)->err = e.what(); return (dsmc_status)e.status; } \
    catch (const std::exception &e) { (ctx)->err = e.what(); return DSMC_EHIP; }

static void fail(int status, const std::string &msg) { throw Error(status, msg); }

// ---------------------------------------------------------------------------
// Parameter preparation
// ---------------------------------------------------------------------------
static TrapDev make_trapdev(const dsmc_config &c, const dsmc_trap &t, int index, int key_base,
                            int atom_begin)
{
    TrapDev d;
    std::memset(&d, 0, sizeof d);
    d.kind = t.kind;
    for (int k = 0; k < 3; ++k) {
        d.nc[k] = t.nc[k];
        d.hw[k] = t.half_width[k];
        d.gmin[k] = -1.0 * t.half_width[k];
        d.len[k] = (2.0 * t.half_width[k]) / (double)t.nc[k];
        d.inv_len[k] = 1.0 / d.len[k];
        d.f_hw[k] = (float)t.half_width[k];
        d.f_inv_len[k] = (float)d.inv_len[k];
    }
    d.f32_index = t.nc[0] <= 256 && t.nc[1] <= 256 && t.nc[2] <= 256;
    d.n_cells = t.nc[0] * t.nc[1] * t.nc[2];
    d.key_base = key_base;
    d.atom_begin = atom_begin;
    d.n_atoms0 = t.n_atoms;
    d.cell_vol = d.len[0] * d.len[1] * d.len[2];
    d.B0 = t.B0; d.dBdx = t.dBdx; d.d2Bdx2 = t.d2Bdx2; d.dBdz = t.dBdz;
    d.mu = t.mu;
    d.neg_mu_over_m = (-t.mu) / c.mass;
    // an empty trap (allowed in a batch) has no cell to collide in
    d.alpha = t.n_atoms > 0 ? t.n_real / (double)t.n_atoms : 0.0;
    d.coll_c1 = d.alpha * c.lpc * c.dt / d.cell_vol;
    d.coll_c1s = d.coll_c1 * c.sigma;
    double Tm = t.T;
    double Tax[3];
    for (int k = 0; k < 3; ++k) Tax[k] = t.T_axis[k] > 0.0 ? t.T_axis[k] : t.T;
    if (t.T_axis[0] > 0.0 || t.T_axis[1] > 0.0 || t.T_axis[2] > 0.0) Tm = (Tax[0] + Tax[1] + Tax[2]) / 3.0;
    d.sigvrmax0 = std::sqrt(16. * kKB * Tm / (kPiRef * c.mass)) * c.sigma;
    d.kT_cut = t.eta_start > 0.0 ? kKB * t.T_cut : 0.0;
    d.eta_start = t.eta_start; d.eta_end = t.eta_end; d.eta_steps = t.eta_steps;
    d.T_init = t.T;
    for (int k = 0; k < 3; ++k) d.T_axis[k] = Tax[k];
    d.periodic = t.kind == DSMC_TRAP_BOX;
    d.stream = (uint32_t)(index + t.philox_stream);
    d.f_B0 = (float)t.B0; d.f_dBdx = (float)t.dBdx; d.f_d2Bdx2 = (float)t.d2Bdx2;
    d.f_dBdz = (float)t.dBdz;
    d.f_neg_mu_over_m = (float)d.neg_mu_over_m;
    d.f_mu_dBdz = (float)(t.mu * t.dBdz);
    return d;
}

static void validate(const dsmc_config *c, const dsmc_trap *traps, int n_traps)
{
    if (!c || !traps) fail(DSMC_EINVAL, "null config or traps");
    if (c->precision != DSMC_FP64_PLAIN && c->precision != DSMC_FP32_KAHAN)
        fail(DSMC_EINVAL, "precision must be DSMC_FP64_PLAIN or DSMC_FP32_KAHAN");
    if (n_traps < 1) fail(DSMC_EINVAL, "n_traps must be >= 1");
    if (c->precision == DSMC_FP64_PLAIN && n_traps != 1)
        fail(DSMC_ENOTSUP, "the FP64 form runs one trap; use DSMC_FP32_KAHAN for batches");
    if (n_traps > fused_max_traps()) fail(DSMC_EINVAL, "too many traps");
    if (!(c->dt > 0.0) || !(c->mass > 0.0) || !(c->sigma >= 0.0))
        fail(DSMC_EINVAL, "dt and mass must be > 0, sigma >= 0");
    if (c->lpc < 1) fail(DSMC_EINVAL, "lpc must be >= 1");
    if (c->obs_capacity < 0 || (long long)c->obs_capacity * n_traps > (1LL << 28))
        fail(DSMC_EINVAL, "obs_capacity out of range");
    if (c->precision == DSMC_FP32_KAHAN && c->lpc != 1)
        fail(DSMC_ENOTSUP, "the FP32 form collides every motion step (lpc = 1)");
    if (c->spin && c->spin_substeps < 1) fail(DSMC_EINVAL, "spin_substeps must be >= 1");
    if ((c->evaporation & DSMC_EVAP_SPIN) && !c->spin)
        fail(DSMC_EINVAL, "spin-loss evaporation needs spin evolution");
    if (c->pair_mode != DSMC_PAIR_UNIFORM && c->pair_mode != DSMC_PAIR_REFERENCE)
        fail(DSMC_EINVAL, "bad pair_mode");
    if (c->sort_mode != DSMC_SORT_INCREMENTAL && c->sort_mode != DSMC_SORT_FULL)
        fail(DSMC_EINVAL, "bad sort_mode");
    if (c->mcwf && (c->precision != DSMC_FP64_PLAIN || !c->spin))
        fail(DSMC_ENOTSUP, "the spin-flip model runs in the FP64 form with spin evolution");
    const int bs = c->block_size;
    if (bs != 0 && bs != 64 && bs != 128 && bs != 256 && bs != 512 && bs != 1024)
        fail(DSMC_EINVAL, "block_size must be 0, 64, 128, 256, 512 or 1024");
    long long total = 0, keys = 0;
    for (int t = 0; t < n_traps; ++t) {
        const dsmc_trap &tr = traps[t];
        if (tr.kind < DSMC_TRAP_IP || tr.kind > DSMC_TRAP_BOX) fail(DSMC_EINVAL, "bad trap kind");
        if (tr.n_atoms < 0) fail(DSMC_EINVAL, "n_atoms < 0");
        for (int k = 0; k < 3; ++k) {
            if (tr.nc[k] < 1) fail(DSMC_EINVAL, "nc must be >= 1");
            if (!(tr.half_width[k] > 0.0)) fail(DSMC_EINVAL, "half_width must be > 0");
        }
        if (!(tr.T > 0.0)) fail(DSMC_EINVAL, "T must be > 0");
        if (tr.kind == DSMC_TRAP_QUADRUPOLE && !(tr.mu * tr.dBdz > 0.0))
            fail(DSMC_EINVAL, "quadrupole needs mu * dBdz > 0");
        if (tr.n_atoms > 0 && !(tr.n_real > 0.0)) fail(DSMC_EINVAL, "n_real must be > 0");
        if (tr.kind == DSMC_TRAP_BOX && (c->spin || (c->evaporation & DSMC_EVAP_SPIN)))
            fail(DSMC_EINVAL, "the periodic box has no field for spin evolution");
        total += tr.n_atoms;
        keys += (long long)tr.nc[0] * tr.nc[1] * tr.nc[2] + 1;
    }
    if (total < 1) fail(DSMC_EINVAL, "no atoms");
    if (total > (1LL << 30)) fail(DSMC_EINVAL, "more than 2^30 atoms");
    if (keys + 1 > (1LL << 31) - 1) fail(DSMC_EINVAL, "too many cells");
}

// ---------------------------------------------------------------------------
// Creation
// ---------------------------------------------------------------------------
static void build_context(dsmc_ctx *x, const dsmc_config *c, const dsmc_trap *traps, int n_traps)
{
    x->cfg = *c;
    x->traps.assign(traps, traps + n_traps);
    x->T = n_traps;
    x->fp32 = c->precision == DSMC_FP32_KAHAN;
    x->block = c->block_size ? c->block_size : 256;
    // DSMC_RESORT_PREP_IN_FINALIZE=0: the re-sort's first phase runs as a
    // kernel of the re-sort instead of in the finalize (results are the same)
    if (const char *e = std::getenv("DSMC_RESORT_PREP_IN_FINALIZE")) x->prep_in_finalize = std::atoi(e) != 0;
    // DSMC_STEP_FUSED=1: dsmc_step runs the collisions inside the move kernel
    // (the fused step) instead of the collide and move kernels one after the
    // other; the results are bitwise the same.  The fused step is slower on
    // MI355X (its collision code holds the kernel to four waves per SIMD,
    // where the move alone streams at six), so it is not the default.
    if (const char *e = std::getenv("DSMC_STEP_FUSED")) x->split_step = std::atoi(e) == 0;
    DSMC_HIP(hipGetDevice(&x->device));
    DSMC_HIP(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));

    int kb = 0, ab = 0;
    for (int t = 0; t < n_traps; ++t) {
        x->key_base.push_back(kb);
        x->atom_begin.push_back(ab);
        x->tdev.push_back(make_trapdev(*c, traps[t], t, kb, ab));
        kb += traps[t].nc[0] * traps[t].nc[1] * traps[t].nc[2] + 1;
        ab += traps[t].n_atoms;
    }
    x->dead_key = (uint32_t)kb;
    x->key_base.push_back(kb);
    x->atom_begin.push_back(ab);
    x->n_cap = ab;
    x->n_keys = kb + 1;
    x->ncells = traps[0].nc[0] * traps[0].nc[1] * traps[0].nc[2];
    int bits = 1;
    while ((1LL << bits) < (long long)x->n_keys) ++bits;
    x->key_bits = bits;
    x->passes = bits <= kRadixBits ? 1 : div_up(bits, kRadixBits);
    const long long per_rec = (long long)sizeof(dsmc_observables) * n_traps;
    x->obs_cap = c->obs_capacity > 0
                     ? c->obs_capacity
                     : (int)std::max<long long>(16, std::min<long long>(65536, (256LL << 20) / per_rec));

    const int N = x->n_cap;
    x->d_traps = x->alloc<TrapDev>(n_traps, false);
    DSMC_HIP(hipMemcpyAsync(x->d_traps, x->tdev.data(), sizeof(TrapDev) * n_traps,
                            hipMemcpyHostToDevice, x->stream));
    x->d_key_base = x->alloc<int>(n_traps + 1, false);
    DSMC_HIP(hipMemcpyAsync(x->d_key_base, x->key_base.data(), sizeof(int) * (n_traps + 1),
                            hipMemcpyHostToDevice, x->stream));
    x->key_a = x->alloc<uint32_t>(N);
    x->key_b = x->alloc<uint32_t>(N);
    x->val_a = x->alloc<uint32_t>(N);
    x->val_b = x->alloc<uint32_t>(N);
    sort_scratch_bind(x->sort, N, x->alloc<char>(sort_scratch_bytes(N)));
    if (x->fp32) x->sigvrmax_f = x->alloc<float>(x->n_keys);
    else x->sigvrmax = x->alloc<double>(x->n_keys);
    x->obs_log = x->alloc<dsmc_observables>((size_t)(x->obs_cap + 1) * n_traps);
    x->step_buf = x->alloc<long long>(2);
    x->step_save = x->alloc<long long>(1);
    x->coll_accum = x->alloc<unsigned long long>(n_traps);
    x->n_dev = x->alloc<int>(4);

    if (!x->fp32) {
        for (int b = 0; b < 2; ++b) {
            for (int k = 0; k < 9; ++k) x->ps[b].d[k] = x->alloc<double>(N);
            x->ps[b].id = x->alloc<int>(N);
            x->ps[b].up = x->alloc<uint8_t>(N);
            for (int k = 0; k < 4; ++k) x->ps[b].psi[k] = c->spin ? x->alloc<double>(N) : nullptr;
        }
        const int C1 = x->ncells + 1;
        x->start_end = x->alloc<int2>(C1);
        x->count = x->alloc<int>(C1);
        x->prefix = x->alloc<int>(C1);
        x->coll_count = x->alloc<int>(C1);
        x->scan_partial = x->alloc<int>(scan_partials(std::max(C1, N)) + 8);
        x->tag = x->alloc<int>(N);
        x->cmap = x->alloc<uint32_t>(N);
        x->evap_pos_d = x->alloc<double>(3 * (size_t)N);
        x->evap_vel_d = x->alloc<double>(3 * (size_t)N);
        x->partials = x->alloc<double>((size_t)observe_plain_blocks(N) * 5 + 8);
        x->evap_accum = x->alloc<double>(2);
        if (c->mcwf) {
            x->pops = x->alloc<double>(2 * (size_t)N);
            x->flips = x->alloc<unsigned long long>(1);
        }
    } else {
        for (int b = 0; b < 2; ++b) {
            x->fs[b].r = x->alloc<f4>(rec_elems(N));
            for (int k = 0; k < 3; ++k) x->fs[b].s[k] = c->spin ? x->alloc<float>(N) : nullptr;
        }
        resort_scratch_bind(x->res, N, x->alloc<char>(resort_scratch_bytes(N), false));
        // one entry per block of the smallest move block size, whole 2048-slot
        // tiles (never-written entries stay zero)
        x->blk_moved = x->alloc<uint32_t>((size_t)div_up(N, 2048) * (2048 / 64) + 64);
        x->blk_dead = x->alloc<uint32_t>((size_t)div_up(N, 2048) * (2048 / 64) + 64);
        x->mask_moved = x->alloc<uint64_t>(resort_mask_words(N));
        x->mask_dead = x->alloc<uint64_t>(resort_mask_words(N));
        x->key_samp = x->alloc<uint32_t>(resort_mask_words(N));
        // (+ 4: the locate reads the samples in 16-byte pieces)
        x->key_tsamp = x->alloc<uint32_t>(resort_mask_words(N) / (kResortTile / 64) + 4);
        x->srt_k[0] = x->key_b; x->srt_v[0] = x->val_b;
        x->srt_k[1] = x->alloc<uint32_t>(N); x->srt_v[1] = x->alloc<uint32_t>(N);
        x->cell_start = x->alloc<uint32_t>(x->n_keys, false);
        x->cell_end = x->alloc<uint32_t>(x->n_keys, false);
        x->ts_buf[0] = x->alloc<int>(n_traps + 1, false);
        x->ts_buf[1] = x->alloc<int>(n_traps + 1, false);
        DSMC_HIP(hipMemcpyAsync(x->trap_start(), x->atom_begin.data(), sizeof(int) * (n_traps + 1),
                                hipMemcpyHostToDevice, x->stream));
        x->acc = x->alloc<double>((size_t)n_traps * kAccK);
        x->rec.stride = obs_record_stride(N);
        x->rec.trap = x->alloc<int>((size_t)x->rec.stride);
        x->rec.sum = x->alloc<double>((size_t)x->rec.stride * kAccK);
        x->rec.acc_over = x->acc;
        x->rec.lvl2 = x->alloc<double>((size_t)n_traps * kMaxFinalizeP * kAccK);
        x->rec.cnt = x->alloc<unsigned int>((size_t)n_traps);
        for (int t = 0; t < n_traps; ++t) x->max_trap_atoms = std::max(x->max_trap_atoms, traps[t].n_atoms);
        x->acc2 = x->alloc<double>((size_t)n_traps * kAccK);
        x->obs_first = x->alloc<int>((size_t)n_traps);
        x->obs_last = x->alloc<int>((size_t)n_traps);
        x->evap_pos_f = x->alloc<float>(3 * (size_t)N);
        x->evap_vel_f = x->alloc<float>(3 * (size_t)N);
    }
    DSMC_HIP(hipMemcpyAsync(x->n_dev, &x->n_cap, sizeof(int), hipMemcpyHostToDevice, x->stream));
    DSMC_HIP(hipStreamSynchronize(x->stream));
}

// ---------------------------------------------------------------------------
// Step pieces: FP64 form
// ---------------------------------------------------------------------------
static CollideParams collide_params(const dsmc_ctx *x)
{
    CollideParams cp;
    std::memset(&cp, 0, sizeof cp);
    cp.seed = x->cfg.seed;
    cp.dt = x->cfg.dt;
    cp.lpc = (double)x->cfg.lpc;
    cp.sigma = x->cfg.sigma;
    cp.pair_mode = x->cfg.pair_mode;
    return cp;
}

static void plain_index(dsmc_ctx *x)
{
    launch_index_plain(x->tdev[0], x->ps[x->cur], x->key_a, x->n_dev, x->n_cap, x->stream, &x->prof);
}

// Sort keys in key_a with identity values; results in (sorted_key, perm).
static void run_sort(dsmc_ctx *x, const int *n_dev)
{
    const bool odd = (x->passes % 2) == 1;
    uint32_t *ok = odd ? x->key_b : x->key_a, *ov = odd ? x->val_b : x->val_a;
    uint32_t *ak = odd ? x->key_a : x->key_b, *av = odd ? x->val_a : x->val_b;
    radix_sort(x->key_a, nullptr, ak, av, ok, ov, n_dev, x->key_bits, x->sort, x->stream, &x->prof);
    x->sorted_key = ok;
    x->perm = ov;
}

static void plain_sort(dsmc_ctx *x)
{
    run_sort(x, x->n_dev);
    launch_gather_plain(x->ps[x->cur ^ 1], x->ps[x->cur], x->perm, x->n_dev, x->n_cap,
                        x->cfg.spin, x->stream, &x->prof);
    x->cur ^= 1;
}

static void plain_table(dsmc_ctx *x)
{
    launch_cell_table_plain(x->sorted_key, x->n_dev, x->n_cap, x->start_end, x->count, x->prefix,
                            x->scan_partial, x->ncells, x->stream, &x->prof);
}

static void plain_collide(dsmc_ctx *x)
{
    launch_collide_plain(x->tdev[0], collide_params(x), x->ps[x->cur], x->sigvrmax, x->prefix,
                         x->coll_count, x->coll_accum, x->ncells, (uint32_t)x->step, x->stream,
                         &x->prof);
    x->coll_steps++;
}

static void plain_spin(dsmc_ctx *x)
{
    SpinParams sp;
    sp.gs = x->cfg.gs;
    sp.dt = x->cfg.dt;
    sp.substeps = x->cfg.spin_substeps;
    sp.dt_spin = x->cfg.dt / x->cfg.spin_substeps;
    sp.interp = x->cfg.spin_interp;
    launch_spin_plain(x->tdev[0], sp, x->ps[x->cur], x->n_dev, x->n_cap, x->stream, &x->prof);
}

// Spin-flip model pieces (FP64 form).  The flip draws use the number of
// collision steps taken so far (this one included) as counter word 1.
static void plain_local_pops(dsmc_ctx *x)
{
    launch_local_pops_plain(x->tdev[0], x->ps[x->cur], x->pops, x->n_dev, x->n_cap, x->stream,
                            &x->prof);
}

static void plain_flip(dsmc_ctx *x)
{
    FlipParams fp;
    std::memset(&fp, 0, sizeof fp);
    fp.seed = x->cfg.seed;
    fp.mass = x->cfg.mass;
    fp.gap = 2.0 * x->tdev[0].mu;
    fp.step = (uint32_t)x->coll_steps;
    launch_flip_plain(x->tdev[0], fp, x->ps[x->cur], x->pops, x->flips, x->n_dev, x->n_cap,
                      x->stream, &x->prof);
}

static void plain_decay(dsmc_ctx *x)
{
    DecayParams dp;
    dp.dt = x->cfg.lpc * x->cfg.dt;
    dp.sigma = x->cfg.sigma;
    dp.mass = x->cfg.mass;
    launch_decay_plain(x->tdev[0], dp, x->ps[x->cur], x->n_dev, x->n_cap, x->stream, &x->prof);
}

static double e_cut_now(const dsmc_ctx *x, long long step)
{
    const TrapDev &t = x->tdev[0];
    if (!(x->cfg.evaporation & DSMC_EVAP_ENERGY) || t.kT_cut <= 0.0) return 0.0;
    const double f = t.eta_steps > 0 ? std::min((double)step / (double)t.eta_steps, 1.0) : 1.0;
    return (t.eta_start + (t.eta_end - t.eta_start) * f) * t.kT_cut;
}

static void plain_evap_tag(dsmc_ctx *x)
{
    EvapParams ep;
    ep.mass = x->cfg.mass;
    ep.e_cut = e_cut_now(x, x->step);
    ep.spin_loss = (x->cfg.evaporation & DSMC_EVAP_SPIN) ? 1 : 0;
    ep.pad = 0;
    launch_evap_tag_plain(x->tdev[0], ep, x->ps[x->cur], x->tag, x->evap_pos_d, x->evap_vel_d,
                          x->evap_accum, x->n_dev, x->n_cap, x->stream, &x->prof);
}

static void plain_compact(dsmc_ctx *x)
{
    compaction_map(x->tag, x->n_dev, x->n_cap, x->cmap, x->n_dev + 2, x->scan_partial, x->stream,
                   &x->prof);
    launch_gather_plain(x->ps[x->cur ^ 1], x->ps[x->cur], x->cmap, x->n_dev + 2, x->n_cap,
                        x->cfg.spin, x->stream, &x->prof);
    x->cur ^= 1;
    DSMC_HIP(hipMemcpyAsync(x->n_dev, x->n_dev + 2, sizeof(int), hipMemcpyDeviceToDevice,
                            x->stream));
}

static void plain_observe(dsmc_ctx *x, dsmc_observables *slot, long long step)
{
    launch_observe_plain(x->tdev[0], x->cfg.mass, x->ps[x->cur], x->n_dev, x->n_cap, x->partials,
                         slot, x->coll_accum, x->evap_accum, step, x->stream, &x->prof);
}

static void plain_step(dsmc_ctx *x)
{
    const int lpc = x->cfg.lpc;
    if (x->step % lpc == 0) {
        plain_index(x);
        plain_sort(x);
        plain_table(x);
        plain_collide(x);
        if (x->cfg.mcwf) plain_local_pops(x);
    }
    if (x->cfg.spin) plain_spin(x);
    launch_move_plain(x->tdev[0], x->cfg.dt, x->ps[x->cur], x->n_dev, x->n_cap, x->stream,
                      &x->prof);
    if (x->cfg.evaporation) {
        plain_evap_tag(x);
        plain_compact(x);
    }
    if (x->cfg.mcwf && (x->step + 1) % lpc == 0) {   // main.cu:501-516
        plain_flip(x);
        plain_decay(x);
    }
    if (x->cfg.spin && (x->step + 1) % lpc == 0)
        launch_normalise_plain(x->ps[x->cur], x->n_dev, x->n_cap, x->stream, &x->prof);
    plain_observe(x, x->obs_log + (x->step % x->obs_cap), x->step);
    x->step++;
}

// ---------------------------------------------------------------------------
// Step pieces: FP32 form (parity p: alive count in n_dev[p], state in fs[p])
// ---------------------------------------------------------------------------
static MoveParams move_params(const dsmc_ctx *x)
{
    MoveParams mp;
    std::memset(&mp, 0, sizeof mp);
    mp.dt = x->cfg.dt;
    mp.f_dt = (float)x->cfg.dt;
    mp.mass = x->cfg.mass;
    mp.f_half_mass = (float)(0.5 * x->cfg.mass);
    mp.gs = x->cfg.gs;
    mp.spin = x->cfg.spin;
    mp.substeps = std::max(1, x->cfg.spin_substeps);
    mp.interp = x->cfg.spin_interp;
    mp.spin_loss = (x->cfg.evaporation & DSMC_EVAP_SPIN) ? 1 : 0;
    mp.energy_cut = (x->cfg.evaporation & DSMC_EVAP_ENERGY) ? 1 : 0;
    mp.n_traps = x->T;
    mp.obs_cap = x->obs_cap;
    mp.dead_key = x->dead_key;
    mp.count_moved = 0;
    mp.blk_moved = x->blk_moved;
    mp.blk_dead = x->blk_dead;
    mp.mask_moved = x->mask_moved;
    mp.mask_dead = x->mask_dead;
    mp.key_samp = x->key_samp;
    mp.key_tsamp = x->key_tsamp;
    return mp;
}

static void fused_index(dsmc_ctx *x)
{
    launch_index_fused(x->d_traps, x->trap_start(), x->T, x->fs[x->cur], x->key_a, x->n_dev + x->cur,
                       x->n_cap, x->dead_key, x->dead_in_state, x->stream, &x->prof);
    x->tbl_valid = false;   // the new keys decide the trap ranges of the next sort
}

// Stable sort of the keys in key_a into the other sort buffer.  When the
// current state is in sorted order of the previous sort's keys (after a
// sorted move), only the atoms whose key changed are placed among the rest;
// the result is the same permutation.
static void fused_sort(dsmc_ctx *x)
{
    const int *n = x->n_dev + x->cur;
    uint32_t *ok = x->srt_k[x->srt_cur ^ 1], *ov = x->srt_v[x->srt_cur ^ 1];
    if (x->srt_valid && x->cfg.sort_mode == DSMC_SORT_INCREMENTAL && resort_supported(x->n_cap))
        resort_incremental(x->key_a, x->srt_k[x->srt_cur], n, x->dead_key, x->blk_moved,
                           x->blk_dead, x->mask_moved, x->mask_dead, x->key_samp, x->key_tsamp,
                           x->block * x->apt, x->prepped, x->res,
                           ok, ov, x->stream, &x->prof);
    else
        radix_sort(x->key_a, nullptr, x->res.alt_k, x->res.alt_v, ok, ov, n, x->key_bits,
                   x->sort, x->stream, &x->prof);
    x->prepped = false;
    x->srt_cur ^= 1;
    x->sorted_key = ok;
    x->perm = ov;
    x->have_sort = true;
    x->srt_valid = false;   // the state is not yet in the new order
}

static void fused_table(dsmc_ctx *x)
{
    TableParams tp;
    tp.n_traps = x->T;
    tp.dead_key = x->dead_key;
    launch_table_fused(x->sorted_key, x->d_key_base, tp, x->n_dev + x->cur, x->n_dev + (x->cur ^ 1),
                       x->trap_start(), x->n_cap, x->stream, &x->prof);
    x->tbl_valid = true;
}

// The trap ranges and alive count of the sorted order: kept by the previous
// sorted move's finalize, else scanned from the sorted keys.
static void fused_need_table(dsmc_ctx *x)
{
    if (x->have_sort && !x->tbl_valid) fused_table(x);
}

static void fused_collide(dsmc_ctx *x)
{
    fused_need_table(x);
    launch_collide_fused(x->d_traps, x->d_key_base, x->T, collide_params(x), x->sorted_key,
                         x->perm, x->n_dev + (x->cur ^ 1), x->fs[x->cur], x->sigvrmax_f,
                         x->coll_accum, x->step_dev(), x->n_cap, x->stream, &x->prof);
}

// coll: the fused step (the collisions of this step run inside the move
// kernel, those of cells across its block boundaries in an edge pass before
// it); needs a pending sort.
static void fused_move(dsmc_ctx *x, bool coll = false)
{
    // atoms removed by the previous move are dropped by a sort: a second move
    // without one would move (and test) them again
    if (!x->have_sort && x->dead_in_state)
        fail(DSMC_ESTATE, "the last move removed atoms: sort before moving again");
    fused_need_table(x);
    // after a sort the alive count of the sorted order is n_dev[cur ^ 1]
    const int *n = x->have_sort ? x->n_dev + (x->cur ^ 1) : x->n_dev + x->cur;
    MoveParams mp = move_params(x);
    // a sorted move also counts, per move block, the atoms whose key changed:
    // the input of the next step's incremental re-sort
    if (x->have_sort) {
        mp.count_moved = 1;
        mp.key_prev = x->sorted_key;
    }
    StepCollide sc;
    if (coll) {
        if (!x->have_sort) fail(DSMC_ESTATE, "the fused step needs a sorted state");
        sc.cp = collide_params(x);
        sc.sigvrmax = x->sigvrmax_f;
        sc.coll_accum = x->coll_accum;
        sc.dry = 0;
        launch_collide_edges(x->d_traps, x->d_key_base, x->T, sc.cp, x->sorted_key, x->perm, n,
                             x->fs[x->cur], x->sigvrmax_f, x->coll_accum, x->step_dev(),
                             64 * x->apt, x->n_cap, x->stream, &x->prof);   // one wave's span
    }
    launch_move_fused(x->block, x->apt, x->d_traps, mp, x->fs[x->cur], x->fs[x->cur ^ 1],
                      x->have_sort ? x->perm : nullptr, x->trap_start(), n, x->key_a,
                      x->evap_pos_f, x->evap_vel_f, x->rec, x->step_dev(), x->n_cap, x->stream,
                      &x->prof, coll ? &sc : nullptr);
    // a sorted move's finalize also leaves the next sort's trap ranges and
    // alive count (the removed atoms go to the end), and runs the first phase
    // of the incremental re-sort that follows
    ResortPrepArgs prep;
    x->prepped = x->prep_in_finalize && x->have_sort &&
                 x->cfg.sort_mode == DSMC_SORT_INCREMENTAL && resort_supported(x->n_cap);
    if (x->prepped)
        prep = resort_prep_args(x->res, n, x->block * x->apt, x->blk_moved, x->blk_dead,
                                x->mask_moved, x->mask_dead);
    launch_obs_finalize(x->T, x->max_trap_atoms, x->block * x->apt, x->cfg.mass, x->trap_start(),
                        x->have_sort ? x->ts_buf[x->tsel ^ 1] : nullptr, x->blk_dead,
                        x->mask_dead, x->rec, x->coll_accum, x->step_dev(),
                        x->step_buf + (x->ssel ^ 1), x->obs_log, x->obs_cap,
                        x->have_sort ? x->n_dev + x->cur : nullptr, prep, x->stream, &x->prof);
    x->ssel ^= 1;
    if (x->have_sort) x->tsel ^= 1;
    if (!x->have_sort)  // keep the count with the state it describes
        DSMC_HIP(hipMemcpyAsync(x->n_dev + (x->cur ^ 1), x->n_dev + x->cur, sizeof(int),
                                hipMemcpyDeviceToDevice, x->stream));
    // a sorted move leaves the state in sorted order of the last sort's keys
    // and has counted the changed keys; after an unsorted move the next sort
    // is a full one
    x->srt_valid = x->have_sort;
    x->tbl_valid = x->have_sort;
    x->have_sort = false;
    x->dead_in_state = x->cfg.evaporation != 0;
    x->cur ^= 1;
}

static void fused_step(dsmc_ctx *x)
{
    fused_sort(x);
    if (x->split_step) {   // the collide and the move as two kernels
        fused_collide(x);
        fused_move(x);
    } else {
        fused_need_table(x);
        fused_move(x, true);
    }
    x->coll_steps++;
    x->step++;
}

static bool graph_ok(const dsmc_ctx *x) { return x->fp32 && x->cfg.use_graph && !x->prof.enabled; }

// Destroy the captured graph.  Its earlier launches may still be in flight on
// the stream, so the stream is drained first: an executable graph is not
// destroyed while the device may still run it.
static void drop_graph(dsmc_ctx *x)
{
    if (!x->graph_exec && !x->graph) return;
    DSMC_HIP(hipStreamSynchronize(x->stream));
    if (x->graph_exec) (void)hipGraphExecDestroy(x->graph_exec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->graph_exec = nullptr;
    x->graph = nullptr;
}

static void fused_graph_steps(dsmc_ctx *x, int pairs)
{
    if (x->graph_exec && (x->graph_srt != x->srt_cur || x->graph_ssel != x->ssel ||
                          x->graph_tsel != x->tsel)) {
        // captured with the other sort (step, trap-range) buffer current
        drop_graph(x);
    }
    if (!x->graph_exec) {
        const int cur0 = x->cur;
        x->graph_srt = x->srt_cur;
        x->graph_ssel = x->ssel;
        x->graph_tsel = x->tsel;
        const long long step0 = x->step, cs0 = x->coll_steps;
        DSMC_HIP(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
        try {
            fused_step(x);
            fused_step(x);
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(x->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        DSMC_HIP(hipStreamEndCapture(x->stream, &x->graph));
        DSMC_HIP(hipGraphInstantiate(&x->graph_exec, x->graph, nullptr, nullptr, 0));
        x->cur = cur0;
        x->srt_cur = x->graph_srt;
        x->ssel = x->graph_ssel;
        x->tsel = x->graph_tsel;
        x->step = step0;
        x->coll_steps = cs0;
    }
    for (int k = 0; k < pairs; ++k) DSMC_HIP(hipGraphLaunch(x->graph_exec, x->stream));
    x->step += 2LL * pairs;
    x->coll_steps += 2LL * pairs;
}

static void do_steps(dsmc_ctx *x, int n)
{
    if (!x->fp32) {
        for (int k = 0; k < n; ++k) plain_step(x);
        return;
    }
    if (x->have_sort) fail(DSMC_ESTATE, "a sort is pending: call dsmc_move_atoms first");
    int k = 0;
    if (graph_ok(x)) {
        // the graph is captured from (and replayed at) state buffer 0 with a
        // valid previous sort
        const bool need_srt = x->cfg.sort_mode == DSMC_SORT_INCREMENTAL;
        while (k < n && (x->cur == 1 || (need_srt && !x->srt_valid))) { fused_step(x); ++k; }
        const int pairs = (n - k) / 2;
        if (pairs > 0) { fused_graph_steps(x, pairs); k += 2 * pairs; }
    }
    for (; k < n; ++k) fused_step(x);
}

// Sync and read the number of slots of the current state buffer (FP32: alive
// atoms plus the atoms the last move removed, which keep their slots until
// the next sort).
static int host_slots(dsmc_ctx *x)
{
    int n = 0;
    DSMC_HIP(hipStreamSynchronize(x->stream));
    DSMC_HIP(hipMemcpy(&n, x->fp32 ? x->n_dev + x->cur : x->n_dev, sizeof(int), hipMemcpyDeviceToHost));
    return n;
}

// FP32 form: the alive slots of the current state in slot order (all slots
// when no move has removed atoms since the last sort); returns the slot count.
static int fused_live(dsmc_ctx *x, std::vector<uint32_t> &live, std::vector<uint32_t> *keys = nullptr)
{
    const int n = host_slots(x);
    live.clear();
    std::vector<uint32_t> k;
    if (x->dead_in_state || keys) {
        k.resize((size_t)n);
        if (n) DSMC_HIP(hipMemcpy(k.data(), x->key_a, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost));
    }
    live.reserve((size_t)n);
    for (int i = 0; i < n; ++i)
        if (!x->dead_in_state || k[(size_t)i] != x->dead_key) live.push_back((uint32_t)i);
    if (keys) {
        keys->resize(live.size());
        for (size_t j = 0; j < live.size(); ++j) (*keys)[j] = k[live[j]];
    }
    return n;
}

// Sync and read the current alive count.
static int host_alive(dsmc_ctx *x)
{
    if (!x->fp32) return host_slots(x);
    if (!x->dead_in_state) return host_slots(x);
    std::vector<uint32_t> live;
    (void)fused_live(x, live);
    return (int)live.size();
}

// Clears one 64-bit counter (a kernel, not a memset: callers may capture it).
__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

static void observe_now(dsmc_ctx *x, std::vector<dsmc_observables> &out)
{
    out.assign(x->T, dsmc_observables{});
    if (!x->fp32) {
        plain_observe(x, x->obs_log + x->obs_cap, x->step);
        DSMC_HIP(hipStreamSynchronize(x->stream));
        DSMC_HIP(hipMemcpy(out.data(), x->obs_log + x->obs_cap, sizeof(dsmc_observables),
                           hipMemcpyDeviceToHost));
        return;
    }
    launch_observe_fused(x->d_traps, x->d_key_base, x->T, x->cfg.mass, x->fs[x->cur], x->key_a,
                         x->n_dev + x->cur, x->n_cap, x->dead_key, x->obs_first, x->obs_last,
                         x->acc2, x->stream);
    std::vector<double> a((size_t)x->T * kAccK);
    DSMC_HIP(hipStreamSynchronize(x->stream));
    DSMC_HIP(hipMemcpy(a.data(), x->acc2, sizeof(double) * a.size(), hipMemcpyDeviceToHost));
    for (int t = 0; t < x->T; ++t) {
        const double *v = a.data() + (size_t)t * kAccK;
        out[t].ek = v[kAccEk];
        out[t].ep = v[kAccEp];
        out[t].v2[0] = v[kAccV2x]; out[t].v2[1] = v[kAccV2y]; out[t].v2[2] = v[kAccV2z];
        out[t].n_alive = (long long)v[kAccN];
        out[t].step = x->step;
    }
}

// Initial per-cell state and labels (initSigvrmax, setUp.cu:287-292).  Every
// transfer goes through the context's own (non-blocking) stream, then waits.
static void reset_cells(dsmc_ctx *x)
{
    std::vector<double> sv(x->n_keys, 0.0);
    for (int t = 0; t < x->T; ++t)
        for (int k = 0; k <= x->tdev[t].n_cells; ++k) sv[x->key_base[t] + k] = x->tdev[t].sigvrmax0;
    const long long zero = 0;
    int nn[2] = { x->n_cap, x->n_cap };
    hipStream_t st = x->stream;
    std::vector<float> svf;
    if (x->fp32) {
        svf.assign(sv.begin(), sv.end());
        DSMC_HIP(hipMemcpyAsync(x->sigvrmax_f, svf.data(), sizeof(float) * svf.size(),
                                hipMemcpyHostToDevice, st));
    } else {
        DSMC_HIP(hipMemcpyAsync(x->sigvrmax, sv.data(), sizeof(double) * sv.size(),
                                hipMemcpyHostToDevice, st));
    }
    DSMC_HIP(hipMemcpyAsync(x->step_dev(), &zero, sizeof zero, hipMemcpyHostToDevice, st));
    DSMC_HIP(hipMemsetAsync(x->coll_accum, 0, sizeof(unsigned long long) * x->T, st));
    DSMC_HIP(hipMemcpyAsync(x->n_dev, nn, sizeof nn, hipMemcpyHostToDevice, st));
    if (x->fp32)
        DSMC_HIP(hipMemcpyAsync(x->trap_start(), x->atom_begin.data(), sizeof(int) * (x->T + 1),
                                hipMemcpyHostToDevice, st));
    DSMC_HIP(hipStreamSynchronize(st));
    x->step = 0;
    x->coll_steps = 0;
    x->cur = 0;
    x->have_sort = false;
    x->srt_valid = false;
    x->dead_in_state = false;
    x->tbl_valid = true;    // all atoms alive, traps at their initial ranges
}

static InitParams init_params(const dsmc_ctx *x)
{
    InitParams ip;
    std::memset(&ip, 0, sizeof ip);
    ip.seed = x->cfg.seed;
    ip.mass = x->cfg.mass;
    ip.gs = x->cfg.gs;
    ip.spin = x->cfg.spin;
    return ip;
}

// (plane, component) of each axis' value and remainder in the FP32 state
// records: position x, y, z = (0,0)+(0,3), (0,1)+(2,0), (0,2)+(2,1); velocity
// x, y, z = (1,0)+(1,3), (1,1)+(2,2), (1,2)+(2,3).
static const int kRecPosHi[3][2] = { { 0, 0 }, { 0, 1 }, { 0, 2 } };
static const int kRecPosLo[3][2] = { { 0, 3 }, { 2, 0 }, { 2, 1 } };
static const int kRecVelHi[3][2] = { { 1, 0 }, { 1, 1 }, { 1, 2 } };
static const int kRecVelLo[3][2] = { { 1, 3 }, { 2, 2 }, { 2, 3 } };

// ---------------------------------------------------------------------------
// Host-side field conversion (device layout -> AoS host)
// ---------------------------------------------------------------------------
template <typename T_>
static std::vector<T_> d2h(const T_ *d, size_t n)
{
    std::vector<T_> h(n);
    if (n) DSMC_HIP(hipMemcpy(h.data(), d, sizeof(T_) * n, hipMemcpyDeviceToHost));
    return h;
}

static void interleave(double *dst, const std::vector<double> *cols, int ncols, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < ncols; ++c) dst[i * ncols + c] = cols[c][i];
}

static size_t field_size(dsmc_ctx *x, int field, int n)
{
    const size_t N = (size_t)n;
    switch (field) {
    case DSMC_FIELD_POS: case DSMC_FIELD_VEL: return N * 3 * sizeof(double);
    case DSMC_FIELD_ACC: return x->fp32 ? 0 : N * 3 * sizeof(double);
    case DSMC_FIELD_ATOM_ID: case DSMC_FIELD_CELL_ID: case DSMC_FIELD_PERM: return N * sizeof(int);
    case DSMC_FIELD_SPIN:
        if (!x->cfg.spin) return 0;
        return N * (x->fp32 ? 3 : 4) * sizeof(double);
    case DSMC_FIELD_EVAP_POS: case DSMC_FIELD_EVAP_VEL: return (size_t)x->n_cap * 3 * sizeof(double);
    case DSMC_FIELD_EVAP_TAG: return x->fp32 ? 0 : N * sizeof(int);
    case DSMC_FIELD_CELL_START_END:
        return x->fp32 ? (size_t)x->n_keys * 2 * sizeof(int) : (size_t)(x->ncells + 1) * 2 * sizeof(int);
    case DSMC_FIELD_CELL_COUNT: case DSMC_FIELD_CELL_PREFIX: case DSMC_FIELD_COLLISION_COUNT:
        return x->fp32 ? 0 : (size_t)(x->ncells + 1) * sizeof(int);
    case DSMC_FIELD_SIGVRMAX: return (size_t)x->n_keys * sizeof(double);
    case DSMC_FIELD_TRAP_START: return (size_t)(x->T + 1) * sizeof(int);
    case DSMC_FIELD_POS_COMP: case DSMC_FIELD_VEL_COMP: return x->fp32 ? N * 3 * sizeof(double) : 0;
    case DSMC_FIELD_OBS: return (size_t)x->obs_cap * x->T * sizeof(dsmc_observables);
    case DSMC_FIELD_SPIN_UP: return N * sizeof(int);
    case DSMC_FIELD_LOCAL_POPS: return x->pops ? (size_t)x->n_cap * 2 * sizeof(double) : 0;
    default: return 0;
    }
}

// Trap ranges of the alive atoms in slot order (traps are contiguous there),
// from their composite keys.
static void trap_ranges_of(const dsmc_ctx *x, const std::vector<uint32_t> &keys, int *out)
{
    std::vector<int> cnt(x->T, 0);
    for (uint32_t k : keys) {
        const int t = (int)(std::upper_bound(x->key_base.begin(), x->key_base.begin() + x->T, (int)k) -
                            x->key_base.begin()) - 1;
        cnt[t]++;
    }
    out[0] = 0;
    for (int t = 0; t < x->T; ++t) out[t + 1] = out[t] + cnt[t];
}

static void copy_field(dsmc_ctx *x, int field, void *host, size_t bytes)
{
    std::vector<uint32_t> live;
    const int n_slots = x->fp32 ? fused_live(x, live) : host_slots(x);
    const int n = x->fp32 ? (int)live.size() : n_slots;
    const size_t need = field_size(x, field, n);
    // a field of this form has a size at one atom; at no alive atoms it is empty
    if (field_size(x, field, 1) == 0) fail(DSMC_ENOTSUP, "field not available in this form");
    if (need == 0) return;
    if (bytes < need) fail(DSMC_EINVAL, "host buffer too small");
    const size_t N = (size_t)n;
    double *hd = (double *)host;
    std::vector<double> cols[4];
    if (!x->fp32) {
        const PlainPtrs &s = x->ps[x->cur];
        switch (field) {
        case DSMC_FIELD_POS: case DSMC_FIELD_VEL: case DSMC_FIELD_ACC: {
            const int o = field == DSMC_FIELD_POS ? 0 : field == DSMC_FIELD_VEL ? 3 : 6;
            for (int c = 0; c < 3; ++c) cols[c] = d2h(s.d[o + c], N);
            interleave(hd, cols, 3, N);
            return;
        }
        case DSMC_FIELD_ATOM_ID: DSMC_HIP(hipMemcpy(host, s.id, N * 4, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_SPIN:
            for (int c = 0; c < 4; ++c) cols[c] = d2h(s.psi[c], N);
            interleave(hd, cols, 4, N);
            return;
        case DSMC_FIELD_EVAP_POS:
            DSMC_HIP(hipMemcpy(host, x->evap_pos_d, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_EVAP_VEL:
            DSMC_HIP(hipMemcpy(host, x->evap_vel_d, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_EVAP_TAG: DSMC_HIP(hipMemcpy(host, x->tag, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_CELL_START_END:
            DSMC_HIP(hipMemcpy(host, x->start_end, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_CELL_COUNT: DSMC_HIP(hipMemcpy(host, x->count, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_CELL_PREFIX: DSMC_HIP(hipMemcpy(host, x->prefix, need, hipMemcpyDeviceToHost)); return;
        case DSMC_FIELD_COLLISION_COUNT:
            DSMC_HIP(hipMemcpy(host, x->coll_count, need, hipMemcpyDeviceToHost)); return;
        default: break;
        }
    } else {
        const FusedPtrs &s = x->fs[x->cur];
        switch (field) {
        case DSMC_FIELD_POS: case DSMC_FIELD_VEL: case DSMC_FIELD_POS_COMP: case DSMC_FIELD_VEL_COMP: {
            const bool pos = field == DSMC_FIELD_POS || field == DSMC_FIELD_POS_COMP;
            const bool comp = field == DSMC_FIELD_POS_COMP || field == DSMC_FIELD_VEL_COMP;
            const std::vector<f4> rec = d2h(s.r, rec_elems((size_t)n_slots));
            for (int c = 0; c < 3; ++c) {
                const int *h = pos ? kRecPosHi[c] : kRecVelHi[c], *l = pos ? kRecPosLo[c] : kRecVelLo[c];
                cols[c].resize(N);
                for (size_t i = 0; i < N; ++i) {
                    const f4 *q = rec.data() + rec_index(live[i]);
                    const float a = q[h[0] * kRecTile][h[1]];
                    const float b = pos ? rem_value(q[l[0] * kRecTile][l[1]]) : q[l[0] * kRecTile][l[1]];
                    cols[c][i] = comp ? (double)b : (double)a + (double)b;
                }
            }
            interleave(hd, cols, 3, N);
            return;
        }
        case DSMC_FIELD_ATOM_ID: {   // the labels travel in the position remainders
            const std::vector<f4> rec = d2h(s.r, rec_elems((size_t)n_slots));
            int *hi_ = (int *)host;
            for (size_t i = 0; i < N; ++i) {
                const f4 *q = rec.data() + rec_index(live[i]);
                hi_[i] = rec_label(q[0][3], q[2 * kRecTile][0], q[2 * kRecTile][1]);
            }
            return;
        }
        case DSMC_FIELD_SPIN:
            for (int c = 0; c < 3; ++c) {
                const std::vector<float> a = d2h(s.s[c], (size_t)n_slots);
                cols[c].resize(N);
                for (size_t i = 0; i < N; ++i) cols[c][i] = a[live[i]];
            }
            interleave(hd, cols, 3, N);
            return;
        case DSMC_FIELD_EVAP_POS: case DSMC_FIELD_EVAP_VEL: {
            std::vector<float> a = d2h(field == DSMC_FIELD_EVAP_POS ? x->evap_pos_f : x->evap_vel_f,
                                       (size_t)x->n_cap * 3);
            for (size_t i = 0; i < a.size(); ++i) hd[i] = a[i];
            return;
        }
        case DSMC_FIELD_CELL_START_END: {
            if (!x->have_sort) fail(DSMC_ESTATE, "the cell table exists between sort and move");
            DSMC_HIP(hipMemsetAsync(x->cell_start, 0xFF, sizeof(uint32_t) * x->n_keys, x->stream));
            DSMC_HIP(hipMemsetAsync(x->cell_end, 0, sizeof(uint32_t) * x->n_keys, x->stream));
            launch_cell_range_fused(x->sorted_key, x->n_dev + (x->cur ^ 1), x->dead_key, x->cell_start,
                                    x->cell_end, x->n_cap, x->stream);
            DSMC_HIP(hipStreamSynchronize(x->stream));
            std::vector<uint32_t> a = d2h(x->cell_start, (size_t)x->n_keys);
            std::vector<uint32_t> b = d2h(x->cell_end, (size_t)x->n_keys);
            int *hi_ = (int *)host;
            for (int k = 0; k < x->n_keys; ++k) { hi_[2 * k] = (int)a[k]; hi_[2 * k + 1] = (int)b[k] - 1; }
            return;
        }
        case DSMC_FIELD_TRAP_START: {
            std::vector<uint32_t> lk;
            (void)fused_live(x, live, &lk);
            trap_ranges_of(x, lk, (int *)host);
            return;
        }
        case DSMC_FIELD_CELL_ID:
            if (!x->have_sort) {   // keys of the alive slots (unsorted order)
                const std::vector<uint32_t> a = d2h(x->key_a, (size_t)n_slots);
                uint32_t *hk = (uint32_t *)host;
                for (size_t i = 0; i < N; ++i) hk[i] = a[live[i]];
                return;
            }
            break;
        default: break;
        }
    }
    switch (field) {
    case DSMC_FIELD_CELL_ID: {
        // FP64 form: the keys of the last index or sort; FP32 form after a
        // sort: the sorted keys (the alive atoms; removed atoms sort last)
        const uint32_t *src = (x->fp32 && !x->have_sort) ? x->key_a : x->sorted_key;
        if (!src) src = x->key_a;
        DSMC_HIP(hipMemcpy(host, src, N * 4, hipMemcpyDeviceToHost));
        return;
    }
    case DSMC_FIELD_PERM:
        if (!x->perm) fail(DSMC_ESTATE, "no sort has run yet");
        DSMC_HIP(hipMemcpy(host, x->perm, N * 4, hipMemcpyDeviceToHost));
        return;
    case DSMC_FIELD_SIGVRMAX:
        if (x->fp32) {
            const std::vector<float> f = d2h(x->sigvrmax_f, (size_t)x->n_keys);
            double *h = (double *)host;
            for (int k = 0; k < x->n_keys; ++k) h[k] = (double)f[k];
        } else {
            DSMC_HIP(hipMemcpy(host, x->sigvrmax, need, hipMemcpyDeviceToHost));
        }
        return;
    case DSMC_FIELD_TRAP_START: {
        int *h = (int *)host;
        h[0] = 0;
        h[1] = n;
        return;
    }
    case DSMC_FIELD_OBS:
        DSMC_HIP(hipMemcpy(host, x->obs_log, need, hipMemcpyDeviceToHost));
        return;
    case DSMC_FIELD_SPIN_UP: {
        int *h = (int *)host;
        if (x->fp32) {   // the FP32 form removes atoms instead of flipping them
            for (size_t i = 0; i < N; ++i) h[i] = 1;
            return;
        }
        const std::vector<uint8_t> u = d2h(x->ps[x->cur].up, N);
        for (size_t i = 0; i < N; ++i) h[i] = u[i] ? 1 : 0;
        return;
    }
    case DSMC_FIELD_LOCAL_POPS:
        DSMC_HIP(hipMemcpy(host, x->pops, need, hipMemcpyDeviceToHost));
        return;
    default: break;
    }
    fail(DSMC_EINVAL, "unknown field");
}

#define PRIM_BEGIN try {
#define PRIM_END                                                              \
    return DSMC_OK;                                                           \
    }                                                                         \
    catch (const Error &e) { g_create_error = e.what(); return (dsmc_status)e.status; } \
    catch (const std::exception &e) { g_create_error = e.what(); return DSMC_EHIP; }

// Temporary device buffer owned by one primitive call.
struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { DSMC_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <typename T_> T_ *as() const { return (T_ *)p; }
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free sort (dsmc_sort_by_key_noalloc): a
// file-scope device array, one copy per device, resident until the process
// ends.  Layout: the pair count (one word in a 256-byte header), the radix
// histograms and digit totals for kNoAllocMaxPairs pairs, then the alternate
// key and value buffers of the passes.  Host code never uses the array's
// name as an address: the address on each device comes from
// hipGetSymbolAddress with that device current (noalloc_scratch()).
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxPairs = 1 << 20;
constexpr size_t kNoAllocHdr = 256;
constexpr size_t kNoAllocSortBytes =
    ((size_t)kRadix * (kNoAllocMaxPairs / kSortTile) * 4 + (size_t)kRadix * 4 + 256 + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes = kNoAllocHdr + kNoAllocSortBytes + 2 * (size_t)kNoAllocMaxPairs * 4;
static_assert(kNoAllocMaxPairs % kSortTile == 0, "whole sort tiles");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (first call there).
static char *noalloc_scratch()
{
    int dev = 0;
    DSMC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) fail(DSMC_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        DSMC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(DSMC_EHIP, "hipGetSymbolAddress returned no address for the sort scratch");
        g_noalloc_ptr[dev] = p;
    }
    return (char *)g_noalloc_ptr[dev];
}

__global__ void k_store_i32(int *p, int v) { *p = v; }

__global__ __launch_bounds__(256) void k_copy_pairs(uint32_t *__restrict__ dk, uint32_t *__restrict__ dv,
                                                    const uint32_t *__restrict__ sk,
                                                    const uint32_t *__restrict__ sv, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        dk[i] = sk[i];
        dv[i] = sv[i];
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int dsmc_abi_version(void) { return DSMC_ABI_VERSION; }

const char *dsmc_last_error(const dsmc_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

dsmc_status dsmc_create(const dsmc_config *cfg, const dsmc_trap *traps, int n_traps, dsmc_ctx **out)
{
    if (!out) return DSMC_EINVAL;
    *out = nullptr;
    dsmc_ctx *x = nullptr;
    try {
        validate(cfg, traps, n_traps);
        x = new dsmc_ctx();
        build_context(x, cfg, traps, n_traps);
        reset_cells(x);
        *out = x;
        g_create_error.clear();
        return DSMC_OK;
    } catch (const Error &e) {
        g_create_error = e.what();
        delete x;
        return (dsmc_status)e.status;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        delete x;
        return DSMC_ENOMEM;
    }
}

dsmc_status dsmc_destroy(dsmc_ctx *ctx)
{
    delete ctx;
    return DSMC_OK;
}

dsmc_status dsmc_init_distribution(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    reset_cells(ctx);
    const InitParams ip = init_params(ctx);
    if (!ctx->fp32) {
        launch_init_plain(ctx->tdev[0], ip, ctx->ps[0], ctx->n_cap, ctx->stream, &ctx->prof);
        plain_index(ctx);
    } else {
        launch_init_fused(ctx->d_traps, ctx->trap_start(), ctx->T, ip, ctx->fs[0], ctx->n_cap,
                          ctx->stream, &ctx->prof);
        fused_index(ctx);
    }
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    DSMC_API_END(ctx)
}

dsmc_status dsmc_set_state(dsmc_ctx *ctx, const double *pos, const double *vel, int n)
{
    DSMC_API_BEGIN(ctx)
    if (!pos || !vel || n != ctx->n_cap) fail(DSMC_EINVAL, "set_state needs n_capacity atoms");
    reset_cells(ctx);
    const size_t N = (size_t)n;
    std::vector<int> ids(N);
    for (size_t i = 0; i < N; ++i) ids[i] = (int)i;
    if (!ctx->fp32) {
        const PlainPtrs &s = ctx->ps[0];
        std::vector<double> cols(6 * N);
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < N; ++i) {
                cols[c * N + i] = pos[3 * i + c];
                cols[(3 + c) * N + i] = vel[3 * i + c];
            }
        for (int c = 0; c < 6; ++c)
            DSMC_HIP(hipMemcpyAsync(s.d[c], cols.data() + c * N, N * 8, hipMemcpyHostToDevice,
                                    ctx->stream));
        DSMC_HIP(hipMemcpyAsync(s.id, ids.data(), N * 4, hipMemcpyHostToDevice, ctx->stream));
        DSMC_HIP(hipMemsetAsync(s.up, 1, N, ctx->stream));
        launch_accel_plain(ctx->tdev[0], s, n, ctx->stream, &ctx->prof);
        plain_index(ctx);
    } else {
        const FusedPtrs &s = ctx->fs[0];
        std::vector<f4> rec(rec_elems(N), f4{ 0.0f, 0.0f, 0.0f, 0.0f });
        for (int c = 0; c < 3; ++c) {
            for (int which = 0; which < 2; ++which) {
                const double *src = which ? vel : pos;
                const int *h = which ? kRecVelHi[c] : kRecPosHi[c];
                const int *l = which ? kRecVelLo[c] : kRecPosLo[c];
                for (size_t i = 0; i < N; ++i) {
                    f4 *q = rec.data() + rec_index((uint32_t)i);
                    const double v = src[3 * i + c];
                    const float hv = (float)v, lv = (float)(v - (double)hv);
                    q[h[0] * kRecTile][h[1]] = hv;
                    // slot i holds atom i: its label in the position remainders
                    q[l[0] * kRecTile][l[1]] = which ? lv : rem_tag(lv, label_part((int)i, c));
                }
            }
        }
        DSMC_HIP(hipMemcpyAsync(s.r, rec.data(), rec.size() * sizeof(f4), hipMemcpyHostToDevice,
                                ctx->stream));
        fused_index(ctx);
    }
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    DSMC_API_END(ctx)
}

dsmc_status dsmc_set_spin(dsmc_ctx *ctx, const double *spin, int n, int components)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->cfg.spin) fail(DSMC_ENOTSUP, "spin evolution is off in this context");
    if (!spin || n != host_alive(ctx)) fail(DSMC_EINVAL, "set_spin needs one row per alive atom");
    const size_t N = (size_t)n;
    if (!ctx->fp32) {
        if (components != 4) fail(DSMC_EINVAL, "the FP64 form stores spinors (4 components)");
        std::vector<double> col(4 * N);
        for (int c = 0; c < 4; ++c) {
            for (size_t i = 0; i < N; ++i) col[c * N + i] = spin[4 * i + c];
            DSMC_HIP(hipMemcpyAsync(ctx->ps[ctx->cur].psi[c], col.data() + c * N, N * 8,
                                    hipMemcpyHostToDevice, ctx->stream));
        }
        DSMC_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        if (components != 3) fail(DSMC_EINVAL, "the FP32 form stores Bloch vectors (3 components)");
        // rows go to the alive slots; slots of removed atoms keep their values
        std::vector<uint32_t> live;
        const size_t S = (size_t)fused_live(ctx, live);
        for (int c = 0; c < 3; ++c) {
            std::vector<float> col = d2h(ctx->fs[ctx->cur].s[c], S);
            for (size_t i = 0; i < N; ++i) col[live[i]] = (float)spin[3 * i + c];
            DSMC_HIP(hipMemcpy(ctx->fs[ctx->cur].s[c], col.data(), S * 4, hipMemcpyHostToDevice));
        }
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_index_atoms(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) {
        if (ctx->have_sort) fail(DSMC_ESTATE, "a sort is pending: call dsmc_move_atoms first");
        fused_index(ctx);
    } else {
        plain_index(ctx);
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_sort_arrays(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) {
        if (ctx->have_sort) fail(DSMC_ESTATE, "a sort is pending: call dsmc_move_atoms first");
        fused_sort(ctx);
    } else {
        plain_sort(ctx);
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_build_cell_table(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->sorted_key) fail(DSMC_ESTATE, "no sort has run yet");
    if (ctx->fp32) {
        if (!ctx->have_sort) fail(DSMC_ESTATE, "build the table right after dsmc_sort_arrays");
        fused_table(ctx);
    } else {
        plain_table(ctx);
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_collide(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) {
        if (!ctx->have_sort) fail(DSMC_ESTATE, "collide needs a sorted state and its table");
        fused_collide(ctx);
        ctx->coll_steps++;
    } else {
        plain_collide(ctx);
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_evolve_spin(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->cfg.spin) fail(DSMC_ENOTSUP, "spin evolution is off in this context");
    if (ctx->fp32) fail(DSMC_ENOTSUP, "the FP32 form evolves spins inside dsmc_move_atoms");
    plain_spin(ctx);
    launch_normalise_plain(ctx->ps[ctx->cur], ctx->n_dev, ctx->n_cap, ctx->stream, &ctx->prof);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_move_atoms(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) {
        fused_move(ctx);
    } else {
        launch_move_plain(ctx->tdev[0], ctx->cfg.dt, ctx->ps[ctx->cur], ctx->n_dev, ctx->n_cap,
                          ctx->stream, &ctx->prof);
        plain_observe(ctx, ctx->obs_log + (ctx->step % ctx->obs_cap), ctx->step);
    }
    ctx->step++;
    DSMC_API_END(ctx)
}

dsmc_status dsmc_calculate_temp(dsmc_ctx *ctx, double *T)
{
    DSMC_API_BEGIN(ctx)
    if (!T) fail(DSMC_EINVAL, "null output");
    std::vector<dsmc_observables> o;
    observe_now(ctx, o);
    const double v2 = o[0].v2[0] + o[0].v2[1] + o[0].v2[2];
    *T = o[0].n_alive > 0 ? ctx->cfg.mass / 3. / kKB * (v2 / (double)o[0].n_alive) : 0.0;
    DSMC_API_END(ctx)
}

dsmc_status dsmc_local_populations(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->pops) fail(DSMC_ENOTSUP, "the spin-flip model is off in this context (cfg.mcwf)");
    plain_local_pops(ctx);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_flip_atoms(dsmc_ctx *ctx, long long *n_flipped)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->pops) fail(DSMC_ENOTSUP, "the spin-flip model is off in this context (cfg.mcwf)");
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, ctx->stream, ctx->flips);
    DSMC_HIP(hipGetLastError());
    plain_flip(ctx);
    if (n_flipped) {
        unsigned long long f = 0;
        DSMC_HIP(hipMemcpyAsync(&f, ctx->flips, sizeof f, hipMemcpyDeviceToHost, ctx->stream));
        DSMC_HIP(hipStreamSynchronize(ctx->stream));
        *n_flipped = (long long)f;
    }
    DSMC_API_END(ctx)
}

dsmc_status dsmc_exponential_decay(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->pops) fail(DSMC_ENOTSUP, "the spin-flip model is off in this context (cfg.mcwf)");
    plain_decay(ctx);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_evaporation_tag(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) fail(DSMC_ENOTSUP, "the FP32 form tags inside dsmc_move_atoms");
    plain_evap_tag(ctx);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_compact(dsmc_ctx *ctx, int *n_alive)
{
    DSMC_API_BEGIN(ctx)
    if (ctx->fp32) fail(DSMC_ENOTSUP, "the FP32 form compacts inside the sort");
    plain_compact(ctx);
    if (n_alive) *n_alive = host_alive(ctx);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_step(dsmc_ctx *ctx, int n_steps)
{
    DSMC_API_BEGIN(ctx)
    if (n_steps < 0) fail(DSMC_EINVAL, "n_steps < 0");
    do_steps(ctx, n_steps);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_sync(dsmc_ctx *ctx)
{
    DSMC_API_BEGIN(ctx)
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    DSMC_API_END(ctx)
}

dsmc_status dsmc_get_info(dsmc_ctx *ctx, dsmc_info *info)
{
    DSMC_API_BEGIN(ctx)
    if (!info) fail(DSMC_EINVAL, "null output");
    info->precision = ctx->cfg.precision;
    info->n_traps = ctx->T;
    info->n_capacity = ctx->n_cap;
    info->n_alive = host_alive(ctx);
    info->n_keys = ctx->n_keys;
    info->key_bits = ctx->key_bits;
    info->radix_passes = ctx->passes;
    info->block_size = ctx->block;
    info->step = ctx->step;
    info->collision_steps = ctx->coll_steps;
    info->obs_capacity = ctx->obs_cap;
    DSMC_API_END(ctx)
}

dsmc_status dsmc_observe(dsmc_ctx *ctx, dsmc_observables *per_trap, int n_traps)
{
    DSMC_API_BEGIN(ctx)
    if (!per_trap || n_traps < ctx->T) fail(DSMC_EINVAL, "need one record per trap");
    std::vector<dsmc_observables> o;
    observe_now(ctx, o);
    std::memcpy(per_trap, o.data(), sizeof(dsmc_observables) * ctx->T);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_copy_out(dsmc_ctx *ctx, int field, void *host, size_t bytes)
{
    DSMC_API_BEGIN(ctx)
    if (!host) fail(DSMC_EINVAL, "null output");
    copy_field(ctx, field, host, bytes);
    DSMC_API_END(ctx)
}

dsmc_status dsmc_field_bytes(dsmc_ctx *ctx, int field, size_t *bytes)
{
    DSMC_API_BEGIN(ctx)
    if (!bytes) fail(DSMC_EINVAL, "null output");
    *bytes = field_size(ctx, field, host_alive(ctx));
    DSMC_API_END(ctx)
}

dsmc_status dsmc_time_steps(dsmc_ctx *ctx, int n_steps, double *ms)
{
    DSMC_API_BEGIN(ctx)
    if (!ms || n_steps < 0) fail(DSMC_EINVAL, "bad arguments");
    hipEvent_t b, e;
    DSMC_HIP(hipEventCreate(&b));
    DSMC_HIP(hipEventCreate(&e));
    try {
        DSMC_HIP(hipStreamSynchronize(ctx->stream));
        DSMC_HIP(hipEventRecord(b, ctx->stream));
        do_steps(ctx, n_steps);
        DSMC_HIP(hipEventRecord(e, ctx->stream));
        DSMC_HIP(hipEventSynchronize(e));
        float f = 0.f;
        DSMC_HIP(hipEventElapsedTime(&f, b, e));
        *ms = f;
    } catch (...) {
        (void)hipEventDestroy(b);
        (void)hipEventDestroy(e);
        throw;
    }
    DSMC_HIP(hipEventDestroy(b));
    DSMC_HIP(hipEventDestroy(e));
    DSMC_API_END(ctx)
}

// ---- several contexts stepped together ------------------------------------
// Independent groups of traps in separate contexts (each with its own stream
// and graph) are stepped in turns of `chunk` steps, so that every stream has
// work queued and the GPU runs one group's latency-bound kernels (re-sort,
// collide, finalize) beside another's bandwidth-bound move.
static void check_group(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk)
{
    if (n_steps < 0 || chunk < 1) fail(DSMC_EINVAL, "n_steps < 0 or chunk < 1");
    for (int i = 0; i < n_ctx; ++i) {
        if (!ctxs[i]) fail(DSMC_EINVAL, "null context in the group");
        if (ctxs[i]->device != ctxs[0]->device) fail(DSMC_EINVAL, "the group's contexts are on different devices");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) fail(DSMC_EINVAL, "a context appears twice in the group");
    }
}

static void group_steps(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk)
{
    for (int done = 0; done < n_steps; done += chunk) {
        const int k = std::min(chunk, n_steps - done);
        for (int i = 0; i < n_ctx; ++i) do_steps(ctxs[i], k);
    }
}

dsmc_status dsmc_step_group(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk)
{
    if (!ctxs || n_ctx < 1) return DSMC_EINVAL;
    DSMC_API_BEGIN(ctxs[0])
    check_group(ctxs, n_ctx, n_steps, chunk);
    group_steps(ctxs, n_ctx, n_steps, chunk);
    DSMC_API_END(ctxs[0])
}

dsmc_status dsmc_time_steps_group(dsmc_ctx *const *ctxs, int n_ctx, int n_steps, int chunk, double *ms)
{
    if (!ctxs || n_ctx < 1) return DSMC_EINVAL;
    DSMC_API_BEGIN(ctxs[0])
    if (!ms) fail(DSMC_EINVAL, "null output");
    check_group(ctxs, n_ctx, n_steps, chunk);
    std::vector<hipEvent_t> ev((size_t)n_ctx + 1, nullptr);
    auto destroy = [&]() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    try {
        for (auto &e : ev) DSMC_HIP(hipEventCreate(&e));
        for (int i = 0; i < n_ctx; ++i) DSMC_HIP(hipStreamSynchronize(ctxs[i]->stream));
        // one start for all: every other stream waits for the first's event
        DSMC_HIP(hipEventRecord(ev[0], ctxs[0]->stream));
        for (int i = 1; i < n_ctx; ++i) DSMC_HIP(hipStreamWaitEvent(ctxs[i]->stream, ev[0], 0));
        group_steps(ctxs, n_ctx, n_steps, chunk);
        for (int i = 0; i < n_ctx; ++i) DSMC_HIP(hipEventRecord(ev[i + 1], ctxs[i]->stream));
        float worst = 0.f;
        for (int i = 0; i < n_ctx; ++i) {
            DSMC_HIP(hipEventSynchronize(ev[i + 1]));
            float f = 0.f;
            DSMC_HIP(hipEventElapsedTime(&f, ev[0], ev[i + 1]));
            worst = std::max(worst, f);
        }
        *ms = worst;
    } catch (...) {
        destroy();
        throw;
    }
    destroy();
    DSMC_API_END(ctxs[0])
}

dsmc_status dsmc_profile_enable(dsmc_ctx *ctx, int enable)
{
    DSMC_API_BEGIN(ctx)
    ctx->prof.collect();
    if (enable) ctx->prof.reset();
    ctx->prof.enabled = enable != 0;
    DSMC_API_END(ctx)
}

dsmc_status dsmc_profile_read(dsmc_ctx *ctx, dsmc_kernel_stat *stats, int max_stats, int *n_stats)
{
    DSMC_API_BEGIN(ctx)
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof.collect();
    int k = 0;
    for (const auto &kv : ctx->prof.stats()) {
        if (stats && k < max_stats) {
            dsmc_kernel_stat &s = stats[k];
            std::memset(&s, 0, sizeof s);
            std::snprintf(s.name, sizeof s.name, "%s", kv.first.c_str());
            s.launches = kv.second.launches;
            s.total_ms = kv.second.total_ms;
            s.bytes_per_launch = kv.second.bytes;
        }
        ++k;
    }
    if (n_stats) *n_stats = k;
    DSMC_API_END(ctx)
}

dsmc_status dsmc_autotune(dsmc_ctx *ctx, int *block_size_out)
{
    DSMC_API_BEGIN(ctx)
    if (!ctx->fp32) {
        if (block_size_out) *block_size_out = 256;
        return DSMC_OK;
    }
    if (ctx->have_sort) fail(DSMC_ESTATE, "a sort is pending: call dsmc_move_atoms first");
    // Time the fused move on the current state without committing it: the move
    // writes the other state buffer and a scratch key buffer, evaporation is
    // off, and the device step counter is restored afterwards.  When the state
    // is in the order of a sort (as in the step loop), the trial moves also
    // compare with that sort's keys and leave the re-sort's masks, counts and
    // samples, as the step's move does; the next sort is then a full one.
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    DSMC_HIP(hipMemcpyAsync(ctx->step_save, ctx->step_dev(), sizeof(long long),
                            hipMemcpyDeviceToDevice, ctx->stream));
    MoveParams mp = move_params(ctx);
    mp.spin_loss = 0;
    mp.energy_cut = 0;
    if (ctx->srt_valid) {
        mp.count_moved = 1;
        mp.key_prev = ctx->sorted_key;
    }
    // (threads per block, atoms per thread): shapes of at least 512 atoms per
    // block.  Smaller blocks multiply the per-block work that trial moves do
    // not weigh (each move block's trap table and sums, the finalize's
    // records, the re-sort's counts): on MI355X a 128 x 2 move made the
    // batched step 26 % slower, and a 64 x 2 move that won the trials made
    // the spin step 2 % slower.
    const int cands[4][2] = { { 512, 1 }, { 1024, 1 }, { 256, 2 }, { 512, 2 } };
    double best = 1e30, pref = 1e30;
    int best_bs = ctx->block, best_apt = ctx->apt;
    // 256 threads x 2 atoms measured best in the step on MI355X (batched,
    // single-trap and spin workloads); the trial moves cannot reproduce the
    // step's cache state, so another candidate must beat it by more than 5 %
    // to be taken
    const int pref_bs = 256, pref_apt = 2;
    hipEvent_t b, e;
    DSMC_HIP(hipEventCreate(&b));
    DSMC_HIP(hipEventCreate(&e));
    const bool was = ctx->prof.enabled;
    ctx->prof.enabled = false;
    // on sorted state the trials time the fused step kernel (its collisions
    // computed but not stored), as the step loop runs it
    StepCollide sc;
    sc.cp = collide_params(ctx);
    sc.sigvrmax = ctx->sigvrmax_f;
    sc.coll_accum = ctx->coll_accum;
    sc.dry = 1;
    const StepCollide *coll = (ctx->srt_valid && !ctx->split_step) ? &sc : nullptr;
    for (const auto &cand : cands) {
        const int bs = cand[0], apt = cand[1];
        float f = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {  // first repetition warms up
            float g = 0.f;
            DSMC_HIP(hipEventRecord(b, ctx->stream));
            for (int k = 0; k < 5; ++k)
                launch_move_fused(bs, apt, ctx->d_traps, mp, ctx->fs[ctx->cur], ctx->fs[ctx->cur ^ 1],
                                  nullptr, ctx->trap_start(), ctx->n_dev + ctx->cur, ctx->res.alt_k,
                                  ctx->evap_pos_f, ctx->evap_vel_f, ctx->rec, ctx->step_dev(),
                                  ctx->n_cap, ctx->stream, nullptr, coll);
            DSMC_HIP(hipEventRecord(e, ctx->stream));
            DSMC_HIP(hipEventSynchronize(e));
            DSMC_HIP(hipEventElapsedTime(&g, b, e));
            if (rep > 0) f = std::min(f, g);
        }
        if (f < best) { best = f; best_bs = bs; best_apt = apt; }
        if (bs == pref_bs && apt == pref_apt) pref = f;
    }
    if (pref <= 1.05 * best) { best_bs = pref_bs; best_apt = pref_apt; }
    ctx->prof.enabled = was;
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(e);
    DSMC_HIP(hipMemcpyAsync(ctx->step_dev(), ctx->step_save, sizeof(long long),
                            hipMemcpyDeviceToDevice, ctx->stream));
    DSMC_HIP(hipMemsetAsync(ctx->acc, 0, sizeof(double) * kAccK * ctx->T, ctx->stream));
    DSMC_HIP(hipStreamSynchronize(ctx->stream));
    if (mp.count_moved) ctx->srt_valid = false;
    if (ctx->block != best_bs || ctx->apt != best_apt) {
        drop_graph(ctx);
        // the per-block counts of the last move are for the old block size
        ctx->srt_valid = false;
    }
    ctx->block = best_bs;
    ctx->apt = best_apt;
    if (block_size_out) *block_size_out = best_bs;
    DSMC_API_END(ctx)
}

// ---- device primitives on caller memory -----------------------------------

dsmc_status dsmc_sort_by_key(unsigned int *d_keys, unsigned int *d_values, int n, int key_bits,
                             void *stream)
{
    PRIM_BEGIN
    if (n < 0 || key_bits < 1 || key_bits > 32 || (n > 0 && (!d_keys || !d_values)))
        fail(DSMC_EINVAL, "bad sort arguments");
    if (n == 0) return DSMC_OK;
    hipStream_t st = (hipStream_t)stream;
    const int passes = key_bits <= kRadixBits ? 1 : div_up(key_bits, kRadixBits);
    DevBuf tk(sizeof(uint32_t) * n), tv(sizeof(uint32_t) * n), nd(sizeof(int)),
        scratch(sort_scratch_bytes(n));
    SortScratch s;
    sort_scratch_bind(s, n, scratch.p);
    DSMC_HIP(hipMemcpyAsync(nd.p, &n, sizeof(int), hipMemcpyHostToDevice, st));
    if (passes % 2 == 0) {
        radix_sort(d_keys, d_values, tk.as<uint32_t>(), tv.as<uint32_t>(), d_keys, d_values,
                   nd.as<int>(), key_bits, s, st, nullptr);
    } else {
        radix_sort(d_keys, d_values, d_keys, d_values, tk.as<uint32_t>(), tv.as<uint32_t>(),
                   nd.as<int>(), key_bits, s, st, nullptr);
        DSMC_HIP(hipMemcpyAsync(d_keys, tk.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        DSMC_HIP(hipMemcpyAsync(d_values, tv.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    DSMC_HIP(hipStreamSynchronize(st));
    PRIM_END
}

dsmc_status dsmc_sort_by_key_noalloc(unsigned int *d_keys, unsigned int *d_values, int n,
                                     int key_bits, void *stream)
{
    PRIM_BEGIN
    if (n < 0 || key_bits < 1 || key_bits > 32 || (n > 0 && (!d_keys || !d_values)))
        fail(DSMC_EINVAL, "bad sort arguments");
    if (n > kNoAllocMaxPairs)
        fail(DSMC_EINVAL, "dsmc_sort_by_key_noalloc sorts at most 1048576 pairs (use dsmc_sort_by_key)");
    char *base = noalloc_scratch();
    if (n == 0) return DSMC_OK;
    hipStream_t st = (hipStream_t)stream;
    int *nd = (int *)base;
    SortScratch s;
    sort_scratch_bind(s, n, base + kNoAllocHdr);
    uint32_t *tk = (uint32_t *)(base + kNoAllocHdr + kNoAllocSortBytes);
    uint32_t *tv = tk + kNoAllocMaxPairs;
    hipLaunchKernelGGL(k_store_i32, dim3(1), dim3(1), 0, st, nd, n);
    DSMC_HIP(hipGetLastError());
    const int passes = key_bits <= kRadixBits ? 1 : div_up(key_bits, kRadixBits);
    if (passes % 2 == 0) {
        radix_sort(d_keys, d_values, tk, tv, d_keys, d_values, nd, key_bits, s, st, nullptr);
    } else {
        radix_sort(d_keys, d_values, d_keys, d_values, tk, tv, nd, key_bits, s, st, nullptr);
        hipLaunchKernelGGL(k_copy_pairs, dim3(std::min(div_up(n, 256), 2048)), dim3(256), 0, st,
                           d_keys, d_values, tk, tv, n);
        DSMC_HIP(hipGetLastError());
    }
    PRIM_END
}

dsmc_status dsmc_exclusive_scan(const int *d_in, int *d_out, int n, void *stream)
{
    PRIM_BEGIN
    if (n < 0 || (n > 0 && (!d_in || !d_out))) fail(DSMC_EINVAL, "bad scan arguments");
    if (n == 0) return DSMC_OK;
    hipStream_t st = (hipStream_t)stream;
    DevBuf part(sizeof(int) * (scan_partials(n) + 8));
    exclusive_scan(d_in, d_out, n, part.as<int>(), st, nullptr);
    DSMC_HIP(hipStreamSynchronize(st));
    PRIM_END
}

dsmc_status dsmc_remove_if(int *d_values, const int *d_stencil, int n, int *n_out, void *stream)
{
    PRIM_BEGIN
    if (n < 0 || (n > 0 && (!d_values || !d_stencil))) fail(DSMC_EINVAL, "bad remove_if arguments");
    if (n == 0) { if (n_out) *n_out = 0; return DSMC_OK; }
    hipStream_t st = (hipStream_t)stream;
    DevBuf map(sizeof(uint32_t) * n), nd(sizeof(int) * 2), part(sizeof(int) * (scan_partials(n) + 8)),
        tmp(sizeof(int) * n);
    DSMC_HIP(hipMemcpyAsync(nd.p, &n, sizeof(int), hipMemcpyHostToDevice, st));
    compaction_map(d_stencil, nd.as<int>(), n, map.as<uint32_t>(), nd.as<int>() + 1, part.as<int>(),
                   st, nullptr);
    DSMC_HIP(hipMemcpyAsync(tmp.p, d_values, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    gather_u32((uint32_t *)d_values, tmp.as<uint32_t>(), map.as<uint32_t>(), nd.as<int>() + 1, n, st,
               nullptr);
    int kept = 0;
    DSMC_HIP(hipMemcpyAsync(&kept, nd.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    DSMC_HIP(hipStreamSynchronize(st));
    if (n_out) *n_out = kept;
    PRIM_END
}

dsmc_status dsmc_philox(unsigned int *d_out, int n, const unsigned int ctr[4],
                        const unsigned int key[2], void *stream)
{
    PRIM_BEGIN
    if (n < 0 || !ctr || !key || (n > 0 && !d_out)) fail(DSMC_EINVAL, "bad philox arguments");
    philox_fill(d_out, n, ctr, key, (hipStream_t)stream);
    DSMC_HIP(hipStreamSynchronize((hipStream_t)stream));
    PRIM_END
}

dsmc_status dsmc_stream_copy(void *d_dst, const void *d_src, size_t bytes, int repeats,
                             double *ms)
{
    PRIM_BEGIN
    if (!d_dst || !d_src || bytes % 16 || repeats < 1) fail(DSMC_EINVAL, "bad stream_copy arguments");
    hipEvent_t b, e;
    DSMC_HIP(hipEventCreate(&b));
    DSMC_HIP(hipEventCreate(&e));
    stream_copy(d_dst, d_src, bytes, nullptr);  // warm-up
    DSMC_HIP(hipEventRecord(b, nullptr));
    for (int r = 0; r < repeats; ++r) stream_copy(d_dst, d_src, bytes, nullptr);
    DSMC_HIP(hipEventRecord(e, nullptr));
    DSMC_HIP(hipEventSynchronize(e));
    float f = 0.f;
    DSMC_HIP(hipEventElapsedTime(&f, b, e));
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(e);
    if (ms) *ms = f / repeats;
    PRIM_END
}

dsmc_status dsmc_gather_copy(void *d_dst, const void *d_src, const unsigned int *d_map, int n_slots,
                             int repeats, double *ms)
{
    PRIM_BEGIN
    if (!d_dst || !d_src || !d_map || n_slots < 1 || repeats < 1)
        fail(DSMC_EINVAL, "bad gather_copy arguments");
    hipEvent_t b, e;
    DSMC_HIP(hipEventCreate(&b));
    DSMC_HIP(hipEventCreate(&e));
    gather_copy(d_dst, d_src, d_map, n_slots, nullptr);  // warm-up
    DSMC_HIP(hipEventRecord(b, nullptr));
    for (int r = 0; r < repeats; ++r) gather_copy(d_dst, d_src, d_map, n_slots, nullptr);
    DSMC_HIP(hipEventRecord(e, nullptr));
    DSMC_HIP(hipEventSynchronize(e));
    float f = 0.f;
    DSMC_HIP(hipEventElapsedTime(&f, b, e));
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(e);
    if (ms) *ms = f / repeats;
    PRIM_END
}

dsmc_status dsmc_device_alloc(void **d_ptr, size_t bytes)
{
    PRIM_BEGIN
    if (!d_ptr) fail(DSMC_EINVAL, "null output");
    DSMC_HIP(hipMalloc(d_ptr, std::max<size_t>(bytes, 16)));
    PRIM_END
}

dsmc_status dsmc_device_free(void *d_ptr)
{
    PRIM_BEGIN
    if (d_ptr) DSMC_HIP(hipFree(d_ptr));
    PRIM_END
}

dsmc_status dsmc_memcpy_htod(void *d_dst, const void *h_src, size_t bytes)
{
    PRIM_BEGIN
    DSMC_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    PRIM_END
}

dsmc_status dsmc_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes)
{
    PRIM_BEGIN
    DSMC_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    PRIM_END
}

}  // extern "C"
