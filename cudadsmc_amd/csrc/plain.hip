// plain.hip -- FP64 drop-in form of the DSMC step, one kernel per reference
// call, structure-of-arrays state kept physically in cell order.
//
// The reference keeps AoS double3 state fixed in memory and permutes an atomID
// indirection list (main.cu:88-140), so after the first sort every per-atom
// kernel gathers through atomID.  Here the state itself is gathered into sorted
// order after each sort (k_gather_plain), so every per-atom kernel streams
// contiguous SoA arrays and collide reads each cell's atoms contiguously; the
// slot -> atom label map is carried along as `id`.
#include "dsmc_device.hpp"
#include "dsmc_internal.hpp"
#include "plain.hpp"

namespace dsmc {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// indexAtoms -> findAtomIndex (collisions.cu:163-185): key[slot].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_index_plain(TrapDev t, const double *__restrict__ x,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ z,
                                                        uint32_t *__restrict__ key,
                                                        const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < *n_dev) key[i] = (uint32_t)(t.key_base + cell_of(t, x[i], y[i], z[i]));
}

// Gather every state array through `map` (dst[i] = src[map[i]]).
__global__ __launch_bounds__(kBlock) void k_gather_plain(PlainPtrs dst, PlainPtrs src,
                                                         const uint32_t *__restrict__ map,
                                                         const int *__restrict__ n_dev, int spin)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t j = map[i];
#pragma unroll
    for (int c = 0; c < 9; ++c) dst.d[c][i] = src.d[c][j];
    dst.id[i] = src.id[j];
    dst.up[i] = src.up[j];
    if (spin) {
#pragma unroll
        for (int c = 0; c < 4; ++c) dst.psi[c][i] = src.psi[c][j];
    }
}

// ---------------------------------------------------------------------------
// Cell table (main.cu:426-442; collisions.cu:228-271).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cell_start_end(const uint32_t *__restrict__ key,
                                                           const int *__restrict__ n_dev,
                                                           int2 *__restrict__ se)
{
    const int n = *n_dev;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = key[s];
    if (s == 0 || key[s - 1] != c) se[c].x = s;
    if (s == n - 1 || key[s + 1] != c) se[c].y = s;
}

__global__ __launch_bounds__(kBlock) void k_cell_count(const int2 *__restrict__ se,
                                                       int *__restrict__ count, int n)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const int2 v = se[c];
    count[c] = v.x == -1 ? 0 : v.y - v.x + 1;
}

// ---------------------------------------------------------------------------
// collide (collisions.cu:291-445): one thread per cell, trials in sequence.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_collide_plain(
    TrapDev t, CollideParams cp, double *__restrict__ vx, double *__restrict__ vy,
    double *__restrict__ vz, double *__restrict__ sigvrmax, const int *__restrict__ prefix,
    int *__restrict__ coll_count, unsigned long long *__restrict__ coll_total, int ncells,
    uint32_t step)
{
    const double cellVolume = t.cell_vol;
    const double crossSection = cp.sigma;
    const double lpc = cp.lpc;
    unsigned long long mine = 0;
    for (int cell = blockIdx.x * kBlock + threadIdx.x; cell < ncells;
         cell += gridDim.x * kBlock) {
        const int n = prefix[cell + 1] - prefix[cell];
        // every cell's count is that of this call: collisions.cu:308,335
        // reset it only in cells of two or more atoms (a cell that empties
        // keeps a stale count there)
        if (n <= 1) { coll_count[cell] = 0; continue; }
        const double Mc = 0.5 * (n - 1) * n;
        const double X = Mc * t.alpha * lpc * cp.dt * sigvrmax[cell] / cellVolume;
        const double lambda = ceil(X) / Mc;
        // reference pair mode keeps collisions.cu:322's truncation of Mc * lambda
        const int Ncol = cp.pair_mode ? (int)(Mc * lambda) : (int)ceil(X);
        const int base = prefix[cell];
        int count = 0;
        double smax = sigvrmax[cell];
        for (int tr = 0; tr < Ncol; ++tr) {
            int i = 0, j = 1;
            double u_acc;
            if (cp.pair_mode == 0) {
                U32x4 o = draw(cp.seed, t.stream, (uint32_t)cell, step, (uint32_t)tr,
                               kPurposeCollide, 0);
                i = (int)uint_below(o.x, (uint32_t)n);
                j = (int)uint_below(o.y, (uint32_t)(n - 1));
                j += (j >= i);
                u_acc = u01(o.z, o.w);
            } else {
                if (n > 2) {
                    i = 0; j = 0;
                    for (uint32_t k = 0; k < 64 && i == j; ++k) {
                        U32x4 o = draw(cp.seed, t.stream, (uint32_t)cell, step, (uint32_t)tr,
                                       kPurposeCollide, 16 + k);
                        i = (int)floor(u01(o.x, o.y) * (n - 1));
                        j = (int)floor(u01(o.z, o.w) * (n - 1));
                    }
                    if (i == j) { i = 0; j = 1; }
                }
                U32x4 o = draw(cp.seed, t.stream, (uint32_t)cell, step, (uint32_t)tr,
                               kPurposeCollide, 1);
                u_acc = u01(o.x, o.y);
            }
            const int a = base + i, b = base + j;
            const double vax = vx[a], vay = vy[a], vaz = vz[a];
            const double vbx = vx[b], vby = vy[b], vbz = vz[b];
            const double dx = vax - vbx, dy = vay - vby, dz = vaz - vbz;
            const double magVrel = sqrt(dx * dx + dy * dy + dz * dz);
            if (magVrel * crossSection > smax) smax = magVrel * crossSection;
            const double ProbCol =
                t.alpha * lpc * cp.dt / cellVolume * magVrel * crossSection / lambda;
            if (ProbCol > u_acc) {
                const double cx = 0.5 * (vax + vbx), cy = 0.5 * (vay + vby), cz = 0.5 * (vaz + vbz);
                double nx, ny, nz;
                sphere_point(cp.seed, t.stream, (uint32_t)cell, step, (uint32_t)tr,
                             kPurposeCollide, 2, nx, ny, nz);
                const double wx = magVrel * nx, wy = magVrel * ny, wz = magVrel * nz;
                vx[a] = cx - 0.5 * wx; vy[a] = cy - 0.5 * wy; vz[a] = cz - 0.5 * wz;
                vx[b] = cx + 0.5 * wx; vy[b] = cy + 0.5 * wy; vz[b] = cz + 0.5 * wz;
                ++count;
            }
        }
        sigvrmax[cell] = smax;
        coll_count[cell] = count;
        mine += (unsigned long long)count;
    }
    mine = wave_sum(mine);
    if ((threadIdx.x & (kWave - 1)) == 0 && mine) atomicAdd(coll_total, mine);
}

// ---------------------------------------------------------------------------
// moveAtoms (moveAtoms.cu:23-91): velocity Verlet with stored acceleration.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_move_plain(TrapDev t, double dt, PlainPtrs s,
                                                       const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    double x = s.d[0][i], y = s.d[1][i], z = s.d[2][i];
    double vx = s.d[3][i], vy = s.d[4][i], vz = s.d[5][i];
    const bool up = s.up[i] != 0;
    D3 a = { s.d[6][i], s.d[7][i], s.d[8][i] };
    vx = vx + (0.5 * a.x) * dt; vy = vy + (0.5 * a.y) * dt; vz = vz + (0.5 * a.z) * dt;
    x = x + vx * dt; y = y + vy * dt; z = z + vz * dt;
    if (t.periodic) {
        x = wrap_periodic(x, t.hw[0]); y = wrap_periodic(y, t.hw[1]); z = wrap_periodic(z, t.hw[2]);
    }
    a = accel(t, x, y, z, up);
    vx = vx + (0.5 * a.x) * dt; vy = vy + (0.5 * a.y) * dt; vz = vz + (0.5 * a.z) * dt;
    s.d[0][i] = x; s.d[1][i] = y; s.d[2][i] = z;
    s.d[3][i] = vx; s.d[4][i] = vy; s.d[5][i] = vz;
    s.d[6][i] = a.x; s.d[7][i] = a.y; s.d[8][i] = a.z;
}

// ---------------------------------------------------------------------------
// evolveWavefunction (wavefunction.cu:11-73) and normalise (:279-299).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_spin_plain(TrapDev t, SpinParams sp, PlainPtrs s,
                                                       const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const double x0 = s.d[0][i], y0 = s.d[1][i], z0 = s.d[2][i];
    double hx = 0.0, hy = 0.0, hz = 0.0;
    if (sp.interp) {  // half-kicked velocity of the coming Verlet step
        hx = s.d[3][i] + (0.5 * s.d[6][i]) * sp.dt;
        hy = s.d[4][i] + (0.5 * s.d[7][i]) * sp.dt;
        hz = s.d[5][i] + (0.5 * s.d[8][i]) * sp.dt;
    }
    double ur = s.psi[0][i], ui = s.psi[1][i], dr = s.psi[2][i], di = s.psi[3][i];
    for (int m = 0; m < sp.substeps; ++m) {
        double x = x0, y = y0, z = z0;
        if (sp.interp) {
            const double tau = (m + 0.5) * sp.dt_spin;
            x = x0 + hx * tau; y = y0 + hy * tau; z = z0 + hz * tau;
        }
        D3 Bn;
        const double B = field_dir(t, x, y, z, Bn);
        const double theta = 0.5 * sp.gs * kMuB * B * sp.dt_spin / kHbar;
        const double sn = sin(theta), cs = cos(theta);
        const double nur = ur * cs + (Bn.x * di - Bn.y * dr + Bn.z * ui) * sn;
        const double nui = ui * cs + ((-Bn.x) * dr - Bn.y * di - Bn.z * ur) * sn;
        const double ndr = dr * cs + (Bn.x * ui + Bn.y * ur - Bn.z * di) * sn;
        const double ndi = di * cs + ((-Bn.x) * ur + Bn.y * ui + Bn.z * dr) * sn;
        ur = nur; ui = nui; dr = ndr; di = ndi;
    }
    s.psi[0][i] = ur; s.psi[1][i] = ui; s.psi[2][i] = dr; s.psi[3][i] = di;
}

__global__ __launch_bounds__(kBlock) void k_normalise_plain(PlainPtrs s,
                                                            const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const double ur = s.psi[0][i], ui = s.psi[1][i], dr = s.psi[2][i], di = s.psi[3][i];
    const double N = sqrt(ur * ur + ui * ui + dr * dr + di * di);
    s.psi[0][i] = ur / N; s.psi[1][i] = ui / N; s.psi[2][i] = dr / N; s.psi[3][i] = di / N;
}

// ---------------------------------------------------------------------------
// Majorana spin-flip model (wavefunction.cu:75-237): local populations before
// the motion substeps of a collision step, flips and decoherence after them.
// Populations are kept per atom label (the reference indexes them by atomID).
// ---------------------------------------------------------------------------
// projectLocalPopulations (wavefunction.cu:156-173): populations of the
// spinor in the spin-up and spin-down states of the local field axis.
__device__ __forceinline__ void local_pops(const TrapDev &t, double x, double y, double z,
                                           double ur, double ui, double dr, double di,
                                           double &pu, double &pd)
{
    D3 Bn;
    (void)field_dir(t, x, y, z, Bn);
    const double re = ur * dr + ui * di, im = ui * dr - ur * di;
    const double w = 0.5 - ur * ur - ui * ui;
    pu = 0.5 + Bn.x * re - Bn.y * im - Bn.z * w;
    pd = 0.5 - Bn.x * re + Bn.y * im + Bn.z * w;
}

__global__ __launch_bounds__(kBlock) void k_local_pops_plain(TrapDev t, PlainPtrs s,
                                                             double *__restrict__ pops,
                                                             const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    double pu, pd;
    local_pops(t, s.d[0][i], s.d[1][i], s.d[2][i], s.psi[0][i], s.psi[1][i], s.psi[2][i],
               s.psi[3][i], pu, pd);
    const size_t id = (size_t)s.id[i];
    pops[2 * id] = pu;
    pops[2 * id + 1] = pd;
}

// flipAtoms (wavefunction.cu:99-154): an atom changes spin state with the
// probability that its population moved into the other state since the
// populations were taken, if its kinetic energy can pay the potential step
// 2 mu |B| (gs muB |B| for the reference's mu); the speed then absorbs the
// step along the old direction.  The uniform is Philox (label, step, 0,
// purpose 3).
__global__ __launch_bounds__(kBlock) void k_flip_plain(TrapDev t, FlipParams fp, PlainPtrs s,
                                                       const double *__restrict__ pops,
                                                       unsigned long long *__restrict__ flips,
                                                       const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long cnt = 0ull;
    if (i < *n_dev) {
        const double x = s.d[0][i], y = s.d[1][i], z = s.d[2][i];
        const double vx = s.d[3][i], vy = s.d[4][i], vz = s.d[5][i];
        const uint32_t id = (uint32_t)s.id[i];
        const bool up = s.up[i] != 0;
        const double opu = pops[2 * (size_t)id], opd = pops[2 * (size_t)id + 1];
        double npu, npd;
        local_pops(t, x, y, z, s.psi[0][i], s.psi[1][i], s.psi[2][i], s.psi[3][i], npu, npd);
        D3 Bn;
        const double magB = field_dir(t, x, y, z, Bn);
        double Pflip, dEp;
        if (up) {
            Pflip = (npd - opd) / opu;
            dEp = fp.gap * magB;
        } else {
            Pflip = (npu - opu) / opd;
            dEp = (-1.0 * fp.gap) * magB;
        }
        const double Ek = 0.5 * fp.mass * (vx * vx + vy * vy + vz * vz);
        const U32x4 o = draw(fp.seed, t.stream, id, fp.step, 0u, kPurposeFlip, 0u);
        if (Pflip > u01(o.x, o.y) && Ek + dEp > 0.0) {
            s.up[i] = up ? 0 : 1;
            const double magVel = sqrt(vx * vx + vy * vy + vz * vz);
            const double speed = sqrt(2.0 * (Ek + dEp) / fp.mass);
            if (magVel > 0.0) {
                s.d[3][i] = speed * (vx / magVel);
                s.d[4][i] = speed * (vy / magVel);
                s.d[5][i] = speed * (vz / magVel);
            } else {   // no direction to keep: along +z
                s.d[3][i] = 0.0; s.d[4][i] = 0.0; s.d[5][i] = speed;
            }
            cnt = 1ull;
        }
    }
    cnt = wave_sum(cnt);
    if ((threadIdx.x & (kWave - 1)) == 0 && cnt) atomicAdd(flips, cnt);
}

// calculateTau (wavefunction.cu:206-219) with getAcc (:221-237), the spin-up
// acceleration: the decoherence time from the position- and velocity-limited
// times of an atom that the field gradient separates.
__device__ __forceinline__ double decoherence_tau(const TrapDev &t, double sigma, double mass,
                                                  double x, double y, double z)
{
    const D3 a = accel(t, x, y, z, true);
    const double relAccel = 2.0 * sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    const double tauPos = 1.163 * sqrt(sigma / relAccel);
    const double tauVel = sqrt(2. / kPiRef) * kHbar / (mass * sigma * relAccel);
    return pow(1 / (tauPos * tauPos * tauPos) + 1 / (tauVel * tauVel * tauVel), -1. / 3.);
}

// exponentialDecay (wavefunction.cu:175-204): psi <- P+ psi + e P- psi with
// P+- = (1 +- B.sigma) / 2 and e = exp(-dt / tau), written out per component.
__global__ __launch_bounds__(kBlock) void k_decay_plain(TrapDev t, DecayParams dp, PlainPtrs s,
                                                        const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_dev) return;
    const double x = s.d[0][i], y = s.d[1][i], z = s.d[2][i];
    const double ur = s.psi[0][i], ui = s.psi[1][i], dr = s.psi[2][i], di = s.psi[3][i];
    D3 Bn;
    (void)field_dir(t, x, y, z, Bn);
    const double tau = decoherence_tau(t, dp.sigma, dp.mass, x, y, z);
    const double om = 1. - exp(-dp.dt / tau);
    const double op = 1. + exp(-dp.dt / tau);
    s.psi[0][i] = 0.5 * ((Bn.x * dr + Bn.y * di + Bn.z * ur) * om + ur * op);
    s.psi[1][i] = 0.5 * ((Bn.x * di - Bn.y * dr + Bn.z * ui) * om + ui * op);
    s.psi[2][i] = 0.5 * ((Bn.x * ur - Bn.y * ui - Bn.z * dr) * om + dr * op);
    s.psi[3][i] = 0.5 * ((Bn.x * ui + Bn.y * ur - Bn.z * di) * om + di * op);
}

// ---------------------------------------------------------------------------
// evaporationTag (evaporation.cu:68-106) plus the energy cut.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_evap_tag_plain(TrapDev t, EvapParams ep, PlainPtrs s,
                                                           int *__restrict__ tag,
                                                           double *__restrict__ evap_pos,
                                                           double *__restrict__ evap_vel,
                                                           double *__restrict__ evap_accum,
                                                           const int *__restrict__ n_dev)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double e_lost = 0.0, n_lost = 0.0;
    if (i < *n_dev) {
        const double x = s.d[0][i], y = s.d[1][i], z = s.d[2][i];
        const double vx = s.d[3][i], vy = s.d[4][i], vz = s.d[5][i];
        const double E = 0.5 * ep.mass * (vx * vx + vy * vy + vz * vz) + potential_energy(t, x, y, z);
        bool lost = false;
        if (ep.spin_loss) {
            D3 Bn;
            (void)field_dir(t, x, y, z, Bn);
            const double ur = s.psi[0][i], ui = s.psi[1][i], dr = s.psi[2][i], di = s.psi[3][i];
            const double proj = 2. * Bn.x * (ur * dr + ui * di) + 2. * Bn.y * (ur * di - ui * dr) +
                                2. * Bn.z * (ur * ur + ui * ui - 0.5);
            lost = proj < 0.0;
        }
        if (ep.e_cut > 0.0) lost = lost || E > ep.e_cut;
        tag[i] = lost ? 1 : 0;
        if (lost) {
            const int id = s.id[i];
            evap_pos[3 * (size_t)id + 0] = x; evap_pos[3 * (size_t)id + 1] = y;
            evap_pos[3 * (size_t)id + 2] = z;
            evap_vel[3 * (size_t)id + 0] = vx; evap_vel[3 * (size_t)id + 1] = vy;
            evap_vel[3 * (size_t)id + 2] = vz;
            e_lost = E;
            n_lost = 1.0;
        }
    }
    e_lost = wave_sum(e_lost);
    n_lost = wave_sum(n_lost);
    if ((threadIdx.x & (kWave - 1)) == 0 && n_lost > 0.0) {
        atomicAdd(&evap_accum[0], e_lost);
        atomicAdd(&evap_accum[1], n_lost);
    }
}

// ---------------------------------------------------------------------------
// FP64 observables (calculateTemp, evaporation.cu:108-189, and the energy of
// energyChecker.py:117-122).  Per-block partials; the last block to finish
// adds them in a fixed order (reproducible) and appends one log record.
// ---------------------------------------------------------------------------
constexpr int kObsK = 5;  // ek, ep, vx^2, vy^2, vz^2

__global__ __launch_bounds__(kBlock) void k_observe_plain(TrapDev t, double mass, PlainPtrs s,
                                                          const int *__restrict__ n_dev,
                                                          double *__restrict__ partials)
{
    __shared__ double tmp[(kBlock / kWave) * kObsK];
    const int n = *n_dev;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[kObsK] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
    if (i < n) {
        const double x = s.d[0][i], y = s.d[1][i], z = s.d[2][i];
        const double vx = s.d[3][i], vy = s.d[4][i], vz = s.d[5][i];
        v[0] = 0.5 * mass * (vx * vx + vy * vy + vz * vz);
        // U_up = mu (|B| - B0); a spin-down atom (after a flip) has
        // U_down = -mu (|B| + B0), so that a flip conserves E_k + U
        const double u = potential_energy(t, x, y, z);
        v[1] = s.up[i] ? u : -u - 2.0 * t.mu * t.B0;
        v[2] = vx * vx; v[3] = vy * vy; v[4] = vz * vz;
    }
    block_sum<kBlock, kObsK>(v, tmp);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kObsK; ++k) partials[(size_t)blockIdx.x * kObsK + k] = v[k];
    }
}

// One block adds the block partials in a fixed order (reproducible) and
// writes the log record; it also takes and clears the collision and
// evaporation accumulators of the step.
__global__ __launch_bounds__(kBlock) void k_observe_finalize_plain(
    const int *__restrict__ n_dev, const double *__restrict__ partials, int nb,
    dsmc_observables *__restrict__ out, unsigned long long *__restrict__ coll_accum,
    double *__restrict__ evap_accum, long long step)
{
    __shared__ double tmp[(kBlock / kWave) * kObsK];
    const int n = *n_dev;
    const int used = (n + kBlock - 1) / kBlock;
    double acc[kObsK] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
    for (int b = threadIdx.x; b < used && b < nb; b += kBlock) {
#pragma unroll
        for (int k = 0; k < kObsK; ++k) acc[k] += partials[(size_t)b * kObsK + k];
    }
    block_sum<kBlock, kObsK>(acc, tmp);
    if (threadIdx.x == 0) {
        out->ek = acc[0];
        out->ep = acc[1];
        out->v2[0] = acc[2]; out->v2[1] = acc[3]; out->v2[2] = acc[4];
        out->e_evap = evap_accum[0];
        out->n_alive = n;
        out->n_evap = (long long)evap_accum[1];
        out->collisions = (long long)*coll_accum;
        out->step = step;
        evap_accum[0] = 0.0; evap_accum[1] = 0.0;
        *coll_accum = 0ull;
    }
}

// ---------------------------------------------------------------------------
// Initial distribution (setUp.cu:129-252): FP64 outputs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_init_plain(TrapDev t, InitParams ip, PlainPtrs s,
                                                       int n)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    double px, py, pz, vx, vy, vz;
    init_atom(t, ip, (uint32_t)a, px, py, pz, vx, vy, vz);
    s.d[0][a] = px; s.d[1][a] = py; s.d[2][a] = pz;
    s.d[3][a] = vx; s.d[4][a] = vy; s.d[5][a] = vz;
    const D3 acc = accel(t, px, py, pz, true);
    s.d[6][a] = acc.x; s.d[7][a] = acc.y; s.d[8][a] = acc.z;
    s.id[a] = a;
    s.up[a] = 1;
    if (ip.spin) {
        // getAlignedSpinUp/Dn (setUp.cu:236-252)
        D3 Bn;
        (void)field_dir(t, px, py, pz, Bn);
        const double rs = 1.0 / sqrt(1 + Bn.x);
        s.psi[0][a] = 0.5 * (1. + Bn.x + Bn.z) * rs;
        s.psi[1][a] = 0.5 * (-Bn.y) * rs;
        s.psi[2][a] = 0.5 * (1. + Bn.x - Bn.z) * rs;
        s.psi[3][a] = 0.5 * Bn.y * rs;
    }
}

// Accelerations for externally supplied positions (dsmc_set_state).
__global__ __launch_bounds__(kBlock) void k_accel_plain(TrapDev t, PlainPtrs s, int n)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    const D3 acc = accel(t, s.d[0][a], s.d[1][a], s.d[2][a], s.up[a] != 0);
    s.d[6][a] = acc.x; s.d[7][a] = acc.y; s.d[8][a] = acc.z;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static int grid_for(int n) { return div_up(n > 0 ? n : 1, kBlock); }

void launch_index_plain(const TrapDev &t, const PlainPtrs &s, uint32_t *key, const int *n_dev,
                        int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "index_plain", (double)n_cap * 28.0, k_index_plain, grid_for(n_cap), kBlock,
                st, t, s.d[0], s.d[1], s.d[2], key, n_dev);
}

void launch_gather_plain(const PlainPtrs &dst, const PlainPtrs &src, const uint32_t *map,
                         const int *n_dev, int n_cap, int spin, hipStream_t st, Profiler *prof)
{
    const double per = 4.0 + 2.0 * (9 * 8 + 4 + 1 + (spin ? 32 : 0));
    DSMC_LAUNCH(prof, "gather_plain", (double)n_cap * per, k_gather_plain, grid_for(n_cap), kBlock,
                st, dst, src, map, n_dev, spin);
}

void launch_cell_table_plain(const uint32_t *key, const int *n_dev, int n_cap, int2 *se,
                             int *count, int *prefix, int *partial, int ncells,
                             hipStream_t st, Profiler *prof)
{
    fill_u32((uint32_t *)se, 0xFFFFFFFFu, 2LL * (ncells + 1), st, prof);
    DSMC_LAUNCH(prof, "cell_start_end", (double)n_cap * 4.0, k_cell_start_end, grid_for(n_cap),
                kBlock, st, key, n_dev, se);
    DSMC_LAUNCH(prof, "cell_count", (double)(ncells + 1) * 12.0, k_cell_count,
                grid_for(ncells + 1), kBlock, st, (const int2 *)se, count, ncells + 1);
    exclusive_scan(count, prefix, ncells + 1, partial, st, prof);
}

void launch_collide_plain(const TrapDev &t, const CollideParams &cp, const PlainPtrs &s,
                          double *sigvrmax, const int *prefix, int *coll_count,
                          unsigned long long *coll_total, int ncells, uint32_t step,
                          hipStream_t st, Profiler *prof)
{
    const int grid = std::min(grid_for(ncells), 8192);
    DSMC_LAUNCH(prof, "collide_plain", (double)ncells * 20.0, k_collide_plain, grid, kBlock, st, t,
                cp, s.d[3], s.d[4], s.d[5], sigvrmax, prefix, coll_count, coll_total, ncells,
                step);
}

void launch_move_plain(const TrapDev &t, double dt, const PlainPtrs &s, const int *n_dev,
                       int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "move_plain", (double)n_cap * 145.0, k_move_plain, grid_for(n_cap), kBlock,
                st, t, dt, s, n_dev);
}

void launch_spin_plain(const TrapDev &t, const SpinParams &sp, const PlainPtrs &s,
                       const int *n_dev, int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "spin_plain", (double)n_cap * 88.0, k_spin_plain, grid_for(n_cap), kBlock,
                st, t, sp, s, n_dev);
}

void launch_normalise_plain(const PlainPtrs &s, const int *n_dev, int n_cap, hipStream_t st,
                            Profiler *prof)
{
    DSMC_LAUNCH(prof, "normalise_plain", (double)n_cap * 64.0, k_normalise_plain,
                grid_for(n_cap), kBlock, st, s, n_dev);
}

void launch_evap_tag_plain(const TrapDev &t, const EvapParams &ep, const PlainPtrs &s, int *tag,
                           double *evap_pos, double *evap_vel, double *evap_accum,
                           const int *n_dev, int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "evap_tag_plain", (double)n_cap * 84.0, k_evap_tag_plain, grid_for(n_cap),
                kBlock, st, t, ep, s, tag, evap_pos, evap_vel, evap_accum, n_dev);
}

void launch_local_pops_plain(const TrapDev &t, const PlainPtrs &s, double *pops, const int *n_dev,
                             int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "local_pops_plain", (double)n_cap * 76.0, k_local_pops_plain,
                grid_for(n_cap), kBlock, st, t, s, pops, n_dev);
}

void launch_flip_plain(const TrapDev &t, const FlipParams &fp, const PlainPtrs &s,
                       const double *pops, unsigned long long *flips, const int *n_dev, int n_cap,
                       hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "flip_plain", (double)n_cap * 101.0, k_flip_plain, grid_for(n_cap), kBlock,
                st, t, fp, s, pops, flips, n_dev);
}

void launch_decay_plain(const TrapDev &t, const DecayParams &dp, const PlainPtrs &s,
                        const int *n_dev, int n_cap, hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "decay_plain", (double)n_cap * 88.0, k_decay_plain, grid_for(n_cap), kBlock,
                st, t, dp, s, n_dev);
}

int observe_plain_blocks(int n_cap) { return grid_for(n_cap); }

void launch_observe_plain(const TrapDev &t, double mass, const PlainPtrs &s, const int *n_dev,
                          int n_cap, double *partials, dsmc_observables *out,
                          unsigned long long *coll_accum, double *evap_accum, long long step,
                          hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "observe_plain", (double)n_cap * 48.0, k_observe_plain, grid_for(n_cap),
                kBlock, st, t, mass, s, n_dev, partials);
    DSMC_LAUNCH(prof, "observe_finalize", (double)grid_for(n_cap) * 40.0, k_observe_finalize_plain,
                1, kBlock, st, n_dev, (const double *)partials, grid_for(n_cap), out, coll_accum,
                evap_accum, step);
}

void launch_init_plain(const TrapDev &t, const InitParams &ip, const PlainPtrs &s, int n,
                       hipStream_t st, Profiler *prof)
{
    DSMC_LAUNCH(prof, "init_plain", (double)n * 150.0, k_init_plain, grid_for(n), kBlock, st, t,
                ip, s, n);
}

void launch_accel_plain(const TrapDev &t, const PlainPtrs &s, int n, hipStream_t st,
                        Profiler *prof)
{
    DSMC_LAUNCH(prof, "accel_plain", (double)n * 48.0, k_accel_plain, grid_for(n), kBlock, st, t,
                s, n);
}

}  // namespace dsmc
