/*
 * dsmc_oracle.c -- CPU FP64 restatement of CUDADSMC's per-timestep DSMC step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): nothing in the
 * product library links or calls this file.  Build: oracle/Makefile
 * (gcc -O2 -fopenmp -ffp-contract=off).  See dsmc_oracle.h for the layout and
 * the list of deliberate deviations from the reference.
 *
 * Reference citations are relative to the CUDADSMC repository root.
 */
#include "dsmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Round: two 32x32->64 products of counter words */
/* 0 and 2 with the Philox multipliers, outputs permuted and XORed with the  */
/* key; the key is bumped by the Weyl constants between rounds.              */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 52 random bits on the grid (k + 1/2) 2^-52: strictly inside (0,1) and
 * symmetric about 1/2.  Every step is exact in binary64. */
double orc_u01(uint32_t hi, uint32_t lo)
{
    uint64_t u = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
    return ((double)u + 0.5) * 0x1p-52;
}

uint32_t orc_uint_below(uint32_t r, uint32_t n)
{
    return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

/* Counter word 3 carries (purpose << 24) | draw index. */
#define ORC_PURPOSE_COLLIDE 1u
#define ORC_PURPOSE_INIT    2u
#define ORC_PURPOSE_FLIP    3u

/* Key = (seed low word, seed high word + stream): each trap has its own
 * Philox stream.  Counter = (a, b, c, (purpose << 24) | draw). */
static inline void orc_draw(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b, uint32_t c,
                            uint32_t purpose, uint32_t draw, uint32_t out[4])
{
    uint32_t ctr[4] = { a, b, c, (purpose << 24) | draw };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) + stream };
    orc_philox(ctr, key, out);
}

/* Marsaglia (1972) point on the unit sphere: (u1, u2) uniform on the disc
 * s = u1^2 + u2^2 < 1, n = (2 u1 sqrt(1-s), 2 u2 sqrt(1-s), 1 - 2 s).
 * Draws first_draw .. first_draw+31; falls back to +z after 32 rejections
 * (probability 0.215^32).  Replaces collisions.cu:537-546. */
static orc_d3 orc_sphere(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b, uint32_t c,
                         uint32_t purpose, uint32_t first_draw)
{
    for (uint32_t k = 0; k < 32; ++k) {
        uint32_t o[4];
        orc_draw(seed, stream, a, b, c, purpose, first_draw + k, o);
        double u1 = 2.0 * orc_u01(o[0], o[1]) - 1.0;
        double u2 = 2.0 * orc_u01(o[2], o[3]) - 1.0;
        double s = u1 * u1 + u2 * u2;
        if (s < 1.0) {
            double q = 2.0 * sqrt(1.0 - s);
            orc_d3 n = { u1 * q, u2 * q, 1.0 - 2.0 * s };
            return n;
        }
    }
    orc_d3 z = { 0.0, 0.0, 1.0 };
    return z;
}

/* Box-Muller: two standard normals from one Philox block. */
static void orc_box_muller(const uint32_t o[4], double *z0, double *z1)
{
    const double two_pi = 6.283185307179586;
    double u1 = orc_u01(o[0], o[1]);
    double u2 = orc_u01(o[2], o[3]);
    double r = sqrt(-2.0 * log(u1));
    *z0 = r * cos(two_pi * u2);
    *z1 = r * sin(two_pi * u2);
}

/* ------------------------------------------------------------------------ */
/* Trap field.                                                               */
/* ------------------------------------------------------------------------ */

/* Ioffe-Pritchard field, src/moveAtoms.cu:135-142 (operand order kept). */
static orc_d3 ip_field(const orc_params *p, orc_d3 r)
{
    double c = 0.5 * p->d2Bdx2;
    orc_d3 B;
    B.x = (p->B0 * 0.0 + p->dBdx * r.x) + c * ((-r.x) * r.z);
    B.y = (p->B0 * 0.0 + p->dBdx * (-r.y)) + c * ((-r.y) * r.z);
    B.z = (p->B0 * 1.0 + p->dBdx * 0.0) + c * (r.z * r.z - 0.5 * (r.x * r.x + r.y * r.y));
    return B;
}

static double quad_rho(orc_d3 r)
{
    return sqrt(r.x * r.x + r.y * r.y + 4.0 * (r.z * r.z));
}

orc_d3 orc_field(const orc_params *p, orc_d3 r)
{
    orc_d3 B = { 0.0, 0.0, 0.0 };
    if (p->kind == ORC_TRAP_IP) {
        B = ip_field(p, r);
    } else if (p->kind == ORC_TRAP_QUAD) {
        /* |B| = B' sqrt(x^2 + y^2 + 4 z^2): energyComparison.py:53 */
        B.x = p->dBdz * r.x;
        B.y = p->dBdz * r.y;
        B.z = p->dBdz * (-2.0 * r.z);
    }
    return B;
}

/* |B| and unit vector B/|B| (moveAtoms.cu:119-133 for IP). */
static double field_mag_dir(const orc_params *p, orc_d3 r, orc_d3 *Bn)
{
    if (p->kind == ORC_TRAP_IP) {
        orc_d3 B = ip_field(p, r);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        Bn->x = B.x / len; Bn->y = B.y / len; Bn->z = B.z / len;
        return len;
    }
    if (p->kind == ORC_TRAP_QUAD) {
        double rho = quad_rho(r);
        if (rho > 0.0) {
            Bn->x = r.x / rho; Bn->y = r.y / rho; Bn->z = (-2.0 * r.z) / rho;
        } else {
            Bn->x = 0.0; Bn->y = 0.0; Bn->z = 1.0;
        }
        return p->dBdz * rho;
    }
    Bn->x = 0.0; Bn->y = 0.0; Bn->z = 1.0;
    return 0.0;
}

/* updateAcc, src/moveAtoms.cu:93-116 (IP) and the quadrupole closed form
 * a = -(mu B'/m) (x, y, 4 z) / rho. */
orc_d3 orc_accel(const orc_params *p, orc_d3 r, int up)
{
    orc_d3 a = { 0.0, 0.0, 0.0 };
    double potential = (-p->mu) / p->mass;
    if (p->kind == ORC_TRAP_IP) {
        orc_d3 B = ip_field(p, r);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        orc_d3 Bn = { B.x / len, B.y / len, B.z / len };
        orc_d3 dx = { p->dBdx - (0.5 * p->d2Bdx2) * r.z, 0.0, ((-0.5) * p->d2Bdx2) * r.x };
        orc_d3 dy = { 0.0, (-p->dBdx) - (0.5 * p->d2Bdx2) * r.z, ((-0.5) * p->d2Bdx2) * r.y };
        orc_d3 dz = { ((-0.5) * p->d2Bdx2) * r.x, ((-0.5) * p->d2Bdx2) * r.y, p->d2Bdx2 * r.z };
        double gx = dx.x * Bn.x + dx.y * Bn.y + dx.z * Bn.z;
        double gy = dy.x * Bn.x + dy.y * Bn.y + dy.z * Bn.z;
        double gz = dz.x * Bn.x + dz.y * Bn.y + dz.z * Bn.z;
        if (up) {
            a.x = potential * gx; a.y = potential * gy; a.z = potential * gz;
        } else {
            a.x = (-1.0 * potential) * gx; a.y = (-1.0 * potential) * gy; a.z = (-1.0 * potential) * gz;
        }
    } else if (p->kind == ORC_TRAP_QUAD) {
        double rho = quad_rho(r);
        if (rho > 0.0) {
            double g = (potential * p->dBdz) / rho;
            if (!up) g = -g;
            a.x = g * r.x; a.y = g * r.y; a.z = g * (4.0 * r.z);
        }
    }
    return a;
}

/* Potential energy: energyChecker.py:121 (IP, spin up), energyComparison.py:53
 * (quadrupole) with the moment explicit. */
double orc_potential(const orc_params *p, orc_d3 r)
{
    if (p->kind == ORC_TRAP_IP) {
        orc_d3 B = ip_field(p, r);
        double len = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
        return p->mu * (len - p->B0);
    }
    if (p->kind == ORC_TRAP_QUAD)
        return p->mu * (p->dBdz * quad_rho(r));
    return 0.0;
}

/* ------------------------------------------------------------------------ */
/* Indexing (src/collisions.cu:163-226, include/vectorMath.cuh:56-59).       */
/* ------------------------------------------------------------------------ */
static int cell_of(const orc_params *p, orc_d3 r)
{
    int idx[3];
    double q[3];
    const double pr[3] = { r.x, r.y, r.z };
    for (int d = 0; d < 3; ++d) {
        double gmin = -1.0 * p->half_width[d];
        double len = (2.0 * p->half_width[d]) / (double)p->nc[d];
        q[d] = floor((pr[d] - gmin) / len);
        /* __double2int_rd + the range test of getCellID; NaN fails both
         * comparisons and lands in the overflow bin */
        if (!(q[d] > -1.0 && q[d] < (double)p->nc[d]))
            return p->nc[0] * p->nc[1] * p->nc[2];
        idx[d] = (int)q[d];
    }
    return idx[2] * p->nc[0] * p->nc[1] + idx[1] * p->nc[0] + idx[0];
}

/* findAtomIndex (collisions.cu:163-185).  d1_quirk = 1 writes the key at the
 * physical index cellID[atomID[slot]] exactly as the reference does (:181);
 * d1_quirk = 0 writes key[slot], the intended slot-aligned semantics. */
void orc_index_atoms(const orc_params *p, const orc_d3 *pos, int *cellID,
                     const int *atomID, int n, int d1_quirk)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        int c = cell_of(p, pos[a]);
        if (d1_quirk) cellID[a] = c; else cellID[s] = c;
    }
}

/* thrust::sort_by_key(cellID, atomID) (collisions.cu:275-287): a stable
 * counting sort; thrust's radix sort of int keys is stable. */
void orc_sort_arrays(int *cellID, int *atomID, int n, int nkeys)
{
    int *cnt = (int *)calloc((size_t)nkeys + 1, sizeof(int));
    int *k2 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *v2 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) cnt[cellID[i] + 1]++;
    for (int k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
    for (int i = 0; i < n; ++i) {
        int d = cnt[cellID[i]]++;
        k2[d] = cellID[i];
        v2[d] = atomID[i];
    }
    memcpy(cellID, k2, sizeof(int) * (size_t)n);
    memcpy(atomID, v2, sizeof(int) * (size_t)n);
    free(cnt); free(k2); free(v2);
}

/* deviceMemset(int2,-1) + cellStartandEndKernel + findNumberOfAtomsInCell +
 * thrust::exclusive_scan (main.cu:426-442, collisions.cu:228-271,
 * cudaHelpers.cu:90-99).  start_end holds (start, end) pairs. */
void orc_cell_table(const int *cellID, int n, int ncells, int *start_end,
                    int *count, int *prefix)
{
    for (int c = 0; c <= ncells; ++c) { start_end[2 * c] = -1; start_end[2 * c + 1] = -1; }
    for (int s = 0; s < n; ++s) {
        if (s == 0) start_end[2 * cellID[s]] = 0;
        else if (cellID[s] != cellID[s - 1]) start_end[2 * cellID[s]] = s;
        if (s == n - 1) start_end[2 * cellID[s] + 1] = n - 1;
        else if (cellID[s] != cellID[s + 1]) start_end[2 * cellID[s] + 1] = s;
    }
    for (int c = 0; c <= ncells; ++c)
        count[c] = start_end[2 * c] == -1 ? 0 : start_end[2 * c + 1] - start_end[2 * c] + 1;
    int run = 0;
    for (int c = 0; c <= ncells; ++c) { prefix[c] = run; run += count[c]; }
}

/* ------------------------------------------------------------------------ */
/* Collisions (src/collisions.cu:291-445, 529-546).                         */
/* ------------------------------------------------------------------------ */
long long orc_collide(const orc_params *p, orc_d3 *vel, double *sigvrmax,
                      const int *prefix, int *collisionCount, const int *atomID,
                      int ncells, uint32_t step)
{
    double lx = (2.0 * p->half_width[0]) / (double)p->nc[0];
    double ly = (2.0 * p->half_width[1]) / (double)p->nc[1];
    double lz = (2.0 * p->half_width[2]) / (double)p->nc[2];
    double cellVolume = lx * ly * lz;
    double crossSection = p->sigma;
    double lpc = (double)p->lpc;
    long long total = 0;

#pragma omp parallel for schedule(dynamic, 512) reduction(+ : total)
    for (int cell = 0; cell < ncells; ++cell) {
        int n = prefix[cell + 1] - prefix[cell];
        /* collisions.cu:308,335 reset the count only in cells of two or more
         * atoms, so a cell that empties keeps a stale count; the count of
         * every cell here is that of this call */
        if (n <= 1) { collisionCount[cell] = 0; continue; }
        double Mc = 0.5 * (n - 1) * n;
        double X = Mc * p->alpha * lpc * p->dt * sigvrmax[cell] / cellVolume;
        double lambda = ceil(X) / Mc;
        /* collisions.cu:322 truncates Mc * (ceil(X) / Mc), which is 1 ulp below
         * ceil(X) for some Mc (n = 23, 34, 59, ...) and then drops a trial; the
         * reference pair mode keeps that, the unbiased mode uses ceil(X). */
        int Ncol = p->pair_mode == ORC_PAIR_REFERENCE ? (int)(Mc * lambda) : (int)ceil(X);
        int base = prefix[cell];
        collisionCount[cell] = 0;
        for (int t = 0; t < Ncol; ++t) {
            uint32_t o[4];
            int i = 0, j = 1;
            double u_acc;
            if (p->pair_mode == ORC_PAIR_UNIFORM) {
                orc_draw(p->seed, p->stream, (uint32_t)cell, step, (uint32_t)t, ORC_PURPOSE_COLLIDE, 0, o);
                i = (int)orc_uint_below(o[0], (uint32_t)n);
                j = (int)orc_uint_below(o[1], (uint32_t)(n - 1));
                j += (j >= i);
                u_acc = orc_u01(o[2], o[3]);
            } else {
                /* collisions.cu:400-413 incl. its floor(u (n-1)) bias (D2) */
                if (n > 2) {
                    i = 0; j = 0;
                    for (uint32_t k = 0; k < 64 && i == j; ++k) {
                        orc_draw(p->seed, p->stream, (uint32_t)cell, step, (uint32_t)t, ORC_PURPOSE_COLLIDE, 16 + k, o);
                        i = (int)floor(orc_u01(o[0], o[1]) * (n - 1));
                        j = (int)floor(orc_u01(o[2], o[3]) * (n - 1));
                    }
                    if (i == j) { i = 0; j = 1; }
                }
                orc_draw(p->seed, p->stream, (uint32_t)cell, step, (uint32_t)t, ORC_PURPOSE_COLLIDE, 1, o);
                u_acc = orc_u01(o[0], o[1]);
            }
            int a = atomID[base + i];
            int b = atomID[base + j];
            orc_d3 va = vel[a], vb = vel[b];
            double dx = va.x - vb.x, dy = va.y - vb.y, dz = va.z - vb.z;
            double magVrel = sqrt(dx * dx + dy * dy + dz * dz);
            if (magVrel * crossSection > sigvrmax[cell])
                sigvrmax[cell] = magVrel * crossSection;
            double ProbCol = p->alpha * lpc * p->dt / cellVolume * magVrel * crossSection / lambda;
            if (ProbCol > u_acc) {
                orc_d3 vcm = { 0.5 * (va.x + vb.x), 0.5 * (va.y + vb.y), 0.5 * (va.z + vb.z) };
                orc_d3 nh = orc_sphere(p->seed, p->stream, (uint32_t)cell, step, (uint32_t)t,
                                       ORC_PURPOSE_COLLIDE, 2);
                orc_d3 nv = { magVrel * nh.x, magVrel * nh.y, magVrel * nh.z };
                vel[a].x = vcm.x - 0.5 * nv.x; vel[a].y = vcm.y - 0.5 * nv.y; vel[a].z = vcm.z - 0.5 * nv.z;
                vel[b].x = vcm.x + 0.5 * nv.x; vel[b].y = vcm.y + 0.5 * nv.y; vel[b].z = vcm.z + 0.5 * nv.z;
                collisionCount[cell]++;
                total++;
            }
        }
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* Motion (src/moveAtoms.cu:23-91): velocity Verlet with stored acceleration. */
/* ------------------------------------------------------------------------ */
static double wrap(double x, double hw)
{
    if (x >= hw) x -= 2.0 * hw;
    else if (x < -hw) x += 2.0 * hw;
    return x;
}

void orc_move_atoms(const orc_params *p, orc_d3 *pos, orc_d3 *vel, orc_d3 *acc,
                    const uint8_t *spinUp, const int *atomID, int n)
{
    const double dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 r = pos[a], v = vel[a], g = acc[a];
        int up = spinUp ? spinUp[a] : 1;
        v.x = v.x + (0.5 * g.x) * dt; v.y = v.y + (0.5 * g.y) * dt; v.z = v.z + (0.5 * g.z) * dt;
        r.x = r.x + v.x * dt; r.y = r.y + v.y * dt; r.z = r.z + v.z * dt;
        if (p->periodic) {
            r.x = wrap(r.x, p->half_width[0]);
            r.y = wrap(r.y, p->half_width[1]);
            r.z = wrap(r.z, p->half_width[2]);
        }
        g = orc_accel(p, r, up);
        v.x = v.x + (0.5 * g.x) * dt; v.y = v.y + (0.5 * g.y) * dt; v.z = v.z + (0.5 * g.z) * dt;
        pos[a] = r; vel[a] = v; acc[a] = g;
    }
}

/* ------------------------------------------------------------------------ */
/* Spin (src/wavefunction.cu:11-73, 279-299).                                */
/* ------------------------------------------------------------------------ */
void orc_evolve_wavefunction(const orc_params *p, const orc_d3 *pos, orc_c *psiUp,
                             orc_c *psiDn, const int *atomID, int n, double dt_spin)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 Bn;
        double B = field_mag_dir(p, pos[a], &Bn);
        double theta = 0.5 * p->gs * ORC_MUB * B * dt_spin / ORC_HBAR;
        double sn = sin(theta), cs = cos(theta);
        orc_c u = psiUp[a], d = psiDn[a], nu, nd;
        nu.re = u.re * cs + (Bn.x * d.im - Bn.y * d.re + Bn.z * u.im) * sn;
        nu.im = u.im * cs + ((-Bn.x) * d.re - Bn.y * d.im - Bn.z * u.re) * sn;
        nd.re = d.re * cs + (Bn.x * u.im + Bn.y * u.re - Bn.z * d.im) * sn;
        nd.im = d.im * cs + ((-Bn.x) * u.re + Bn.y * u.im + Bn.z * d.re) * sn;
        psiUp[a] = nu; psiDn[a] = nd;
    }
}

void orc_normalise(orc_c *psiUp, orc_c *psiDn, const int *atomID, int n)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_c u = psiUp[a], d = psiDn[a];
        double N = sqrt(u.re * u.re + u.im * u.im + d.re * d.re + d.im * d.im);
        psiUp[a].re = u.re / N; psiUp[a].im = u.im / N;
        psiDn[a].re = d.re / N; psiDn[a].im = d.im / N;
    }
}

/* ------------------------------------------------------------------------ */
/* Majorana spin-flip model (src/wavefunction.cu:75-237, main.cu:457-462,    */
/* 501-516).  Populations are stored per atom label, two doubles each.       */
/* ------------------------------------------------------------------------ */
/* projectLocalPopulations (wavefunction.cu:156-173). */
static void local_pops(const orc_params *p, orc_d3 r, orc_c u, orc_c d, double *pu, double *pd)
{
    orc_d3 Bn;
    (void)field_mag_dir(p, r, &Bn);
    double re = u.re * d.re + u.im * d.im, im = u.im * d.re - u.re * d.im;
    double w = 0.5 - u.re * u.re - u.im * u.im;
    *pu = 0.5 + Bn.x * re - Bn.y * im - Bn.z * w;
    *pd = 0.5 - Bn.x * re + Bn.y * im + Bn.z * w;
}

/* getLocalPopulations (wavefunction.cu:75-97). */
void orc_local_populations(const orc_params *p, const orc_d3 *pos, const orc_c *psiUp,
                           const orc_c *psiDn, const int *atomID, int n, double *pops)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        local_pops(p, pos[a], psiUp[a], psiDn[a], &pops[2 * (size_t)a], &pops[2 * (size_t)a + 1]);
    }
}

/* flipAtoms (wavefunction.cu:99-154): flip probability from the change of the
 * populations, potential step 2 mu |B| (gs muB |B| for mu = gs muB / 2), the
 * speed absorbs it along the old direction (along +z from rest).  Uniform:
 * Philox counter (label, step, 0, (3 << 24) | 0). */
long long orc_flip_atoms(const orc_params *p, const orc_d3 *pos, orc_d3 *vel, const orc_c *psiUp,
                         const orc_c *psiDn, const double *pops, uint8_t *spinUp,
                         const int *atomID, int n, uint32_t step)
{
    long long flips = 0;
#pragma omp parallel for schedule(static) reduction(+ : flips)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 r = pos[a], v = vel[a];
        double opu = pops[2 * (size_t)a], opd = pops[2 * (size_t)a + 1], npu, npd;
        local_pops(p, r, psiUp[a], psiDn[a], &npu, &npd);
        orc_d3 Bn;
        double magB = field_mag_dir(p, r, &Bn);
        double gap = 2.0 * p->mu, Pflip, dEp;
        int up = spinUp[a] != 0;
        if (up) {
            Pflip = (npd - opd) / opu;
            dEp = gap * magB;
        } else {
            Pflip = (npu - opu) / opd;
            dEp = (-1.0 * gap) * magB;
        }
        double Ek = 0.5 * p->mass * (v.x * v.x + v.y * v.y + v.z * v.z);
        uint32_t o[4];
        orc_draw(p->seed, p->stream, (uint32_t)a, step, 0u, ORC_PURPOSE_FLIP, 0u, o);
        if (Pflip > orc_u01(o[0], o[1]) && Ek + dEp > 0.0) {
            spinUp[a] = up ? 0 : 1;
            double magVel = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
            double speed = sqrt(2.0 * (Ek + dEp) / p->mass);
            if (magVel > 0.0) {
                vel[a].x = speed * (v.x / magVel);
                vel[a].y = speed * (v.y / magVel);
                vel[a].z = speed * (v.z / magVel);
            } else {
                vel[a].x = 0.0; vel[a].y = 0.0; vel[a].z = speed;
            }
            flips += 1;
        }
    }
    return flips;
}

/* calculateTau (wavefunction.cu:206-219) with getAcc (:221-237) = the spin-up
 * acceleration of orc_accel. */
double orc_decoherence_tau(const orc_params *p, orc_d3 r)
{
    orc_d3 g = orc_accel(p, r, 1);
    double relAccel = 2.0 * sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
    double tauPos = 1.163 * sqrt(p->sigma / relAccel);
    double tauVel = sqrt(2. / ORC_PI) * ORC_HBAR / (p->mass * p->sigma * relAccel);
    return pow(1 / (tauPos * tauPos * tauPos) + 1 / (tauVel * tauVel * tauVel), -1. / 3.);
}

/* exponentialDecay (wavefunction.cu:175-204). */
void orc_exponential_decay(const orc_params *p, const orc_d3 *pos, orc_c *psiUp, orc_c *psiDn,
                           const int *atomID, int n, double dt)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 Bn;
        (void)field_mag_dir(p, pos[a], &Bn);
        double tau = orc_decoherence_tau(p, pos[a]);
        double om = 1. - exp(-dt / tau), op = 1. + exp(-dt / tau);
        orc_c u = psiUp[a], d = psiDn[a], nu, nd;
        nu.re = 0.5 * ((Bn.x * d.re + Bn.y * d.im + Bn.z * u.re) * om + u.re * op);
        nu.im = 0.5 * ((Bn.x * d.im - Bn.y * d.re + Bn.z * u.im) * om + u.im * op);
        nd.re = 0.5 * ((Bn.x * u.re - Bn.y * u.im - Bn.z * d.re) * om + d.re * op);
        nd.im = 0.5 * ((Bn.x * u.im + Bn.y * u.re - Bn.z * d.im) * om + d.im * op);
        psiUp[a] = nu; psiDn[a] = nd;
    }
}

/* Bloch-vector form of the same rotation: the SU(2) operator
 * cos(theta) - i sin(theta) Bn.sigma of wavefunction.cu:47-70 rotates
 * s = (2 Re u*d, 2 Im u*d, |u|^2 - |d|^2) by phi = 2 theta about Bn
 * (Rodrigues), followed by renormalisation of |s| to 1. */
void orc_rotate_bloch(const orc_params *p, const orc_d3 *pos, orc_d3 *sv,
                      const int *atomID, int n, double dt_spin)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 Bn;
        double B = field_mag_dir(p, pos[a], &Bn);
        double phi = p->gs * ORC_MUB * B * dt_spin / ORC_HBAR;
        double c = cos(phi), sn = sin(phi);
        orc_d3 v = sv[a];
        double nd = Bn.x * v.x + Bn.y * v.y + Bn.z * v.z;
        orc_d3 cr = { Bn.y * v.z - Bn.z * v.y, Bn.z * v.x - Bn.x * v.z, Bn.x * v.y - Bn.y * v.x };
        orc_d3 w = { v.x * c + cr.x * sn + Bn.x * nd * (1.0 - c),
                     v.y * c + cr.y * sn + Bn.y * nd * (1.0 - c),
                     v.z * c + cr.z * sn + Bn.z * nd * (1.0 - c) };
        double L = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
        sv[a].x = w.x / L; sv[a].y = w.y / L; sv[a].z = w.z / L;
    }
}

/* ------------------------------------------------------------------------ */
/* Evaporation and FP64 reductions (src/evaporation.cu:68-189).             */
/* ------------------------------------------------------------------------ */
void orc_evaporation_tag(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                         orc_d3 *evapPos, orc_d3 *evapVel, const orc_c *psiUp,
                         const orc_c *psiDn, const int *atomID, int *evapTag, int n,
                         double e_cut)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        int lost = 0;
        if (psiUp && psiDn) {
            orc_d3 Bn;
            field_mag_dir(p, pos[a], &Bn);
            orc_c u = psiUp[a], d = psiDn[a];
            double proj = 2. * Bn.x * (u.re * d.re + u.im * d.im) +
                          2. * Bn.y * (u.re * d.im - u.im * d.re) +
                          2. * Bn.z * (u.re * u.re + u.im * u.im - 0.5);
            lost = proj < 0.0;
        }
        if (e_cut > 0.0) {
            orc_d3 v = vel[a];
            double E = 0.5 * p->mass * (v.x * v.x + v.y * v.y + v.z * v.z) + orc_potential(p, pos[a]);
            lost |= E > e_cut;
        }
        evapTag[s] = lost;
        if (lost) { evapPos[a] = pos[a]; evapVel[a] = vel[a]; }
    }
}

/* thrust::remove_if(atomID, atomID+N, evapTag, identity) (main.cu:493-498):
 * stable; returns the new length. */
int orc_remove_if(int *atomID, const int *evapTag, int n)
{
    int w = 0;
    for (int s = 0; s < n; ++s)
        if (!evapTag[s]) atomID[w++] = atomID[s];
    return w;
}

/* calculateTemp (evaporation.cu:108-128): T = m <v^2> / (3 kB). */
double orc_calculate_temp(const orc_params *p, const orc_d3 *vel, const int *atomID, int n)
{
    double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
    for (int s = 0; s < n; ++s) {
        orc_d3 v = vel[atomID[s]];
        sum += v.x * v.x + v.y * v.y + v.z * v.z;
    }
    return p->mass / 3. / ORC_KB * (sum / n);
}

void orc_energy(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                const int *atomID, int n, double out[3])
{
    orc_energy_spin(p, pos, vel, NULL, atomID, n, out);
}

/* With spin states: U_down = -mu (|B| + B0) = -U_up - 2 mu B0, so that a flip
 * (potential step 2 mu |B|) conserves E_k + U.  spinUp NULL: all up. */
void orc_energy_spin(const orc_params *p, const orc_d3 *pos, const orc_d3 *vel,
                     const uint8_t *spinUp, const int *atomID, int n, double out[3])
{
    double ek = 0.0, ep = 0.0, v2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ek, ep, v2)
    for (int s = 0; s < n; ++s) {
        int a = atomID[s];
        orc_d3 v = vel[a];
        double w = v.x * v.x + v.y * v.y + v.z * v.z;
        v2 += w;
        ek += 0.5 * p->mass * w;
        double u = orc_potential(p, pos[a]);
        ep += (!spinUp || spinUp[a]) ? u : -u - 2.0 * p->mu * p->B0;
    }
    out[0] = ek; out[1] = ep; out[2] = v2;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions (src/setUp.cu:129-252, 287-292).                       */
/* Draw layout (counter (atom, 0, attempt, (2 << 24) | draw)):              */
/*   draws 0,1: velocities (Box-Muller); 2,3: position radius (quadrupole)  */
/*   or box coordinates; 4..35: quadrupole direction (Marsaglia);           */
/*   IP rejection attempt k uses counter word 2 = k and draws 8,9,10.       */
/* ------------------------------------------------------------------------ */
static orc_d3 sample_position(const orc_params *p, double T, uint32_t atom)
{
    orc_d3 r = { 0.0, 0.0, 0.0 };
    uint32_t o[4];
    if (p->kind == ORC_TRAP_QUAD) {
        /* density exp(-mu B' rho / kT): in (x, y, 2z) the radius is
         * Gamma(3, r_th) and the direction uniform. */
        double rth = ORC_KB * T / (p->mu * p->dBdz);
        orc_draw(p->seed, p->stream, atom, 0, 0, ORC_PURPOSE_INIT, 2, o);
        double u1 = orc_u01(o[0], o[1]), u2 = orc_u01(o[2], o[3]);
        orc_draw(p->seed, p->stream, atom, 0, 0, ORC_PURPOSE_INIT, 3, o);
        double u3 = orc_u01(o[0], o[1]);
        double rad = -rth * (log(u1) + log(u2) + log(u3));
        orc_d3 nh = orc_sphere(p->seed, p->stream, atom, 0, 0, ORC_PURPOSE_INIT, 4);
        r.x = rad * nh.x; r.y = rad * nh.y; r.z = 0.5 * (rad * nh.z);
    } else if (p->kind == ORC_TRAP_IP) {
        /* selectAtomInThermalDistribution (setUp.cu:178-205): Gaussian
         * proposal of width W/3, accepted with exp(-U/kT). */
        for (uint32_t k = 0; k < 100000u; ++k) {
            double n1, n2, n3, spare;
            orc_draw(p->seed, p->stream, atom, 0, k, ORC_PURPOSE_INIT, 8, o);
            orc_box_muller(o, &n1, &n2);
            orc_draw(p->seed, p->stream, atom, 0, k, ORC_PURPOSE_INIT, 9, o);
            orc_box_muller(o, &n3, &spare);
            orc_d3 c = { n1 * p->half_width[0] / 3., n2 * p->half_width[1] / 3., n3 * p->half_width[2] / 3. };
            orc_d3 B = ip_field(p, c);
            double magB = sqrt(B.x * B.x + B.y * B.y + B.z * B.z);
            double U = 0.5 * (magB - p->B0) * p->gs * ORC_MUB;
            double Pr = exp(-U / ORC_KB / T);
            orc_draw(p->seed, p->stream, atom, 0, k, ORC_PURPOSE_INIT, 10, o);
            if (orc_u01(o[0], o[1]) < Pr) { r = c; break; }
        }
    } else {
        orc_draw(p->seed, p->stream, atom, 0, 0, ORC_PURPOSE_INIT, 2, o);
        double ux = orc_u01(o[0], o[1]), uy = orc_u01(o[2], o[3]);
        orc_draw(p->seed, p->stream, atom, 0, 0, ORC_PURPOSE_INIT, 3, o);
        double uz = orc_u01(o[0], o[1]);
        r.x = -p->half_width[0] + (2.0 * p->half_width[0]) * ux;
        r.y = -p->half_width[1] + (2.0 * p->half_width[1]) * uy;
        r.z = -p->half_width[2] + (2.0 * p->half_width[2]) * uz;
    }
    return r;
}

void orc_init_distribution(const orc_params *p, double T, const double Taniso[3],
                           orc_d3 *pos, orc_d3 *vel, orc_d3 *acc, orc_c *psiUp,
                           orc_c *psiDn, int *atomID, uint8_t *spinUp, int n)
{
#pragma omp parallel for schedule(static)
    for (int a = 0; a < n; ++a) {
        uint32_t o[4];
        double z0, z1, z2, spare;
        pos[a] = sample_position(p, T, (uint32_t)a);
        orc_draw(p->seed, p->stream, (uint32_t)a, 0, 0, ORC_PURPOSE_INIT, 0, o);
        orc_box_muller(o, &z0, &z1);
        orc_draw(p->seed, p->stream, (uint32_t)a, 0, 0, ORC_PURPOSE_INIT, 1, o);
        orc_box_muller(o, &z2, &spare);
        double tx = Taniso ? Taniso[0] : T, ty = Taniso ? Taniso[1] : T, tz = Taniso ? Taniso[2] : T;
        vel[a].x = z0 * sqrt(ORC_KB * tx / p->mass);
        vel[a].y = z1 * sqrt(ORC_KB * ty / p->mass);
        vel[a].z = z2 * sqrt(ORC_KB * tz / p->mass);
        if (acc) acc[a] = orc_accel(p, pos[a], 1);
        if (psiUp && psiDn) {
            /* getAlignedSpinUp/Dn (setUp.cu:236-252) */
            orc_d3 Bn;
            field_mag_dir(p, pos[a], &Bn);
            double rs = 1.0 / sqrt(1 + Bn.x);
            psiUp[a].re = 0.5 * (1. + Bn.x + Bn.z) * rs; psiUp[a].im = 0.5 * (-Bn.y) * rs;
            psiDn[a].re = 0.5 * (1. + Bn.x - Bn.z) * rs; psiDn[a].im = 0.5 * Bn.y * rs;
        }
        if (atomID) atomID[a] = a;
        if (spinUp) spinUp[a] = 1;
    }
}

/* initSigvrmax (setUp.cu:287-292): sqrt(16 kT / (pi m)) * sigma, with the
 * cross-section sigma = 8 pi a^2 taken from the parameters. */
double orc_sigvrmax_init(const orc_params *p, double T)
{
    return sqrt(16. * ORC_KB * T / (ORC_PI * p->mass)) * p->sigma;
}

/* ------------------------------------------------------------------------ */
/* The step loop (main.cu:410-524) for runs without spin or evaporation:    */
/* index -> sort -> cell table -> collide, then lpc x (move, temperature).   */
/* ------------------------------------------------------------------------ */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double wall_seconds(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}

void orc_run(const orc_params *p, orc_d3 *pos, orc_d3 *vel, orc_d3 *acc,
             int *atomID, int *cellID, double *sigvrmax, int n, int n_collision_steps,
             uint32_t step0, orc_run_stats *stats)
{
    int ncells = p->nc[0] * p->nc[1] * p->nc[2];
    int *start_end = (int *)malloc(sizeof(int) * 2 * (size_t)(ncells + 1));
    int *count = (int *)malloc(sizeof(int) * (size_t)(ncells + 1));
    int *prefix = (int *)malloc(sizeof(int) * (size_t)(ncells + 1));
    int *ccount = (int *)calloc((size_t)ncells + 1, sizeof(int));
    long long coll = 0, atom_steps = 0;
    double t0 = wall_seconds();
    for (int cs = 0; cs < n_collision_steps; ++cs) {
        orc_index_atoms(p, pos, cellID, atomID, n, 0);
        orc_sort_arrays(cellID, atomID, n, ncells + 1);
        orc_cell_table(cellID, n, ncells, start_end, count, prefix);
        coll += orc_collide(p, vel, sigvrmax, prefix, ccount, atomID, ncells,
                            step0 + (uint32_t)(cs * p->lpc));
        for (int k = 0; k < p->lpc; ++k) {
            orc_move_atoms(p, pos, vel, acc, NULL, atomID, n);
            (void)orc_calculate_temp(p, vel, atomID, n);
            atom_steps += n;
        }
    }
    double t1 = wall_seconds();
    if (stats) {
        stats->seconds = t1 - t0;
        stats->atom_steps = atom_steps;
        stats->collisions = coll;
        stats->threads = orc_num_threads();
    }
    free(start_end); free(count); free(prefix); free(ccount);
}
