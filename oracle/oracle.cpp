/*
 * oracle.cpp — FP64 CPU oracle of the iRBC adhesive-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  No product code
 * links or loads it.
 *
 * Each function names the reference lines it restates.  Expressions keep
 * the reference's operation order (left-to-right evaluation, the same
 * parenthesisation) because the oracle is pinned bit-for-bit against a
 * one-thread build of the reference; that is the only reason the arithmetic
 * reads alike.  Data layout, control structure and the stream plumbing are
 * this file's own: SoA arrays, CSR cell lists, an explicit tape interface.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* MT19937 (Matsumoto & Nishimura 1998) as seeded by the reference's
 * CRandomMersenne: the standard init_genrand recurrence, then 37 outputs
 * discarded (Src/mersenne.cpp:27-42, 73-112). */
struct Mt19937 {
    uint32_t s[624];
    int idx = 624;

    void seed(int32_t sd)
    {
        s[0] = (uint32_t)sd;
        for (int i = 1; i < 624; ++i)
            s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
        idx = 624;
        for (int i = 0; i < 37; ++i) bits();
    }
    void refill()
    {
        static const uint32_t mag[2] = {0u, 0x9908B0DFu};
        int k = 0;
        for (; k < 624 - 397; ++k) {
            uint32_t y = (s[k] & 0x80000000u) | (s[k + 1] & 0x7FFFFFFFu);
            s[k] = s[k + 397] ^ (y >> 1) ^ mag[y & 1u];
        }
        for (; k < 623; ++k) {
            uint32_t y = (s[k] & 0x80000000u) | (s[k + 1] & 0x7FFFFFFFu);
            s[k] = s[k + 397 - 624] ^ (y >> 1) ^ mag[y & 1u];
        }
        uint32_t y = (s[623] & 0x80000000u) | (s[0] & 0x7FFFFFFFu);
        s[623] = s[396] ^ (y >> 1) ^ mag[y & 1u];
        idx = 0;
    }
    uint32_t bits()
    {
        if (idx >= 624) refill();
        uint32_t y = s[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9D2C5680u;
        y ^= (y << 15) & 0xEFC60000u;
        y ^= y >> 18;
        return y;
    }
    /* Random(): 32 bits scaled by 2^-32 (Src/mersenne.cpp:108-112) */
    double uniform() { return (double)bits() * (1.0 / (65536.0 * 65536.0)); }
    /* IRandom(min,max) (Src/mersenne.cpp:115-125) */
    int irandom(int lo, int hi)
    {
        if (hi <= lo) return hi == lo ? lo : (int)0x80000000;
        int r = (int)((double)(uint32_t)(hi - lo + 1) * uniform() + lo);
        if (r > hi) r = hi;
        return r;
    }
};

/* Marsaglia polar method with the cached second variate, as the reference's
 * StochasticLib1::Normal (Src/stoc1.cpp:746-765).  Returns standard normals;
 * the caller applies z*s + m exactly as Normal(m, s) does. */
struct Polar {
    Mt19937 u;
    double x2 = 0.0;
    int have = 0;
    double next()
    {
        if (have) {
            have = 0;
            return x2;
        }
        double x1, w;
        do {
            x1 = 2. * u.uniform() - 1.;
            x2 = 2. * u.uniform() - 1.;
            w = x1 * x1 + x2 * x2;
        } while (w >= 1. || w < 1E-30);
        w = std::sqrt(std::log(w) * (-2. / w));
        x1 *= w;
        x2 *= w;
        have = 1;
        return x1;
    }
};

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
 * Weyl key schedule.  The counter-stream mode below keys it exactly as the
 * device's production streams do, so the oracle can run free over many
 * steps on the same draws: counter {index lo, stream, index hi, sub}, key
 * {seed lo, seed hi}; a uniform is the top 53 bits of a 64-bit half over
 * 2^53; a normal pair is Box-Muller of two uniforms. */
struct Philox {
    static void block(uint32_t c[4], uint32_t k0, uint32_t k1)
    {
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
            c[0] = n0;
            c[1] = n1;
            c[2] = n2;
            c[3] = n3;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    static void u2(uint64_t seed, uint32_t stream, uint64_t index, uint32_t sub, double &u0, double &u1)
    {
        uint32_t c[4] = {(uint32_t)index, stream, (uint32_t)(index >> 32), sub};
        block(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint64_t a = ((uint64_t)c[0] << 32 | c[1]) >> 11, b = ((uint64_t)c[2] << 32 | c[3]) >> 11;
        u0 = (double)a * (1.0 / 9007199254740992.0);
        u1 = (double)b * (1.0 / 9007199254740992.0);
    }
    static void n2(uint64_t seed, uint32_t stream, uint64_t index, uint32_t sub, double &z0, double &z1)
    {
        double u0, u1;
        u2(seed, stream, index, sub, u0, u1);
        const double r = std::sqrt(-2.0 * std::log(1.0 - u0));
        const double t = 6.283185307179586 * u1;
        z0 = r * std::cos(t);
        z1 = r * std::sin(t);
    }
};
/* stream ids of the counter-based draws */
enum { PX_SHIFT = 1, PX_AXIS = 2, PX_WALL = 3, PX_MEM = 4, PX_BOND = 5 };

inline void cross(const double A[3], const double B[3], double C[3])
{
    /* cp() (Src/memfuncs.cpp:47-52) */
    C[0] = (A[1] * B[2]) - (B[1] * A[2]);
    C[1] = (A[2] * B[0]) - (B[2] * A[0]);
    C[2] = (A[0] * B[1]) - (B[0] * A[1]);
}

/* bbound() (Src/solfuncs.cpp:603-622) */
inline double wrap_loop(double x, double lx)
{
    if (x > lx)
        while (x > lx) x -= lx;
    if (x < 0)
        while (x < 0) x += lx;
    return x;
}

} // namespace

struct orc {
    irbc_params p;
    double Lx, Ly, Lz;
    int64_t ncell;

    /* solvent, SoA (define.h:82-94) */
    std::vector<double> x, y, z, vx, vy, vz;
    std::vector<int32_t> pid, indio;

    /* cells (define.h:97-117) */
    std::vector<double> vcm, ncc, rand1, rand2, dvv, the;
    std::vector<int32_t> nc;
    std::vector<int64_t> sol_off;
    std::vector<int64_t> sol_idx;
    std::vector<std::vector<int32_t>> mem;
    std::vector<std::vector<int32_t>> llist;

    /* membrane (define.h:120-179) */
    int nv = 0;
    std::vector<double> mx, my, mz, mvx, mvy, mvz, fx, fy, fz, poff, pon;
    std::vector<int32_t> ix, iy, mpid, bondi, bondl, knob;
    std::vector<std::vector<std::pair<double, int>>> pairs;
    std::vector<int32_t> prep_live; /* took the free-knob branch in the last bond_prep */
    std::vector<int32_t> e1, e2, es1, es2, iedge;
    std::vector<double> l0, kappap;
    std::vector<int32_t> t1, t2, t3;
    std::vector<double> tarea, cmx, cmy, cmz;
    double Sarea = 0, Volume = 0, Sarea0 = 0, Volume0 = 0, theta0 = 0;
    double center[3] = {0, 0, 0};
    int64_t nsolb = 0;

    /* parasite */
    double para[3] = {0, 0, 0}, paraV[3] = {0, 0, 0}, paraft[3] = {0, 0, 0}, parafdt[3] = {0, 0, 0};
    int32_t paradx = 0, parady = 0, paraid = 0;

    /* ligands (maincode.cpp:77-93) */
    int64_t nlig = 0;
    int32_t nligy = 0;
    std::vector<double> xlig, ylig;
    std::vector<int32_t> ligi;

    /* streams */
    Mt19937 ran;
    Polar sto;
    int tapes = 0;
    /* counter-based mode: every draw from Philox keyed by (seed, step,
     * cell or vertex), as the device's production streams */
    int philox = 0;
    uint64_t pseed = 0;
    /* FP32 storage mode: solvent positions and velocities rounded to float
     * where the device stores them (upload, the rotated velocities before
     * streaming, the end of each streaming step) */
    int fp32 = 0;
    void round_solvent()
    {
        if (!fp32) return;
        const int64_t n = (int64_t)x.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] = (double)(float)x[i];
            y[i] = (double)(float)y[i];
            z[i] = (double)(float)z[i];
            vx[i] = (double)(float)vx[i];
            vy[i] = (double)(float)vy[i];
            vz[i] = (double)(float)vz[i];
        }
    }
    double shift_u[3] = {0, 0, 0};
    std::vector<double> cell_u, wall_z, mem_z, bond_u;
    int64_t bond_pos = 0;

    /* driver counters (maincode.cpp:450-460) */
    int64_t step = 0;
    int32_t colsteps = -1, upd = 0;
    double wallspeed = 0.0;

    /* cindex[][][] lookup (maincode.cpp:112-136, define.h:330-334): in-range
     * indices map to i*Ly*(Lz+1)+j*(Lz+1)+k; the array's extra slab at
     * i = Lx, j = Ly or k = Lz+1 is never written and reads as cell 0. */
    int32_t cell_of(int fxi, int fyi, int fzi) const
    {
        if (fxi >= 0 && fxi < p.Lx && fyi >= 0 && fyi < p.Ly && fzi >= 0 && fzi <= p.Lz)
            return (int32_t)(((int64_t)fxi * p.Ly + fyi) * (p.Lz + 1) + fzi);
        return 0;
    }
    void cell_ijk(int64_t c, int &i, int &j, int &k) const
    {
        k = (int)(c % (p.Lz + 1));
        j = (int)((c / (p.Lz + 1)) % p.Ly);
        i = (int)(c / ((int64_t)(p.Lz + 1) * p.Ly));
    }
    /* draws */
    double d_shift(int k)
    {
        if (tapes) return shift_u[k];
        if (philox) {
            double u0, u1, u2, u3;
            Philox::u2(pseed, PX_SHIFT, (uint64_t)step, 0u, u0, u1);
            Philox::u2(pseed, PX_SHIFT, (uint64_t)step, 1u, u2, u3);
            return shift_u[k] = (k == 0 ? u0 : k == 1 ? u1 : u2);
        }
        double u = ran.uniform();
        shift_u[k] = u;
        return u;
    }
    double d_cell(int64_t c, int k)
    {
        if (tapes) return cell_u[2 * c + k];
        if (philox) {
            double u0, u1;
            Philox::u2(pseed, PX_AXIS, (uint64_t)step, (uint32_t)c, u0, u1);
            return cell_u[2 * c + k] = (k ? u1 : u0);
        }
        double u = ran.uniform();
        cell_u[2 * c + k] = u;
        return u;
    }
    double px_normal(uint32_t stream, int64_t c, int k) const
    {
        double a, b;
        Philox::n2(pseed, stream, (uint64_t)step, 2u * (uint32_t)c + (uint32_t)(k >> 1), a, b);
        return (k & 1) ? b : a;
    }
    double d_wall(int64_t c, int k)
    {
        if (tapes) return wall_z[3 * c + k];
        if (philox) return wall_z[3 * c + k] = px_normal(PX_WALL, c, k);
        double zz = sto.next();
        wall_z[3 * c + k] = zz;
        return zz;
    }
    double d_mem(int64_t c, int k)
    {
        if (tapes) return mem_z[3 * c + k];
        if (philox) return mem_z[3 * c + k] = px_normal(PX_MEM, c, k);
        double zz = sto.next();
        mem_z[3 * c + k] = zz;
        return zz;
    }
    double d_bond()
    {
        if (tapes) return bond_pos < (int64_t)bond_u.size() ? bond_u[bond_pos++] : 0.0;
        double u = ran.uniform();
        bond_u.push_back(u);
        return u;
    }
    /* the k-th draw of vertex v in this step's bond dynamics: keyed by
     * (v, k) in the counter-based mode, the next draw of the sequence else */
    double d_bond_at(int v, int k)
    {
        if (!philox || tapes) return d_bond();
        double u0, u1;
        Philox::u2(pseed, PX_BOND, (uint64_t)step, ((uint32_t)v << 4) | (uint32_t)(k >> 1), u0, u1);
        const double u = (k & 1) ? u1 : u0;
        bond_u.push_back(u);
        return u;
    }

    /* unwrapped vertex coordinates, x + ix*Lx (Src/memfuncs.cpp passim) */
    double ux(int v) const { return mx[v] + ix[v] * Lx; }
    double uy(int v) const { return my[v] + iy[v] * Ly; }

    double areatri(int a1, int a2, int a3) const
    {
        /* Areatri (Src/memfuncs.cpp:144-157) */
        double A[3], B[3], eta[3];
        A[0] = ux(a2) - ux(a1);
        A[1] = uy(a2) - uy(a1);
        A[2] = mz[a2] - mz[a1];
        B[0] = ux(a3) - ux(a1);
        B[1] = uy(a3) - uy(a1);
        B[2] = mz[a3] - mz[a1];
        cross(A, B, eta);
        return 0.5 * std::sqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
    }
    double voltri(int a1, int a2, int a3) const
    {
        /* Voltri + Normaldt (Src/memfuncs.cpp:110-142) */
        double A[3] = {ux(a1), uy(a1), mz[a1]};
        double B[3] = {ux(a2), uy(a2), mz[a2]};
        double C[3] = {ux(a3), uy(a3), mz[a3]};
        double Cn[3], D[3], E[3], F[3];
        Cn[0] = (A[0] + B[0] + C[0]) / 3.0;
        Cn[1] = (A[1] + B[1] + C[1]) / 3.0;
        Cn[2] = (A[2] + B[2] + C[2]) / 3.0;
        normaldt(a1, a2, a3, D);
        (void)E;
        (void)F;
        return (1.0 / 6.0) * ((D[0] * Cn[0]) + (D[1] * Cn[1]) + (D[2] * Cn[2]));
    }
    void normaldt(int a1, int a2, int a3, double X[3]) const
    {
        double A[3], B[3];
        A[0] = ux(a2) - ux(a1);
        A[1] = uy(a2) - uy(a1);
        A[2] = mz[a2] - mz[a1];
        B[0] = ux(a3) - ux(a1);
        B[1] = uy(a3) - uy(a1);
        B[2] = mz[a3] - mz[a1];
        cross(A, B, X);
    }
    /* liglist membership of ligand l for cell c (maincode.cpp:138-154):
     * only cells with zmax <= 2; fabs() before the periodic wrap. */
    bool lig_member(int64_t c, int64_t l) const
    {
        int i, j, k;
        cell_ijk(c, i, j, k);
        double zmax = -1.0 + (double)k + 1.0;
        if (!(zmax <= 2.0)) return false;
        double xmn = (double)i, xmxx = (double)i + 1.0;
        double ymn = (double)j, ymxx = (double)j + 1.0;
        double A0 = std::fabs(((xmn + xmxx) * 0.5) - xlig[l]);
        double A1 = std::fabs(((ymn + ymxx) * 0.5) - ylig[l]);
        A0 -= Lx * std::round(A0 / Lx);
        A1 -= Ly * std::round(A1 / Ly);
        return (A0 < 1.5) && (A1 < 1.5);
    }
};

static int g_threads = 0;

extern "C" {

void orc_set_threads(int n)
{
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

orc *orc_create(const irbc_params *p)
{
    orc *o = new orc();
    o->p = *p;
    o->Lx = p->Lx;
    o->Ly = p->Ly;
    o->Lz = p->Lz;
    o->ncell = irbc_ncells(p);
    int64_t n = o->ncell;
    o->vcm.assign(3 * n, 0.0);
    o->ncc.assign(n, 0.0);
    o->rand1.assign(n, 0.0);
    o->rand2.assign(n, 0.0);
    o->dvv.assign(n, 0.0);
    o->the.assign(n, 0.0);
    o->nc.assign(n, 0);
    o->sol_off.assign(n + 1, 0);
    o->mem.assign(n, {});
    o->llist.assign(n, {});
    o->cell_u.assign(2 * n, 0.0);
    o->wall_z.assign(3 * n, 0.0);
    o->mem_z.assign(3 * n, 0.0);
    /* ligand grid (maincode.cpp:77-93): x outer, y inner, spacing 0.5 */
    double dx = p->lig_spacing, xm = 0.0;
    while (xm < o->Lx) {
        double ym = 0.0;
        int32_t ny = 0;
        while (ym < o->Ly) {
            o->xlig.push_back(xm);
            o->ylig.push_back(ym);
            o->ligi.push_back(0);
            ym += dx;
            ++ny;
        }
        o->nligy = ny;
        xm += dx;
    }
    o->nlig = (int64_t)o->xlig.size();
    o->ran.seed(12345);
    o->sto.u.seed(12345);
    return o;
}

void orc_destroy(orc *o) { delete o; }

int64_t orc_ncells(const orc *o) { return o->ncell; }

/* ---------------- streams ---------------- */

void orc_seed_mt(orc *o, int32_t seed_ran, int32_t seed_sto)
{
    o->ran.seed(seed_ran);
    o->sto.u.seed(seed_sto);
    o->sto.have = 0;
    o->tapes = 0;
}

void orc_use_tapes(orc *o, int on) { o->tapes = on; }

void orc_use_philox(orc *o, uint64_t seed)
{
    o->philox = 1;
    o->pseed = seed;
    o->tapes = 0;
}

void orc_philox_u2(uint64_t seed, uint32_t stream, uint64_t index, uint32_t sub, double *u)
{
    Philox::u2(seed, stream, index, sub, u[0], u[1]);
}

void orc_philox_block(uint32_t *c, uint32_t k0, uint32_t k1) { Philox::block(c, k0, k1); }

void orc_set_fp32(orc *o, int on)
{
    o->fp32 = on ? 1 : 0;
    o->round_solvent();
}

void orc_set_collision_tape(orc *o, const double *shift_u, const double *cell_u,
                            const double *wall_z, const double *mem_z)
{
    std::memcpy(o->shift_u, shift_u, 3 * sizeof(double));
    std::memcpy(o->cell_u.data(), cell_u, 2 * o->ncell * sizeof(double));
    std::memcpy(o->wall_z.data(), wall_z, 3 * o->ncell * sizeof(double));
    std::memcpy(o->mem_z.data(), mem_z, 3 * o->ncell * sizeof(double));
}

void orc_get_collision_tape(const orc *o, double *shift_u, double *cell_u, double *wall_z,
                            double *mem_z)
{
    std::memcpy(shift_u, o->shift_u, 3 * sizeof(double));
    std::memcpy(cell_u, o->cell_u.data(), 2 * o->ncell * sizeof(double));
    std::memcpy(wall_z, o->wall_z.data(), 3 * o->ncell * sizeof(double));
    std::memcpy(mem_z, o->mem_z.data(), 3 * o->ncell * sizeof(double));
}

void orc_set_bond_tape(orc *o, const double *u, int64_t n)
{
    o->bond_u.assign(u, u + n);
    o->bond_pos = 0;
}

int64_t orc_get_bond_tape(const orc *o, double *u, int64_t cap)
{
    int64_t n = (int64_t)o->bond_u.size();
    if (u) std::memcpy(u, o->bond_u.data(), (size_t)std::min(n, cap) * sizeof(double));
    return n;
}

double orc_mt_uniform(orc *o) { return o->ran.uniform(); }
uint32_t orc_mt_bits(orc *o) { return o->ran.bits(); }
double orc_sto_normal(orc *o, double m, double s) { return o->sto.next() * s + m; }

/* ---------------- state ---------------- */

int orc_set_mesh(orc *o, int nv, const double *xyz, int ne, const int32_t *edges, int nt,
                 const int32_t *tris)
{
    /* topology and rest state of the mesh as maincode.cpp:157-375 builds it,
     * for vertex positions that are already placed in the box */
    o->nv = nv;
    auto zero = [&](std::vector<double> &v) { v.assign(nv, 0.0); };
    auto izero = [&](std::vector<int32_t> &v) { v.assign(nv, 0); };
    zero(o->mx); zero(o->my); zero(o->mz);
    zero(o->mvx); zero(o->mvy); zero(o->mvz);
    zero(o->fx); zero(o->fy); zero(o->fz);
    zero(o->poff); zero(o->pon);
    izero(o->ix); izero(o->iy); izero(o->mpid); izero(o->bondi); izero(o->knob);
    o->bondl.assign(nv, 12345678); /* randno (define.h:37) */
    o->pairs.assign(nv, {});
    o->prep_live.assign(nv, 0);
    for (int i = 0; i < nv; ++i) {
        o->mx[i] = xyz[3 * i];
        o->my[i] = xyz[3 * i + 1];
        o->mz[i] = xyz[3 * i + 2];
    }
    o->e1.resize(ne); o->e2.resize(ne); o->l0.resize(ne); o->kappap.resize(ne);
    for (int i = 0; i < ne; ++i) {
        o->e1[i] = edges[2 * i];
        o->e2[i] = edges[2 * i + 1];
        int a1 = o->e1[i], a2 = o->e2[i];
        double distx = (o->mx[a1] - o->mx[a2]) * (o->mx[a1] - o->mx[a2]);
        double disty = (o->my[a1] - o->my[a2]) * (o->my[a1] - o->my[a2]);
        double distz = (o->mz[a1] - o->mz[a2]) * (o->mz[a1] - o->mz[a2]);
        o->l0[i] = std::sqrt(distx + disty + distz); /* maincode.cpp:247-250 */
    }
    o->t1.resize(nt); o->t2.resize(nt); o->t3.resize(nt);
    o->cmx.assign(nt, 0.0); o->cmy.assign(nt, 0.0); o->cmz.assign(nt, 0.0);
    for (int i = 0; i < nt; ++i) {
        o->t1[i] = tris[3 * i];
        o->t2[i] = tris[3 * i + 1];
        o->t3[i] = tris[3 * i + 2];
    }
    /* edgest: the vertex opposite the edge in the first and in the last
     * triangle that holds it (maincode.cpp:276-342) */
    o->es1.assign(ne, 0);
    o->es2.assign(ne, 0);
    for (int i = 0; i < ne; ++i) {
        int a1 = o->e1[i], a2 = o->e2[i];
        auto holds = [&](int j) {
            int f1 = (a1 == o->t1[j]) + (a1 == o->t2[j]) + (a1 == o->t3[j]);
            int f2 = (a2 == o->t1[j]) + (a2 == o->t2[j]) + (a2 == o->t3[j]);
            return f1 > 0 && f2 > 0;
        };
        auto third = [&](int j, int32_t &dst) {
            if (o->t1[j] != a1 && o->t1[j] != a2) dst = o->t1[j];
            if (o->t2[j] != a1 && o->t2[j] != a2) dst = o->t2[j];
            if (o->t3[j] != a1 && o->t3[j] != a2) dst = o->t3[j];
        };
        for (int j = 0; j < nt; ++j)
            if (holds(j)) { third(j, o->es1[i]); break; }
        for (int j = nt - 1; j >= 0; --j)
            if (holds(j)) { third(j, o->es2[i]); break; }
    }
    orc_totalva0(o);
    /* kappap per edge (maincode.cpp:347-361) */
    double elap = o->p.elap, shear = o->p.shear;
    for (int i = 0; i < ne; ++i) {
        double ll0 = o->l0[i];
        double lmax = 2.5 * ll0;
        double x0 = ll0 / lmax;
        double onemx0 = 1.0 - x0;
        double shear1 = 1.0 / (4.0 * onemx0 * onemx0);
        double shear2 = x0 / (2.0 * onemx0 * onemx0 * onemx0);
        double shear4 = (1.732 / (4.0 * elap * lmax * x0));
        double sh = shear4 * shear2 - shear4 * shear1 + shear4 * 0.25;
        o->kappap[i] = (shear - sh) * 4.0 * ll0 * ll0 * ll0 / (3.0 * 1.732);
    }
    /* tarea (maincode.cpp:371-375) */
    o->tarea.resize(nt);
    for (int i = 0; i < nt; ++i) o->tarea[i] = o->areatri(o->t1[i], o->t2[i], o->t3[i]);
    /* theta0 from the vertex count (maincode.cpp:207-209) */
    double Pi = o->p.Pi;
    double numer = 1.73205080757 * ((double)nv - 2.0) - 5.0 * Pi;
    double denom = 1.73205080757 * ((double)nv - 2.0) - 3.0 * Pi;
    o->theta0 = 180.0 * std::acos(numer / denom) / Pi;
    o->iedge.assign(ne, 1);
    orc_centerofmass(o);
    orc_makeiedge(o);
    return 0;
}

void orc_set_solvent(orc *o, int64_t n, const double *x, const double *y, const double *z,
                     const double *vx, const double *vy, const double *vz, const int32_t *indio)
{
    o->x.assign(x, x + n);
    o->y.assign(y, y + n);
    o->z.assign(z, z + n);
    o->vx.assign(vx, vx + n);
    o->vy.assign(vy, vy + n);
    o->vz.assign(vz, vz + n);
    o->pid.assign(n, 0);
    if (indio)
        o->indio.assign(indio, indio + n);
    else
        o->indio.assign(n, 1);
    o->round_solvent();
}

int64_t orc_n_solvent(const orc *o) { return (int64_t)o->x.size(); }

void orc_get_solvent(const orc *o, double *x, double *y, double *z, double *vx, double *vy,
                     double *vz, int32_t *pid, int32_t *indio)
{
    size_t n = o->x.size();
    if (x) std::memcpy(x, o->x.data(), n * 8);
    if (y) std::memcpy(y, o->y.data(), n * 8);
    if (z) std::memcpy(z, o->z.data(), n * 8);
    if (vx) std::memcpy(vx, o->vx.data(), n * 8);
    if (vy) std::memcpy(vy, o->vy.data(), n * 8);
    if (vz) std::memcpy(vz, o->vz.data(), n * 8);
    if (pid) std::memcpy(pid, o->pid.data(), n * 4);
    if (indio) std::memcpy(indio, o->indio.data(), n * 4);
}

int orc_n_vertices(const orc *o) { return o->nv; }

void orc_set_vertices(orc *o, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *fx,
                      const double *fy, const double *fz, const int32_t *ix, const int32_t *iy,
                      const int32_t *pid, const int32_t *bondi, const int32_t *bondl,
                      const int32_t *knob)
{
    int n = o->nv;
    auto cp = [n](std::vector<double> &d, const double *s) { if (s) d.assign(s, s + n); };
    auto ci = [n](std::vector<int32_t> &d, const int32_t *s) { if (s) d.assign(s, s + n); };
    cp(o->mx, x); cp(o->my, y); cp(o->mz, z);
    cp(o->mvx, vx); cp(o->mvy, vy); cp(o->mvz, vz);
    cp(o->fx, fx); cp(o->fy, fy); cp(o->fz, fz);
    ci(o->ix, ix); ci(o->iy, iy); ci(o->mpid, pid);
    ci(o->bondi, bondi); ci(o->bondl, bondl); ci(o->knob, knob);
}

void orc_get_vertices(const orc *o, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *fx, double *fy, double *fz, double *poff, double *pon,
                      int32_t *ix, int32_t *iy, int32_t *pid, int32_t *bondi, int32_t *bondl,
                      int32_t *knob)
{
    size_t n = o->nv;
    auto cp = [n](double *d, const std::vector<double> &s) { if (d) std::memcpy(d, s.data(), n * 8); };
    auto ci = [n](int32_t *d, const std::vector<int32_t> &s) { if (d) std::memcpy(d, s.data(), n * 4); };
    cp(x, o->mx); cp(y, o->my); cp(z, o->mz);
    cp(vx, o->mvx); cp(vy, o->mvy); cp(vz, o->mvz);
    cp(fx, o->fx); cp(fy, o->fy); cp(fz, o->fz);
    cp(poff, o->poff); cp(pon, o->pon);
    ci(ix, o->ix); ci(iy, o->iy); ci(pid, o->mpid);
    ci(bondi, o->bondi); ci(bondl, o->bondl); ci(knob, o->knob);
}

void orc_set_parasite(orc *o, const double *d, const int32_t *i)
{
    for (int k = 0; k < 3; ++k) {
        o->para[k] = d[k];
        o->paraV[k] = d[3 + k];
        o->paraft[k] = d[6 + k];
        o->parafdt[k] = d[9 + k];
    }
    o->paradx = i[0];
    o->parady = i[1];
    o->paraid = i[2];
}

void orc_get_parasite(const orc *o, double *d, int32_t *i)
{
    for (int k = 0; k < 3; ++k) {
        d[k] = o->para[k];
        d[3 + k] = o->paraV[k];
        d[6 + k] = o->paraft[k];
        d[9 + k] = o->parafdt[k];
    }
    i[0] = o->paradx;
    i[1] = o->parady;
    i[2] = o->paraid;
}

int64_t orc_n_ligands(const orc *o) { return o->nlig; }
void orc_set_ligands(orc *o, const int32_t *l) { o->ligi.assign(l, l + o->nlig); }
void orc_get_ligands(const orc *o, int32_t *l) { std::memcpy(l, o->ligi.data(), o->nlig * 4); }
void orc_get_ligand_xy(const orc *o, double *x, double *y)
{
    std::memcpy(x, o->xlig.data(), o->nlig * 8);
    std::memcpy(y, o->ylig.data(), o->nlig * 8);
}

void orc_get_cells(const orc *o, double *vcm, int32_t *nc, double *ncc, double *rand1,
                   double *rand2, double *dvv, double *the)
{
    size_t n = o->ncell;
    if (vcm) std::memcpy(vcm, o->vcm.data(), 3 * n * 8);
    if (nc) std::memcpy(nc, o->nc.data(), n * 4);
    if (ncc) std::memcpy(ncc, o->ncc.data(), n * 8);
    if (rand1) std::memcpy(rand1, o->rand1.data(), n * 8);
    if (rand2) std::memcpy(rand2, o->rand2.data(), n * 8);
    if (dvv) std::memcpy(dvv, o->dvv.data(), n * 8);
    if (the) std::memcpy(the, o->the.data(), n * 8);
}

void orc_get_globals(const orc *o, double *g)
{
    g[0] = o->Sarea;
    g[1] = o->Volume;
    g[2] = o->Sarea0;
    g[3] = o->Volume0;
    g[4] = o->theta0;
    g[5] = o->wallspeed;
    g[6] = o->center[0];
    g[7] = o->center[1];
    g[8] = o->center[2];
    g[9] = (double)o->nsolb;
}

void orc_get_topology(const orc *o, double *l0, double *kappap, int32_t *edgest, int32_t *iedge,
                      double *tarea)
{
    size_t ne = o->e1.size(), nt = o->t1.size();
    if (l0) std::memcpy(l0, o->l0.data(), ne * 8);
    if (kappap) std::memcpy(kappap, o->kappap.data(), ne * 8);
    if (edgest)
        for (size_t i = 0; i < ne; ++i) {
            edgest[2 * i] = o->es1[i];
            edgest[2 * i + 1] = o->es2[i];
        }
    if (iedge) std::memcpy(iedge, o->iedge.data(), ne * 4);
    if (tarea) std::memcpy(tarea, o->tarea.data(), nt * 8);
}

int64_t orc_llist_total(const orc *o)
{
    int64_t t = 0;
    for (auto &l : o->llist) t += (int64_t)l.size();
    return t;
}

void orc_get_llist(const orc *o, int64_t *off, int32_t *ent)
{
    int64_t t = 0;
    for (int64_t c = 0; c < o->ncell; ++c) {
        off[c] = t;
        for (int32_t v : o->llist[c]) ent[t++] = v;
    }
    off[o->ncell] = t;
}

int64_t orc_get_pairs(const orc *o, int32_t *cnt, int32_t *lig, int64_t cap)
{
    int64_t t = 0;
    for (int v = 0; v < o->nv; ++v) {
        int c = o->prep_live[v] ? (int)o->pairs[v].size() : 0;
        if (cnt) cnt[v] = c;
        for (int k = 0; k < c; ++k) {
            if (lig && t < cap) lig[t] = o->pairs[v][k].second;
            ++t;
        }
    }
    return t;
}

/* ---------------- collision ---------------- */

static void rotate_one(const orc *o, int64_t c, double &vx, double &vy, double &vz)
{
    /* CollisionDyn body (Src/solfuncs.cpp:707-733) */
    double thet = o->rand1[c];
    double zz = o->rand2[c];
    double unitx = std::sqrt(1.0 - (zz * zz)) * std::cos(thet);
    double unity = std::sqrt(1.0 - (zz * zz)) * std::sin(thet);
    double unitz = zz;
    double scalev = o->the[c];
    const double *vc = &o->vcm[3 * c];
    double d[3] = {vx - vc[0], vy - vc[1], vz - vc[2]};
    double u[3] = {unitx, unity, unitz};
    double xxx = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
    double dvp[3] = {xxx * u[0], xxx * u[1], xxx * u[2]};
    double dvn[3] = {d[0] - dvp[0], d[1] - dvp[1], d[2] - dvp[2]};
    double dvpn[3];
    cross(dvn, u, dvpn);
    double ca = o->p.cosalpha, sa = o->p.sinalpha;
    vx = vc[0] + scalev * (ca * dvn[0] + sa * dvpn[0] + dvp[0]);
    vy = vc[1] + scalev * (ca * dvn[1] + sa * dvpn[1] + dvp[1]);
    vz = vc[2] + scalev * (ca * dvn[2] + sa * dvpn[2] + dvp[2]);
}

void orc_collision(orc *o)
{
    /* CollisionStep (Src/solfuncs.cpp:109-323) */
    const irbc_params &P = o->p;
    const double a = 1.0;
    double rpos[3];
    rpos[0] = a * o->d_shift(0);
    rpos[1] = a * o->d_shift(1);
    rpos[2] = a * o->d_shift(2);
    int64_t nc = o->ncell;
    if (!o->tapes) {
        std::fill(o->wall_z.begin(), o->wall_z.end(), 0.0);
        std::fill(o->mem_z.begin(), o->mem_z.end(), 0.0);
    }
    /* reset (:120-136) */
    std::fill(o->vcm.begin(), o->vcm.end(), 0.0);
    std::fill(o->nc.begin(), o->nc.end(), 0);
    std::fill(o->ncc.begin(), o->ncc.end(), 0.0);
    std::fill(o->dvv.begin(), o->dvv.end(), 0.0);
    std::fill(o->the.begin(), o->the.end(), 0.0);
    for (auto &m : o->mem) m.clear();
    /* shifted-grid keys (:139-162); lists keep particle order (stable CSR) */
    int64_t ns = (int64_t)o->x.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ns; ++i) {
        int fxi = (int)std::floor(wrap_loop(o->x[i] - rpos[0], o->Lx));
        int fyi = (int)std::floor(wrap_loop(o->y[i] - rpos[1], o->Ly));
        int fzi = (int)std::floor(o->z[i] - rpos[2] + 1.0);
        o->pid[i] = o->cell_of(fxi, fyi, fzi);
    }
    std::fill(o->sol_off.begin(), o->sol_off.end(), 0);
    for (int64_t i = 0; i < ns; ++i) o->sol_off[o->pid[i] + 1]++;
    for (int64_t c = 0; c < nc; ++c) o->sol_off[c + 1] += o->sol_off[c];
    o->sol_idx.assign(ns, 0);
    {
        std::vector<int64_t> cur(o->sol_off.begin(), o->sol_off.end() - 1);
        for (int64_t i = 0; i < ns; ++i) o->sol_idx[cur[o->pid[i]]++] = i;
    }
    for (int i = 0; i < o->nv; ++i) {
        int fxi = (int)std::floor(wrap_loop(o->mx[i] - rpos[0], o->Lx));
        int fyi = (int)std::floor(wrap_loop(o->my[i] - rpos[1], o->Ly));
        int fzi = (int)std::floor(o->mz[i] - rpos[2] + 1.0);
        o->mpid[i] = o->cell_of(fxi, fyi, fzi);
        o->mem[o->mpid[i]].push_back(i);
    }
    {
        int fxi = (int)std::floor(wrap_loop(o->para[0] - rpos[0], o->Lx));
        int fyi = (int)std::floor(wrap_loop(o->para[1] - rpos[1], o->Ly));
        int fzi = (int)std::floor(o->para[2] - rpos[2] + 1.0);
        o->paraid = o->cell_of(fxi, fyi, fzi);
    }
    const double mass = P.mass, pmass = P.pmass, PRmass = P.PRmass, navg = P.navg, Th = P.Theta;
    /* solvent sums (:164-176) */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        double *v = &o->vcm[3 * c];
        for (int64_t q = o->sol_off[c]; q < o->sol_off[c + 1]; ++q) {
            int64_t i = o->sol_idx[q];
            v[0] += mass * o->vx[i];
            v[1] += mass * o->vy[i];
            v[2] += mass * o->vz[i];
            o->ncc[c] += (1 * mass);
            o->nc[c] += 1;
        }
    }
    /* wall ghosts (:178-206): top layer zmax >= Lz, bottom layers zmin <= 0 */
    for (int64_t c = 0; c < nc; ++c) {
        int i, j, k;
        o->cell_ijk(c, i, j, k);
        double zmin = -1.0 + (double)k, zmax = -1.0 + (double)k + 1.0;
        int nn = o->nc[c];
        double *v = &o->vcm[3 * c];
        if (zmax >= o->Lz) {
            if (nn < navg) {
                double s = std::sqrt((navg - nn) * Th / mass);
                v[0] += mass * (o->d_wall(c, 0) * s + (navg - nn) * o->wallspeed);
                v[1] += mass * (o->d_wall(c, 1) * s + 0);
                v[2] += mass * (o->d_wall(c, 2) * s + 0);
                o->ncc[c] = navg * mass;
            }
        } else if (zmin <= 0) {
            if (nn < navg) {
                double s = std::sqrt((navg - nn) * Th / mass);
                v[0] += mass * (o->d_wall(c, 0) * s + 0);
                v[1] += mass * (o->d_wall(c, 1) * s + 0);
                v[2] += mass * (o->d_wall(c, 2) * s + 0);
                o->ncc[c] = navg * mass;
            }
        }
    }
    /* membrane sums (:208-220), parasite (:223-228) */
    for (int64_t c = 0; c < nc; ++c) {
        double *v = &o->vcm[3 * c];
        for (int32_t q : o->mem[c]) {
            v[0] += pmass * o->mvx[q];
            v[1] += pmass * o->mvy[q];
            v[2] += pmass * o->mvz[q];
            o->ncc[c] += (1 * pmass);
            o->nc[c] += 1;
        }
    }
    {
        int64_t c = o->paraid;
        o->vcm[3 * c] += PRmass * o->paraV[0];
        o->vcm[3 * c + 1] += PRmass * o->paraV[1];
        o->vcm[3 * c + 2] += PRmass * o->paraV[2];
        o->ncc[c] += (1 * PRmass);
        o->nc[c] += 1;
    }
    /* membrane-cell ghosts (:230-245) */
    for (int64_t c = 0; c < nc; ++c) {
        int nn = o->nc[c];
        if ((nn < navg) && (o->mem[c].size() > 0)) {
            double s = std::sqrt((navg - nn) * Th / mass);
            double *v = &o->vcm[3 * c];
            v[0] += mass * (o->d_mem(c, 0) * s + 0);
            v[1] += mass * (o->d_mem(c, 1) * s + 0);
            v[2] += mass * (o->d_mem(c, 2) * s + 0);
            o->ncc[c] += (navg - nn) * mass;
        }
    }
    /* vcm /= (int)ncc (:248-261) */
    for (int64_t c = 0; c < nc; ++c) {
        int nn = (int)o->ncc[c];
        if (nn != 0) {
            o->vcm[3 * c] /= (double)nn;
            o->vcm[3 * c + 1] /= (double)nn;
            o->vcm[3 * c + 2] /= (double)nn;
        }
    }
    /* random axes (:262-271) */
    for (int64_t c = 0; c < nc; ++c) {
        o->rand1[c] = 2.0 * P.Pi * o->d_cell(c, 0);
        o->rand2[c] = -1.0 + (2.0 * o->d_cell(c, 1));
    }
    /* dvv (:273-307) */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        const double *v = &o->vcm[3 * c];
        for (int64_t q = o->sol_off[c]; q < o->sol_off[c + 1]; ++q) {
            int64_t i = o->sol_idx[q];
            double A0 = o->vx[i] - v[0], A1 = o->vy[i] - v[1], A2 = o->vz[i] - v[2];
            double aa = (A0 * A0 + A1 * A1 + A2 * A2);
            o->dvv[c] += (mass * aa);
        }
    }
    for (int64_t c = 0; c < nc; ++c) {
        const double *v = &o->vcm[3 * c];
        for (int32_t q : o->mem[c]) {
            double A0 = o->mvx[q] - v[0], A1 = o->mvy[q] - v[1], A2 = o->mvz[q] - v[2];
            double aa = (A0 * A0 + A1 * A1 + A2 * A2);
            o->dvv[c] += (pmass * aa);
        }
    }
    {
        int64_t c = o->paraid;
        const double *v = &o->vcm[3 * c];
        double A0 = o->paraV[0] - v[0], A1 = o->paraV[1] - v[1], A2 = o->paraV[2] - v[2];
        double aa = (A0 * A0 + A1 * A1 + A2 * A2);
        o->dvv[c] += (PRmass * aa);
    }
    /* thermostat factor (:310-318) */
    for (int64_t c = 0; c < nc; ++c) {
        o->the[c] = 1.0;
        if ((o->nc[c] > 0) && (o->dvv[c] != 0))
            o->the[c] = std::sqrt(3.0 * ((double)o->nc[c] - 1.0) * Th / (o->dvv[c]));
    }
    /* rotations (:700-810) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ns; ++i) {
        rotate_one(o, o->pid[i], o->vx[i], o->vy[i], o->vz[i]);
        if (o->fp32) {
            o->vx[i] = (double)(float)o->vx[i];
            o->vy[i] = (double)(float)o->vy[i];
            o->vz[i] = (double)(float)o->vz[i];
        }
    }
    for (int i = 0; i < o->nv; ++i) rotate_one(o, o->mpid[i], o->mvx[i], o->mvy[i], o->mvz[i]);
    rotate_one(o, o->paraid, o->paraV[0], o->paraV[1], o->paraV[2]);
}

/* ---------------- cell -> triangle lists ---------------- */

void orc_llistupdate(orc *o)
{
    /* llistupdate (Src/solfuncs.cpp:326-365); the centroid terms do not
     * depend on the cell and are evaluated once per triangle with the same
     * arithmetic.  A list keeps its stale tail beyond the new count. */
    int nt = (int)o->t1.size();
    std::vector<double> cx(nt), cy(nt), cz(nt);
    for (int j = 0; j < nt; ++j) {
        int a1 = o->t1[j], a2 = o->t2[j], a3 = o->t3[j];
        cx[j] = std::fmod((((o->mx[a1] + o->ix[a1] * o->Lx) + (o->mx[a2] + o->ix[a2] * o->Lx) +
                            (o->mx[a3] + o->ix[a3] * o->Lx)) / 3.0), o->Lx);
        cy[j] = std::fmod((((o->my[a1] + o->iy[a1] * o->Ly) + (o->my[a2] + o->iy[a2] * o->Ly) +
                            (o->my[a3] + o->iy[a3] * o->Ly)) / 3.0), o->Ly);
        cz[j] = (o->mz[a1] + o->mz[a2] + o->mz[a3]) / 3.0;
    }
    const double a = 1.0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < o->ncell; ++c) {
        int i, j, k;
        o->cell_ijk(c, i, j, k);
        double xmn = (double)i * a, xmxx = (double)i * a + a;
        double ymn = (double)j * a, ymxx = (double)j * a + a;
        double zmn = -1.0 + (double)k * a, zmxx = -1.0 + (double)k * a + a;
        double cellx = (xmn + xmxx) * 0.5, celly = (ymn + ymxx) * 0.5, cellz = (zmn + zmxx) * 0.5;
        std::vector<int32_t> &L = o->llist[c];
        int count = 0;
        int ntold = (int)L.size();
        for (int t = 0; t < nt; ++t) {
            double xxx1 = cx[t] - cellx;
            double yyy1 = cy[t] - celly;
            xxx1 -= o->Lx * std::round(xxx1 / o->Lx);
            yyy1 -= o->Ly * std::round(yyy1 / o->Ly);
            double zzz = std::fabs(cz[t] - cellz);
            double A1aa = xxx1 * xxx1 + yyy1 * yyy1 + zzz * zzz;
            if (A1aa < 4.0 * a * a) {
                count += 1;
                if (count > ntold)
                    L.push_back(t);
                else
                    L[count - 1] = t;
            }
        }
    }
}

void orc_tricm(orc *o)
{
    /* tricm (Src/memfuncs.cpp:18-31) */
    for (size_t i = 0; i < o->t1.size(); ++i) {
        int a1 = o->t1[i], a2 = o->t2[i], a3 = o->t3[i];
        o->cmx[i] = ((o->mx[a1] + o->ix[a1] * o->Lx) + (o->mx[a2] + o->ix[a2] * o->Lx) +
                     (o->mx[a3] + o->ix[a3] * o->Lx)) / 3.0;
        o->cmy[i] = ((o->my[a1] + o->iy[a1] * o->Ly) + (o->my[a2] + o->iy[a2] * o->Ly) +
                     (o->my[a3] + o->iy[a3] * o->Ly)) / 3.0;
        o->cmz[i] = (o->mz[a1] + o->mz[a2] + o->mz[a3]) / 3.0;
        o->cmx[i] = std::fmod(o->cmx[i], o->Lx);
        o->cmy[i] = std::fmod(o->cmy[i], o->Ly);
    }
}

/* ---------------- streaming + bounce-back ---------------- */

int orc_solve_p3(double *x, double aa, double b, double c)
{
    /* SolveP3 (Src/poly34.cpp:26-55): trigonometric form for three real
     * roots, Cardano otherwise */
    const double TwoPi = 6.28318530717958648;
    const double eps = 1e-14;
    double a2 = aa * aa;
    double q = (a2 - 3 * b) / 9;
    double r = (aa * (2 * a2 - 9 * b) + 27 * c) / 54;
    double r2 = r * r;
    double q3 = q * q * q;
    double A, B;
    if (r2 < q3) {
        double t = r / std::sqrt(q3);
        if (t < -1) t = -1;
        if (t > 1) t = 1;
        t = std::acos(t);
        aa /= 3;
        q = -2 * std::sqrt(q);
        x[0] = q * std::cos(t / 3) - aa;
        x[1] = q * std::cos((t + TwoPi) / 3) - aa;
        x[2] = q * std::cos((t - TwoPi) / 3) - aa;
        return 3;
    }
    A = -std::pow(std::fabs(r) + std::sqrt(r2 - q3), 1. / 3);
    if (r < 0) A = -A;
    B = (A == 0) ? 0 : q / A;
    aa /= 3;
    x[0] = (A + B) - aa;
    x[1] = -0.5 * (A + B) - aa;
    x[2] = 0.5 * std::sqrt(3.) * (A - B);
    if (std::fabs(x[2]) < eps) {
        x[2] = x[1];
        return 2;
    }
    return 1;
}

double orc_solve(double aa, double bb, double cc, double dtmd)
{
    /* Solve (Src/solfuncs.cpp:369-396): smallest root in [0, dtmd], else the
     * next, else the largest */
    double dt[3];
    int n = orc_solve_p3(dt, aa, bb, cc);
    double tt;
    if ((n == 3) || (n == 2)) {
        std::sort(dt, dt + 3);
        tt = dt[0];
        if ((tt < 0) || (tt > dtmd)) {
            tt = dt[1];
            if ((tt < 0) || (tt > dtmd)) tt = dt[2];
        }
    } else {
        tt = dt[0];
    }
    return tt;
}

/* diagnostics of the bounce-back contact times (not part of the
 * reference): hits, hits at a negative contact time, the smallest contact
 * time of a hit, plane crossings examined */
static std::atomic<int64_t> g_hits{0}, g_hits_neg{0}, g_cross{0};
static std::atomic<double> g_tmin{0.0};

void orc_hit_stats(double *out)
{
    out[0] = (double)g_hits.exchange(0);
    out[1] = (double)g_hits_neg.exchange(0);
    out[2] = g_tmin.exchange(0.0);
    out[3] = (double)g_cross.exchange(0);
}

double orc_reflection(orc *o, int64_t i)
{
    /* reflectionFed(i, 0) (Src/solfuncs.cpp:400-598): bounce-back off the
     * moving triangles listed for the particle's unshifted cell */
    const double a = 1.0, dtmd = o->p.dtmd, rstep = 0.0;
    const double Lx = o->Lx, Ly = o->Ly;
    double fulldt = rstep;
    int fxi = (int)std::floor(wrap_loop(o->x[i], Lx));
    int fyi = (int)std::floor(wrap_loop(o->y[i], Ly));
    int fzi = (int)std::floor(o->z[i] + 1.0);
    int32_t cid = o->cell_of(fxi, fyi, fzi);
    const std::vector<int32_t> &L = o->llist[cid];
    for (size_t j = 0; j < L.size(); ++j) {
        int flg = 0;
        int tj = L[j];
        int p1 = o->t1[tj], p2 = o->t2[tj], p3 = o->t3[tj];
        double solvx = o->x[i], solvy = o->y[i], solvz = o->z[i];
        double A1[3], B1[3], C1[3], A2[3], B2[3], C2[3];
        A1[0] = o->mx[p1] + o->mvx[p1] * rstep;
        A1[1] = o->my[p1] + o->mvy[p1] * rstep;
        A1[2] = o->mz[p1] + o->mvz[p1] * rstep;
        B1[0] = o->mx[p2] + o->mvx[p2] * rstep;
        B1[1] = o->my[p2] + o->mvy[p2] * rstep;
        B1[2] = o->mz[p2] + o->mvz[p2] * rstep;
        C1[0] = o->mx[p3] + o->mvx[p3] * rstep;
        C1[1] = o->my[p3] + o->mvy[p3] * rstep;
        C1[2] = o->mz[p3] + o->mvz[p3] * rstep;
        /* periodic stencil (:468-496) */
        double rsx = Lx - (3.0 * a), lsx = 3.0 * a, rsy = Ly - (3.0 * a), lsy = 3.0 * a;
        if (solvx < lsx) {
            A1[0] -= (A1[0] > rsx) ? Lx : 0;
            B1[0] -= (B1[0] > rsx) ? Lx : 0;
            C1[0] -= (C1[0] > rsx) ? Lx : 0;
        } else if (solvx > rsx) {
            A1[0] += (A1[0] < lsx) ? Lx : 0;
            B1[0] += (B1[0] < lsx) ? Lx : 0;
            C1[0] += (C1[0] < lsx) ? Lx : 0;
        }
        if (solvy < lsy) {
            A1[1] -= (A1[1] > rsy) ? Ly : 0;
            B1[1] -= (B1[1] > rsy) ? Ly : 0;
            C1[1] -= (C1[1] > rsy) ? Ly : 0;
        } else if (solvy > rsy) {
            A1[1] += (A1[1] < lsy) ? Ly : 0;
            B1[1] += (B1[1] < lsy) ? Ly : 0;
            C1[1] += (C1[1] < lsy) ? Ly : 0;
        }
        double hdt = dtmd - rstep;
        A2[0] = A1[0] + o->mvx[p1] * hdt;
        A2[1] = A1[1] + o->mvy[p1] * hdt;
        A2[2] = A1[2] + o->mvz[p1] * hdt;
        B2[0] = B1[0] + o->mvx[p2] * hdt;
        B2[1] = B1[1] + o->mvy[p2] * hdt;
        B2[2] = B1[2] + o->mvz[p2] * hdt;
        C2[0] = C1[0] + o->mvx[p3] * hdt;
        C2[1] = C1[1] + o->mvy[p3] * hdt;
        C2[2] = C1[2] + o->mvz[p3] * hdt;
        double X[3], Xdt[3], e[3], f[3];
        e[0] = B1[0] - A1[0]; e[1] = B1[1] - A1[1]; e[2] = B1[2] - A1[2];
        f[0] = C1[0] - A1[0]; f[1] = C1[1] - A1[1]; f[2] = C1[2] - A1[2];
        cross(e, f, X);
        e[0] = B2[0] - A2[0]; e[1] = B2[1] - A2[1]; e[2] = B2[2] - A2[2];
        f[0] = C2[0] - A2[0]; f[1] = C2[1] - A2[1]; f[2] = C2[2] - A2[2];
        cross(e, f, Xdt);
        double PSt[3], PSdt[3];
        PSt[0] = solvx - A1[0];
        PSt[1] = solvy - A1[1];
        PSt[2] = solvz - A1[2];
        PSdt[0] = (solvx + o->vx[i] * hdt) - A2[0];
        PSdt[1] = (solvy + o->vy[i] * hdt) - A2[1];
        PSdt[2] = (solvz + o->vz[i] * hdt) - A2[2];
        double bt = X[0] * PSt[0] + X[1] * PSt[1] + X[2] * PSt[2];
        double bdt = Xdt[0] * PSdt[0] + Xdt[1] * PSdt[1] + Xdt[2] * PSdt[2];
        if ((o->indio[i] == 1) && (bt < 0)) flg = 1;
        if ((o->indio[i] == 0) && (bt > 0)) flg = 1;
        if ((bt * bdt <= 0) && (flg == 0)) {
            double a10[3], a20[3], d1[3], d2[3], a1[3], a2[3], b1[3], b2[3], b3[3], d11[3], d22[3];
            a10[0] = B1[0] - A1[0]; a10[1] = B1[1] - A1[1]; a10[2] = B1[2] - A1[2];
            a20[0] = C1[0] - A1[0]; a20[1] = C1[1] - A1[1]; a20[2] = C1[2] - A1[2];
            d1[0] = o->mvx[p2] - o->mvx[p1]; d1[1] = o->mvy[p2] - o->mvy[p1]; d1[2] = o->mvz[p2] - o->mvz[p1];
            d2[0] = o->mvx[p3] - o->mvx[p1]; d2[1] = o->mvy[p3] - o->mvy[p1]; d2[2] = o->mvz[p3] - o->mvz[p1];
            a1[0] = PSt[0]; a1[1] = PSt[1]; a1[2] = PSt[2];
            a2[0] = o->vx[i] - o->mvx[p1]; a2[1] = o->vy[i] - o->mvy[p1]; a2[2] = o->vz[i] - o->mvz[p1];
            cross(a10, a20, b1);
            cross(a10, d2, d11);
            cross(d1, a20, d22);
            cross(d1, d2, b3);
            b2[0] = d11[0] + d22[0]; b2[1] = d11[1] + d22[1]; b2[2] = d11[2] + d22[2];
            double a2b3 = a2[0] * b3[0] + a2[1] * b3[1] + a2[2] * b3[2];
            double a1b3 = a1[0] * b3[0] + a1[1] * b3[1] + a1[2] * b3[2];
            double a2b2 = a2[0] * b2[0] + a2[1] * b2[1] + a2[2] * b2[2];
            double a1b2 = a1[0] * b2[0] + a1[1] * b2[1] + a1[2] * b2[2];
            double a2b1 = a2[0] * b1[0] + a2[1] * b1[1] + a2[2] * b1[2];
            double a1b1 = a1[0] * b1[0] + a1[1] * b1[1] + a1[2] * b1[2];
            double tt = orc_solve((a1b3 + a2b2) / a2b3, (a1b2 + a2b1) / a2b3, a1b1 / a2b3, dtmd);
            double pd[3];
            PSdt[0] = (solvx + o->vx[i] * tt) - (A1[0] + o->mvx[p1] * tt);
            PSdt[1] = (solvy + o->vy[i] * tt) - (A1[1] + o->mvy[p1] * tt);
            PSdt[2] = (solvz + o->vz[i] * tt) - (A1[2] + o->mvz[p1] * tt);
            pd[0] = (solvx + o->vx[i] * tt);
            pd[1] = (solvy + o->vy[i] * tt);
            pd[2] = (solvz + o->vz[i] * tt);
            a10[0] = (B1[0] + o->mvx[p2] * tt) - (A1[0] + o->mvx[p1] * tt);
            a10[1] = (B1[1] + o->mvy[p2] * tt) - (A1[1] + o->mvy[p1] * tt);
            a10[2] = (B1[2] + o->mvz[p2] * tt) - (A1[2] + o->mvz[p1] * tt);
            a20[0] = (C1[0] + o->mvx[p3] * tt) - (A1[0] + o->mvx[p1] * tt);
            a20[1] = (C1[1] + o->mvy[p3] * tt) - (A1[1] + o->mvy[p1] * tt);
            a20[2] = (C1[2] + o->mvz[p3] * tt) - (A1[2] + o->mvz[p1] * tt);
            double pa1 = PSdt[0] * a10[0] + PSdt[1] * a10[1] + PSdt[2] * a10[2];
            double pa2 = PSdt[0] * a20[0] + PSdt[1] * a20[1] + PSdt[2] * a20[2];
            double a1a2 = a10[0] * a20[0] + a10[1] * a20[1] + a10[2] * a20[2];
            double a12 = a10[0] * a10[0] + a10[1] * a10[1] + a10[2] * a10[2];
            double a22 = a20[0] * a20[0] + a20[1] * a20[1] + a20[2] * a20[2];
            double denom = (a12 * a22) - (a1a2 * a1a2);
            double eta = ((pa1 * a22) - (pa2 * a1a2)) / denom;
            double zeta = ((pa2 * a12) - (pa1 * a1a2)) / denom;
            g_cross.fetch_add(1, std::memory_order_relaxed);
            if ((eta >= 0) && (zeta >= 0) && (eta + zeta <= 1) && (tt < (dtmd - rstep))) {
                g_hits.fetch_add(1, std::memory_order_relaxed);
                if (tt < 0) {
                    g_hits_neg.fetch_add(1, std::memory_order_relaxed);
                    double cur = g_tmin.load();
                    while (tt < cur && !g_tmin.compare_exchange_weak(cur, tt)) {}
                }
                fulldt += tt;
                double vd[3];
                vd[0] = (1.0 - eta - zeta) * o->mvx[p1] + eta * o->mvx[p2] + zeta * o->mvx[p3];
                vd[1] = (1.0 - eta - zeta) * o->mvy[p1] + eta * o->mvy[p2] + zeta * o->mvy[p3];
                vd[2] = (1.0 - eta - zeta) * o->mvz[p1] + eta * o->mvz[p2] + zeta * o->mvz[p3];
                o->vx[i] = 2.0 * vd[0] - o->vx[i];
                o->vy[i] = 2.0 * vd[1] - o->vy[i];
                o->vz[i] = 2.0 * vd[2] - o->vz[i];
                o->x[i] = pd[0];
                o->y[i] = pd[1];
                o->z[i] = pd[2];
            }
        }
    }
    return fulldt;
}

void orc_stream(orc *o, double wallspeed)
{
    /* solvent streaming loop (maincode.cpp:510-562) */
    o->wallspeed = wallspeed;
    const double dtmd = o->p.dtmd, Lz = o->Lz, Lx = o->Lx, Ly = o->Ly;
    int64_t ns = (int64_t)o->x.size();
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < ns; ++i) {
        double fullts = orc_reflection(o, i);
        double rem = dtmd - fullts;
        double &x = o->x[i], &y = o->y[i], &z = o->z[i];
        double &vx = o->vx[i], &vy = o->vy[i], &vz = o->vz[i];
        x += vx * rem;
        y += vy * rem;
        z += vz * rem;
        double ul = Lz, ll = 0;
        if (z > ul || z < ll) {
            double wl = (z > ul) ? ul : ll;
            double A[3], B[3];
            A[0] = x - vx * rem;
            A[1] = y - vy * rem;
            A[2] = z - vz * rem;
            double s = (wl - A[2]) / (z - A[2]);
            B[0] = A[0] + s * (x - A[0]);
            B[1] = A[1] + s * (y - A[1]);
            B[2] = A[2] + s * (z - A[2]);
            double dist = std::sqrt(((B[0] - A[0]) * (B[0] - A[0])) + ((B[1] - A[1]) * (B[1] - A[1])) +
                                    ((B[2] - A[2]) * (B[2] - A[2])));
            double dtt = dist / std::sqrt((vx * vx) + (vy * vy) + (vz * vz));
            if (z > ul) {
                vx = wallspeed;
                vy = 0;
                vz = 0;
            } else {
                vx *= -1.0;
                vy *= -1.0;
                vz *= -1.0;
            }
            x = B[0] + vx * (rem - dtt);
            y = B[1] + vy * (rem - dtt);
            z = B[2] + vz * (rem - dtt);
        }
        x = wrap_loop(x, Lx);
        y = wrap_loop(y, Ly);
    }
    o->round_solvent();
}

/* ---------------- membrane ---------------- */

void orc_kick1(orc *o)
{
    /* maincode.cpp:487-496 */
    double h = o->p.dtmd, m2 = 2.0 * o->p.pmass;
    for (int i = 0; i < o->nv; ++i) {
        o->mvx[i] += (o->fx[i] * h / m2);
        o->mvy[i] += (o->fy[i] * h / m2);
        o->mvz[i] += (o->fz[i] * h / m2);
        o->fx[i] = 0.0;
        o->fy[i] = 0.0;
        o->fz[i] = 0.0;
    }
}

void orc_drift(orc *o)
{
    /* membrane drift + bboundmx/my (maincode.cpp:566-574, Src/solfuncs.cpp:625-659) */
    double h = o->p.dtmd;
    for (int i = 0; i < o->nv; ++i) {
        o->mx[i] += (o->mvx[i] * h);
        o->my[i] += (o->mvy[i] * h);
        o->mz[i] += (o->mvz[i] * h);
        double xx = o->mx[i];
        if (xx > o->Lx) { xx -= o->Lx; o->ix[i] += 1; }
        if (xx < 0) { xx += o->Lx; o->ix[i] -= 1; }
        o->mx[i] = xx;
        double yy = o->my[i];
        if (yy > o->Ly) { yy -= o->Ly; o->iy[i] += 1; }
        if (yy < 0) { yy += o->Ly; o->iy[i] -= 1; }
        o->my[i] = yy;
    }
    /* parasite drift (maincode.cpp:576-598) */
    double PR = o->p.PRmass;
    for (int k = 0; k < 3; ++k)
        o->para[k] += (o->paraV[k] * h) + (o->paraft[k] * h * h / (2.0 * PR));
    if (o->para[0] > o->Lx) { o->para[0] -= o->Lx; o->paradx += 1; }
    else if (o->para[0] < 0) { o->para[0] += o->Lx; o->paradx -= 1; }
    if (o->para[1] > o->Ly) { o->para[1] -= o->Ly; o->parady += 1; }
    else if (o->para[1] < 0) { o->para[1] += o->Ly; o->parady -= 1; }
}

void orc_totalva(orc *o)
{
    /* TotalVA (Src/memfuncs.cpp:161-190): two ordered sums */
    double totA = 0, totV = 0;
    int nt = (int)o->t1.size();
    for (int i = 0; i < nt; ++i) totA += o->areatri(o->t1[i], o->t2[i], o->t3[i]);
    o->Sarea = totA;
    for (int i = 0; i < nt; ++i) totV += o->voltri(o->t1[i], o->t2[i], o->t3[i]);
    o->Volume = totV;
}

void orc_totalva0(orc *o)
{
    /* TotalVA0 (Src/memfuncs.cpp:192-211): one interleaved loop */
    double totA = 0, totV = 0;
    for (size_t i = 0; i < o->t1.size(); ++i) {
        totA += o->areatri(o->t1[i], o->t2[i], o->t3[i]);
        totV += o->voltri(o->t1[i], o->t2[i], o->t3[i]);
    }
    o->Sarea0 = totA;
    o->Volume0 = totV;
}

void orc_pfdt(orc *o)
{
    /* pfdt: WCA between the parasite bead and each vertex (Src/memfuncs.cpp:443-484) */
    double cutoff = o->p.paraR, epsilon = 0.1, sigma = 0.890898718 * cutoff;
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < o->nv; ++i) {
        double A[3], X[3];
        A[0] = (o->para[0] + o->paradx * o->Lx) - (o->mx[i] + o->ix[i] * o->Lx);
        A[1] = (o->para[1] + o->parady * o->Ly) - (o->my[i] + o->iy[i] * o->Ly);
        A[2] = o->para[2] - o->mz[i];
        double aa = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
        if (aa <= cutoff * cutoff) {
            double ex = (sigma * sigma * sigma * sigma * sigma * sigma) / (aa * aa * aa);
            X[0] = (24.0 * epsilon / aa) * ((2.0 * ex * ex) - ex) * A[0];
            X[1] = (24.0 * epsilon / aa) * ((2.0 * ex * ex) - ex) * A[1];
            X[2] = (24.0 * epsilon / aa) * ((2.0 * ex * ex) - ex) * A[2];
            sx += X[0];
            sy += X[1];
            sz += X[2];
            o->fx[i] -= X[0];
            o->fy[i] -= X[1];
            o->fz[i] -= X[2];
        }
    }
    o->parafdt[0] += sx;
    o->parafdt[1] += sy;
    o->parafdt[2] += sz;
}

void orc_forces(orc *o)
{
    /* force_calculate_dt (Src/memfuncs.cpp:214-441) at one thread: springs,
     * area/volume penalties and bending accumulate into ff, then f += ff */
    int nv = o->nv, ne = (int)o->e1.size(), nt = (int)o->t1.size();
    std::vector<double> ffx(nv, 0.0), ffy(nv, 0.0), ffz(nv, 0.0);
    const double elap = o->p.elap;
    auto P = [&](int v, double r[3]) { r[0] = o->ux(v); r[1] = o->uy(v); r[2] = o->mz[v]; };
    for (int i = 0; i < ne; ++i) {
        int l1 = o->e1[i], l2 = o->e2[i];
        double l0 = o->l0[i];
        double A[3];
        A[0] = o->ux(l1) - o->ux(l2);
        A[1] = o->uy(l1) - o->uy(l2);
        A[2] = o->mz[l1] - o->mz[l2];
        double aa = std::sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        A[0] /= aa;
        A[1] /= aa;
        A[2] /= aa;
        double xx = aa / (2.5 * l0);
        double xxx = 1.0 / (4.0 * (1.0 - xx) * (1.0 - xx));
        ffx[l1] += (-1.0 / (elap)) * (xxx - 0.25 + xx) * A[0];
        ffy[l1] += (-1.0 / (elap)) * (xxx - 0.25 + xx) * A[1];
        ffz[l1] += (-1.0 / (elap)) * (xxx - 0.25 + xx) * A[2];
        ffx[l2] += (1.0 / (elap)) * (xxx - 0.25 + xx) * A[0];
        ffy[l2] += (1.0 / (elap)) * (xxx - 0.25 + xx) * A[1];
        ffz[l2] += (1.0 / (elap)) * (xxx - 0.25 + xx) * A[2];
        double kp = o->kappap[i];
        ffx[l1] += kp * A[0] / (aa * aa);
        ffy[l1] += kp * A[1] / (aa * aa);
        ffz[l1] += kp * A[2] / (aa * aa);
        ffx[l2] += -kp * A[0] / (aa * aa);
        ffy[l2] += -kp * A[1] / (aa * aa);
        ffz[l2] += -kp * A[2] / (aa * aa);
    }
    const double kappasa = o->p.kappasa, kappav = o->p.kappav, kappad = o->p.kappad;
    for (int i = 0; i < nt; ++i) {
        int l1 = o->t1[i], l2 = o->t2[i], l3 = o->t3[i];
        double beta = -kappasa * (o->Sarea - o->Sarea0) / o->Sarea0;
        double betav = -kappav * (o->Volume - o->Volume0) / (o->Volume0);
        double area = o->areatri(l1, l2, l3);
        double alpha = beta / (4.0 * area);
        double alphad = -kappad * (area - o->tarea[i]) / (4.0 * o->tarea[i] * area);
        double p1[3], p2[3], p3[3], A21[3], A13[3], A32[3], eta[3], C[3], B[3];
        P(l1, p1); P(l2, p2); P(l3, p3);
        for (int k = 0; k < 3; ++k) {
            A21[k] = p2[k] - p1[k];
            A13[k] = p1[k] - p3[k];
            A32[k] = p3[k] - p2[k];
        }
        cross(A13, A21, eta);
        cross(eta, A32, C);
        ffx[l1] += alpha * C[0]; ffy[l1] += alpha * C[1]; ffz[l1] += alpha * C[2];
        ffx[l1] += alphad * C[0]; ffy[l1] += alphad * C[1]; ffz[l1] += alphad * C[2];
        cross(eta, A13, C);
        ffx[l2] += alpha * C[0]; ffy[l2] += alpha * C[1]; ffz[l2] += alpha * C[2];
        ffx[l2] += alphad * C[0]; ffy[l2] += alphad * C[1]; ffz[l2] += alphad * C[2];
        cross(eta, A21, C);
        ffx[l3] += alpha * C[0]; ffy[l3] += alpha * C[1]; ffz[l3] += alpha * C[2];
        ffx[l3] += alphad * C[0]; ffy[l3] += alphad * C[1]; ffz[l3] += alphad * C[2];
        for (int k = 0; k < 3; ++k) B[k] = (p1[k] + p2[k] + p3[k]) / 3.0;
        cross(B, A32, C);
        ffx[l1] += betav * ((eta[0] / 3.0) + C[0]) / 6.0;
        ffy[l1] += betav * ((eta[1] / 3.0) + C[1]) / 6.0;
        ffz[l1] += betav * ((eta[2] / 3.0) + C[2]) / 6.0;
        cross(B, A13, C);
        ffx[l2] += betav * ((eta[0] / 3.0) + C[0]) / 6.0;
        ffy[l2] += betav * ((eta[1] / 3.0) + C[1]) / 6.0;
        ffz[l2] += betav * ((eta[2] / 3.0) + C[2]) / 6.0;
        cross(B, A21, C);
        ffx[l3] += betav * ((eta[0] / 3.0) + C[0]) / 6.0;
        ffy[l3] += betav * ((eta[1] / 3.0) + C[1]) / 6.0;
        ffz[l3] += betav * ((eta[2] / 3.0) + C[2]) / 6.0;
    }
    const double kappab = o->p.kappab, Pi = o->p.Pi;
    int l2 = 0, l3 = 0;
    for (int i = 0; i < ne; ++i) {
        if (o->iedge[i] == 1) { l2 = o->e1[i]; l3 = o->e2[i]; }
        if (o->iedge[i] == -1) { l3 = o->e1[i]; l2 = o->e2[i]; }
        int l1 = o->es1[i], l4 = o->es2[i];
        double p1[3], p2[3], p3[3], p4[3];
        P(l1, p1); P(l2, p2); P(l3, p3); P(l4, p4);
        double a21[3], a31[3], a24[3], a34[3], a32[3], t1[3], t2[3];
        for (int k = 0; k < 3; ++k) {
            a21[k] = p2[k] - p1[k];
            a31[k] = p3[k] - p1[k];
            a24[k] = p2[k] - p4[k];
            a34[k] = p3[k] - p4[k];
            a32[k] = p3[k] - p2[k];
            t1[k] = (p1[k] + p2[k] + p3[k]) / 3.0;
            t2[k] = (p4[k] + p2[k] + p3[k]) / 3.0;
        }
        double eta[3], zeta[3];
        cross(a21, a31, eta);
        cross(a34, a24, zeta);
        double etam = std::sqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
        double zetam = std::sqrt(zeta[0] * zeta[0] + zeta[1] * zeta[1] + zeta[2] * zeta[2]);
        double costh = (eta[0] * zeta[0] + eta[1] * zeta[1] + eta[2] * zeta[2]) / (etam * zetam);
        if (costh > 1.0) costh = 1.0;
        if (costh < -1.0) costh = -1.0;
        if (costh == 1.0) costh -= 0.0005;
        if (costh == -1.0) costh += 0.0005;
        double sinth = std::sqrt(1.0 - costh * costh);
        double tt0 = t1[0] - t2[0];
        double tt1 = t1[1] - t2[1];
        double factor = (((eta[0] - zeta[0]) * tt0) + ((eta[1] - zeta[1]) * tt1) +
                         ((eta[2] - zeta[2]) * (t1[2] - t2[2])));
        double fact = (factor >= 0.0) ? 1.0 : -1.0;
        sinth = fact * sinth;
        double betab = kappab * (sinth * std::cos(o->theta0 * Pi / 180.0) -
                                 costh * std::sin(o->theta0 * Pi / 180.0)) / sinth;
        double b11 = -(betab * costh) / (etam * etam);
        double b22 = -(betab * costh) / (zetam * zetam);
        double b12 = betab / (etam * zetam);
        double C[3], D[3], E[3], F[3];
        cross(eta, a32, C);
        cross(zeta, a32, D);
        ffx[l1] += b11 * C[0] + b12 * D[0];
        ffy[l1] += b11 * C[1] + b12 * D[1];
        ffz[l1] += b11 * C[2] + b12 * D[2];
        ffx[l4] += -b12 * C[0] - b22 * D[0];
        ffy[l4] += -b12 * C[1] - b22 * D[1];
        ffz[l4] += -b12 * C[2] - b22 * D[2];
        cross(eta, a31, C);
        cross(zeta, a31, D);
        cross(eta, a34, E);
        cross(zeta, a34, F);
        ffx[l2] += -b11 * C[0] + b12 * (E[0] - D[0]) + b22 * F[0];
        ffy[l2] += -b11 * C[1] + b12 * (E[1] - D[1]) + b22 * F[1];
        ffz[l2] += -b11 * C[2] + b12 * (E[2] - D[2]) + b22 * F[2];
        cross(eta, a21, C);
        cross(zeta, a21, D);
        cross(eta, a24, E);
        cross(zeta, a24, F);
        ffx[l3] += b11 * C[0] + b12 * (-E[0] + D[0]) - b22 * F[0];
        ffy[l3] += b11 * C[1] + b12 * (-E[1] + D[1]) - b22 * F[1];
        ffz[l3] += b11 * C[2] + b12 * (-E[2] + D[2]) - b22 * F[2];
    }
    for (int i = 0; i < nv; ++i) {
        o->fx[i] += ffx[i];
        o->fy[i] += ffy[i];
        o->fz[i] += ffz[i];
    }
}

void orc_external(orc *o)
{
    /* gravity, bond springs, wall repulsion (maincode.cpp:606-665) */
    int nv = o->nv;
    for (int i = 0; i < nv; ++i) o->fz[i] += -o->p.Fg / (double)nv;
    const double ks = o->p.ks, el0 = o->p.el0;
    for (int i = 0; i < nv; ++i) {
        if (o->bondi[i] == 1) {
            int jj = o->bondl[i];
            double A[3];
            A[0] = o->mx[i] - o->xlig[jj];
            A[1] = o->my[i] - o->ylig[jj];
            A[2] = o->mz[i];
            A[0] -= o->Lx * std::round(A[0] / o->Lx);
            A[1] -= o->Ly * std::round(A[1] / o->Ly);
            double aa = std::sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
            A[0] /= aa;
            A[1] /= aa;
            A[2] /= aa;
            if (aa > el0) {
                o->fx[i] += -ks * (aa - el0) * A[0];
                o->fy[i] += -ks * (aa - el0) * A[1];
                o->fz[i] += -ks * (aa - el0) * A[2];
            }
        }
    }
    for (int i = 0; i < nv; ++i) {
        double c = o->mz[i];
        if (c < 0.5) o->fz[i] += 300.0 * std::exp(-50.0 * c);
    }
    for (int i = 0; i < nv; ++i) {
        double c = o->Lz - o->mz[i];
        if (c < 0.5) o->fz[i] -= 300.0 * std::exp(-50.0 * c);
    }
}

void orc_bond_prep(orc *o)
{
    /* unbinding probability for bonded knobs and sorted candidate ligands
     * for free knobs near the wall (maincode.cpp:668-723) */
    const double ks = o->p.ks, el0 = o->p.el0, dthr = o->p.dthr, dtmd = o->p.dtmd;
    for (int ii = 0; ii < o->nv; ++ii) {
        o->prep_live[ii] = 0;
        if (o->bondi[ii] == 1) {
            int p1modx = o->ix[ii], p1mody = o->iy[ii];
            int jj = o->bondl[ii];
            double A[3], B[3];
            A[0] = (o->mx[ii] + p1modx * o->Lx) - (o->xlig[jj] + p1modx * o->Lx);
            A[1] = (o->my[ii] + p1mody * o->Ly) - (o->ylig[jj] + p1mody * o->Ly);
            A[2] = o->mz[ii] - 0.0;
            double xx = std::sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
            if (xx > el0) {
                B[0] = -ks * (xx - el0) * A[0] / xx;
                B[1] = -ks * (xx - el0) * A[1] / xx;
                B[2] = -ks * (xx - el0) * A[2] / xx;
            } else {
                B[0] = B[1] = B[2] = 0.0;
            }
            double bb = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
            double koff = o->p.offrate * std::exp(bb / o->p.FD);
            o->poff[ii] = 1.0 - std::exp(-koff * dtmd);
        } else if ((o->mz[ii] < dthr) && (o->bondi[ii] == 0) && (o->knob[ii] == 1)) {
            auto &pr = o->pairs[ii];
            pr.clear();
            o->prep_live[ii] = 1;
            int c = o->mpid[ii];
            double kon = o->p.onrate;
            o->pon[ii] = 1.0 - std::exp(-kon * dtmd);
            int p1modx = o->ix[ii], p1mody = o->iy[ii];
            for (int64_t j = 0; j < o->nlig; ++j) {
                if (!o->lig_member(c, j)) continue;
                double A[3];
                A[0] = (o->mx[ii] + p1modx * o->Lx) - (o->xlig[j] + p1modx * o->Lx);
                A[1] = (o->my[ii] + p1mody * o->Ly) - (o->ylig[j] + p1mody * o->Ly);
                A[2] = o->mz[ii] - 0.0;
                double aaa = (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
                if (aaa <= dthr * dthr) pr.push_back(std::make_pair(aaa, (int)j));
            }
            std::sort(pr.begin(), pr.end());
        }
    }
}

void orc_bond_dynamics(orc *o)
{
    /* bond_dynamics (maincode.cpp:16-54): vertices in order against the
     * shared ligand occupancy; a free knob keeps drawing after a bind */
    if (!o->tapes) o->bond_u.clear();
    o->bond_pos = 0;
    for (int ii = 0; ii < o->nv; ++ii) {
        if (o->bondi[ii] == 1) {
            double randoff = o->d_bond_at(ii, 0);
            int jj = o->bondl[ii];
            if (o->poff[ii] > randoff) {
                o->bondi[ii] = 0;
                o->bondl[ii] = 12345678;
                o->ligi[jj] = 0;
            }
        } else if ((o->mz[ii] < o->p.dthr) && (o->bondi[ii] == 0) && (o->knob[ii] == 1)) {
            for (size_t k = 0; k < o->pairs[ii].size(); ++k) {
                int j = o->pairs[ii][k].second;
                double pon = o->pon[ii];
                double randon = o->d_bond_at(ii, (int)k);
                if ((pon > randon) && (o->ligi[j] == 0)) {
                    o->ligi[j] = 1;
                    o->bondi[ii] = 1;
                    o->bondl[ii] = j;
                }
            }
        }
    }
}

void orc_kick2(orc *o)
{
    /* maincode.cpp:727-744 */
    double h = o->p.dtmd, m2 = 2.0 * o->p.pmass, PR = o->p.PRmass;
    for (int i = 0; i < o->nv; ++i) {
        o->mvx[i] += (o->fx[i] * h / m2);
        o->mvy[i] += (o->fy[i] * h / m2);
        o->mvz[i] += (o->fz[i] * h / m2);
    }
    for (int k = 0; k < 3; ++k) {
        o->paraV[k] += ((o->paraft[k] + o->parafdt[k]) * h / (2.0 * PR));
        o->paraft[k] = o->parafdt[k];
        o->parafdt[k] = 0.0;
    }
}

void orc_centerofmass(orc *o)
{
    /* centerofmass (Src/memfuncs.cpp:55-74) */
    double rr[3] = {0, 0, 0};
    for (int i = 0; i < o->nv; ++i) {
        rr[0] += o->mx[i] + o->ix[i] * o->Lx;
        rr[1] += o->my[i] + o->iy[i] * o->Ly;
        rr[2] += o->mz[i];
    }
    o->center[0] = std::fmod(rr[0] / (double)o->nv, o->Lx);
    o->center[1] = std::fmod(rr[1] / (double)o->nv, o->Ly);
    o->center[2] = std::fmod(rr[2] / (double)o->nv, o->Lz);
}

void orc_makeiedge(orc *o)
{
    /* makeiedge (Src/memfuncs.cpp:533-588): orientation of each hinge
     * relative to the centre of mass; wrapped coordinates */
    for (size_t i = 0; i < o->e1.size(); ++i) {
        int l2 = o->e1[i], l3 = o->e2[i], l1 = o->es1[i], l4 = o->es2[i];
        auto V = [&](int v, double r[3]) { r[0] = o->mx[v]; r[1] = o->my[v]; r[2] = o->mz[v]; };
        double p1[3], p2[3], p3[3], p4[3];
        V(l1, p1); V(l2, p2); V(l3, p3); V(l4, p4);
        double a21[3], a31[3], a24[3], a34[3], t1[3], t2[3], tt1[3], tt2[3], eta[3], zeta[3];
        for (int k = 0; k < 3; ++k) {
            a21[k] = p2[k] - p1[k];
            a31[k] = p3[k] - p1[k];
            a24[k] = p2[k] - p4[k];
            a34[k] = p3[k] - p4[k];
            t1[k] = (p1[k] + p2[k] + p3[k]) / 3.0;
            t2[k] = (p4[k] + p2[k] + p3[k]) / 3.0;
            tt1[k] = t1[k] - o->center[k];
            tt2[k] = t2[k] - o->center[k];
        }
        cross(a21, a31, eta);
        cross(a34, a24, zeta);
        double che1 = eta[0] * tt1[0] + eta[1] * tt1[1] + eta[2] * tt1[2];
        double che2 = zeta[0] * tt2[0] + zeta[1] * tt2[1] + zeta[2] * tt2[2];
        if ((che1 >= 0) && (che2 >= 0))
            o->iedge[i] = 1;
        else if ((che1 < 0) && (che2 < 0))
            o->iedge[i] = -1;
        else
            o->iedge[i] = 1;
    }
}

void orc_set_wallspeed(orc *o, double w) { o->wallspeed = w; }

void orc_classify(orc *o)
{
    /* the flag loop of sortsolvent_initial1 (Src/solfuncs.cpp:17-94):
     * approximate ray parity (+x ray, triangle centroid ahead of the particle) */
    orc_tricm(o);
    int64_t ns = (int64_t)o->x.size();
    int nt = (int)o->t1.size();
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < ns; ++i) {
        int count = 0;
        double Rd[3] = {1.0, 0.0, 0.0};
        double Q1[3], Q2[3];
        Q1[0] = o->x[i] + Rd[0] * 0.3;
        Q1[1] = o->y[i] + Rd[1] * 0.3;
        Q1[2] = o->z[i] + Rd[2] * 0.3;
        Q2[0] = o->x[i] + Rd[0] * 0.7;
        Q2[1] = o->y[i] + Rd[1] * 0.7;
        Q2[2] = o->z[i] + Rd[2] * 0.7;
        for (int j = 0; j < nt; ++j) {
            int p1 = o->t1[j], p2 = o->t2[j], p3 = o->t3[j];
            double X[3];
            o->normaldt(p1, p2, p3, X);
            double dotP1Q1 = (o->mx[p1] - Q1[0]) * X[0] + (o->my[p1] - Q1[1]) * X[1] + (o->mz[p1] - Q1[2]) * X[2];
            double dotQ2Q1 = (Q2[0] - Q1[0]) * X[0] + (Q2[1] - Q1[1]) * X[1] + (Q2[2] - Q1[2]) * X[2];
            double P0[3];
            P0[0] = Q1[0] + ((dotP1Q1 / dotQ2Q1) * (Q2[0] - Q1[0]));
            P0[1] = Q1[1] + ((dotP1Q1 / dotQ2Q1) * (Q2[1] - Q1[1]));
            P0[2] = Q1[2] + ((dotP1Q1 / dotQ2Q1) * (Q2[2] - Q1[2]));
            double v1[3] = {o->mx[p1], o->my[p1], o->mz[p1]};
            double v2[3] = {o->mx[p2], o->my[p2], o->mz[p2]};
            double v3[3] = {o->mx[p3], o->my[p3], o->mz[p3]};
            double P2P1[3], P3P2[3], P1P3[3], P0P1[3], P0P2[3], P0P3[3], f1[3], f2[3], f3[3];
            for (int k = 0; k < 3; ++k) {
                P2P1[k] = v2[k] - v1[k];
                P3P2[k] = v3[k] - v2[k];
                P1P3[k] = v1[k] - v3[k];
                P0P1[k] = P0[k] - v1[k];
                P0P2[k] = P0[k] - v2[k];
                P0P3[k] = P0[k] - v3[k];
            }
            cross(P2P1, P0P1, f1);
            cross(P3P2, P0P2, f2);
            cross(P1P3, P0P3, f3);
            double ch1 = f1[0] * X[0] + f1[1] * X[1] + f1[2] * X[2];
            double ch2 = f2[0] * X[0] + f2[1] * X[1] + f2[2] * X[2];
            double ch3 = f3[0] * X[0] + f3[1] * X[1] + f3[2] * X[2];
            if ((ch1 >= 0) && (ch2 >= 0) && (ch3 >= 0) && (o->cmx[j] >= o->x[i])) count += 1;
        }
        o->indio[i] = ((count % 2) == 0) ? 1 : 0;
    }
}

void orc_sortsolvent(orc *o)
{
    /* sortsolvent_initial1 (Src/solfuncs.cpp:17-106): flags, then std::sort
     * of the whole particle array by the flag */
    orc_classify(o);
    int64_t ns = (int64_t)o->x.size();
    /* std::sort on the flag alone: sorting (flag, index) records with the
     * same comparator reproduces libstdc++'s permutation of the particles */
    struct Rec {
        int32_t f;
        int64_t i;
        bool operator<(const Rec &r) const { return f < r.f; }
    };
    std::vector<Rec> rec(ns);
    for (int64_t i = 0; i < ns; ++i) rec[i] = Rec{o->indio[i], i};
    std::sort(rec.begin(), rec.end());
    auto perm = [&](std::vector<double> &v) {
        std::vector<double> t(ns);
        for (int64_t k = 0; k < ns; ++k) t[k] = v[rec[k].i];
        v.swap(t);
    };
    perm(o->x); perm(o->y); perm(o->z); perm(o->vx); perm(o->vy); perm(o->vz);
    for (int64_t k = 0; k < ns; ++k) o->indio[k] = rec[k].f;
    o->nsolb = 0;
    for (int64_t i = 0; i + 1 < ns; ++i)
        if ((o->indio[i] == 0) && (o->indio[i + 1] == 1)) { o->nsolb = i; break; }
}

/* ---------------- driver ---------------- */

void orc_set_counters(orc *o, int64_t step, int32_t colsteps, int32_t upd)
{
    o->step = step;
    o->colsteps = colsteps;
    o->upd = upd;
}

void orc_get_counters(const orc *o, int64_t *step, int32_t *colsteps, int32_t *upd)
{
    *step = o->step;
    *colsteps = o->colsteps;
    *upd = o->upd;
}

void orc_md_step(orc *o)
{
    /* one iteration of the time loop (maincode.cpp:461-764) without output */
    o->wallspeed = irbc_wallspeed(&o->p, o->step);
    o->colsteps += 1;
    o->upd += 1;
    if (o->colsteps == o->p.csteps) {
        orc_collision(o);
        o->colsteps = 0;
    }
    orc_kick1(o);
    if (o->upd == o->p.uplist) {
        orc_llistupdate(o);
        o->upd = 0;
    }
    orc_stream(o, o->wallspeed);
    orc_drift(o);
    orc_totalva(o);
    orc_pfdt(o);
    orc_forces(o);
    orc_external(o);
    orc_bond_prep(o);
    orc_bond_dynamics(o);
    orc_kick2(o);
    o->step += 1;
}

/* ---------------- README-run init ---------------- */

int orc_init_readme(orc *o, int32_t seed, int64_t ns, int nv, const double *xyz_raw, int ne,
                    const int32_t *edges, int nt, const int32_t *tris, int nknobs, double *stats)
{
    /* maincode.cpp:58-446 at one thread: RanGen[0] and sto[0] share the seed */
    orc_seed_mt(o, seed, seed);
    const irbc_params &P = o->p;
    if (nknobs > nv) return -1; /* the reference would loop forever (maincode.cpp:403) */
    /* solvent (:96-107) */
    double vmax = std::sqrt(3.0 * P.Theta / P.mass);
    o->x.resize(ns); o->y.resize(ns); o->z.resize(ns);
    o->vx.resize(ns); o->vy.resize(ns); o->vz.resize(ns);
    for (int64_t j = 0; j < ns; ++j) {
        o->x[j] = (o->Lx * o->ran.uniform());
        o->y[j] = (o->Ly * o->ran.uniform());
        o->z[j] = (o->Lz * o->ran.uniform());
        o->vx[j] = -vmax + 2.0 * vmax * o->ran.uniform();
        o->vy[j] = -vmax + 2.0 * vmax * o->ran.uniform();
        o->vz[j] = -vmax + 2.0 * vmax * o->ran.uniform();
    }
    o->pid.assign(ns, 0);
    o->indio.assign(ns, 1);
    /* vertices (:165-206) */
    vmax = std::sqrt(3.0 * P.Theta / (P.pmass));
    std::vector<double> xyz(xyz_raw, xyz_raw + 3 * nv), vel(3 * nv), zc(nv);
    for (int i = 0; i < nv; ++i) {
        zc[i] = xyz[3 * i + 2];
        vel[3 * i] = -vmax + 2.0 * vmax * o->ran.uniform();
        vel[3 * i + 1] = -vmax + 2.0 * vmax * o->ran.uniform();
        vel[3 * i + 2] = -vmax + 2.0 * vmax * o->ran.uniform();
    }
    std::sort(zc.begin(), zc.end());
    double Center[3];
    Center[2] = -zc.front() + 0.3;
    Center[0] = o->Lx / 2.0 + 10.0;
    Center[1] = o->Ly / 2.0;
    for (int i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] += Center[k];
    /* parasite (:212-225) */
    o->para[0] = Center[0] - 2.0;
    o->para[1] = Center[1];
    o->para[2] = Center[2];
    vmax = std::sqrt(3.0 * P.Theta / (P.PRmass));
    for (int k = 0; k < 3; ++k) o->paraV[k] = -vmax + 2.0 * vmax * o->ran.uniform();
    o->paradx = o->parady = 0;
    for (int k = 0; k < 3; ++k) o->paraft[k] = o->parafdt[k] = 0.0;
    /* mesh topology (:228-375) */
    orc_set_mesh(o, nv, xyz.data(), ne, edges, nt, tris);
    for (int i = 0; i < nv; ++i) {
        o->mvx[i] = vel[3 * i];
        o->mvy[i] = vel[3 * i + 1];
        o->mvz[i] = vel[3 * i + 2];
    }
    double me = 0, me2 = 0;
    for (int i = 0; i < ne; ++i) {
        me += o->l0[i];
        me2 += o->l0[i] * o->l0[i];
    }
    double shear = P.shear;
    double areaC = 2.0 * shear + P.kappasa + P.kappad;
    if (stats) {
        stats[0] = me / (double)ne;
        stats[1] = std::sqrt((me2 / (double)ne) - ((me / (double)ne) * (me / (double)ne)));
        stats[2] = o->Sarea0;
        stats[3] = o->Volume0;
        stats[4] = shear;
        stats[5] = (4.0 * areaC * shear) / (areaC + shear);
        stats[6] = (areaC - shear) / (areaC + shear);
    }
    /* (:383-391) */
    orc_centerofmass(o);
    orc_makeiedge(o);
    orc_totalva(o);
    orc_pfdt(o);
    orc_forces(o);
    orc_tricm(o);
    orc_sortsolvent(o);
    orc_llistupdate(o);
    /* knob placement (:394-433): area-weighted triangle, random corner; a
     * corner that is already a knob lets the scan continue to the next
     * triangle with a fresh corner draw */
    std::vector<double> parea(nt);
    for (int i = 0; i < nt; ++i) parea[i] = o->tarea[i] / o->Sarea0;
    int knob_count = 0;
    while (knob_count < nknobs) {
        double rad = o->ran.uniform();
        double aaa = 0;
        for (int j = 0; j < nt; ++j) {
            aaa += parea[j];
            if (aaa > rad) {
                int radi = o->ran.irandom(1, 3);
                if ((radi == 1) && (o->knob[o->t1[j]] == 0)) {
                    o->knob[o->t1[j]] = 1;
                    knob_count += 1;
                    break;
                } else if ((radi == 2) && (o->knob[o->t2[j]] == 0)) {
                    o->knob[o->t2[j]] = 1;
                    knob_count += 1;
                    break;
                } else if ((radi == 3) && (o->knob[o->t3[j]] == 0)) {
                    o->knob[o->t3[j]] = 1;
                    knob_count += 1;
                    break;
                }
            }
        }
    }
    int cnt = 0;
    for (int i = 0; i < nv; ++i) cnt += o->knob[i] == 1;
    if (stats) stats[7] = cnt;
    o->step = 0;
    o->colsteps = -1;
    o->upd = 0;
    return 0;
}

} /* extern "C" */
