/*
 * irbc_dev.h — device-side data layout shared by every kernel of the
 * MI355X iRBC step.
 *
 * A context holds a batch of replicas packed replica-major in one set of
 * device arrays.  DevRep is one replica's descriptor: its offsets into the
 * packed arrays, its parameters, and the dynamic per-step state that device
 * code advances (so a captured sequence of launches needs no host writes).
 *
 * Precision: solvent positions and velocities are FP32 in HBM and all
 * arithmetic on them is FP64 in registers; per-cell momentum sums are 64-bit
 * fixed point at 2^-40, order-independent (each term is rounded once, and
 * is exact for |v| >= 2^-17); the membrane, parasite and bonds are FP64
 * throughout.
 */
#ifndef IRBC_DEV_H
#define IRBC_DEV_H

#include <stdint.h>

#define IRBC_WG 256            /* particle and cell tile */
#define IRBC_MEM_WG 512        /* one membrane workgroup per replica (2 waves per SIMD: no spills) */
#define IRBC_SCAN_BLOCK 4096   /* elements per scan block (256 threads x 16) */
#define IRBC_MAX_CAND 16       /* ligand candidates per knob (3x3 window in practice) */
#define IRBC_TRI_KIN 18         /* doubles per triangle: 3 positions, 3 half-kicked velocities */
#define IRBC_LLIST_PER_TRI 64  /* a triangle lies within 2a of at most 4x4x4 cells */
#define IRBC_RANDNO 12345678   /* unbound-ligand sentinel (define.h:37) */
#define IRBC_MAX_REPLICAS 4096 /* per context: kernels flattened over a batch keep a prefix per replica in LDS */
#define IRBC_FIX_SCALE 1099511627776.0      /* 2^40: fixed-point unit of cell sums */
#define IRBC_FIX_INV (1.0 / 1099511627776.0)
#ifndef IRBC_SORT_EVERY_DEFAULT
#define IRBC_SORT_EVERY_DEFAULT 1 /* collisions per counting sort (IRBC_OPT_SORT_EVERY) */
#endif

/* status bits raised by device code, read by the host after a sync */
#define IRBC_ST_NAN 1
#define IRBC_ST_CAND_OVERFLOW 2
#define IRBC_ST_LLIST_OVERFLOW 4
#define IRBC_ST_RANGE 8
#define IRBC_ST_STALL 16       /* a bounded in-kernel wait ran out (the step's results are invalid) */
#define IRBC_ST_XCD 32         /* a replica's membrane-chain workgroups ran on two XCDs while handing
                                * off through one XCD's L2 (the step's results are invalid) */

/* counter-based stream ids (Philox4x32-10 counters: {index, stream, hi, lo}) */
#define IRBC_RS_SHIFT 1u
#define IRBC_RS_AXIS 2u
#define IRBC_RS_WALL 3u
#define IRBC_RS_MEM 4u
#define IRBC_RS_BOND 5u
#define IRBC_RS_INIT 6u     /* a fresh solvent's positions (irbc_init_solvent) */
#define IRBC_RS_INITV 7u    /* and velocities */

#define IRBC_SERIES_REC 8      /* doubles per time-series record */
/* lanes per vertex in the membrane chain's gather.  More lanes fetch a
 * vertex's terms in fewer round trips but add workgroups that hold wave
 * slots while they wait; measured on the 64x64x48 replica (one queue / two
 * queues with graphs / 64-replica batch, us per step): 1: 59.0 / 61.8 /
 * 378.5, 2: 59.9 / 60.6 / 382.2, 4: 62.1 / 59.1 / 391.5; again once the
 * streaming split weighted the XCDs by their chains (one queue / batch):
 * 1: 55.4 / 346.3, 2: 55.4 / 350.4, 4: 56.5 / 359.2. */
#ifndef IRBC_GATHER_LANES
#define IRBC_GATHER_LANES 1
#endif

/* vertex drift up to which the cells' near masks stand (see k_cell_mask) */
#define IRBC_DRIFT_MAX 0.25
/* distance budget of the near masks: a sub-cell is marked when it lies
 * within this of a listed triangle's sphere and slab; an unmarked
 * particle streams freely while its step travel plus the vertex drift
 * stays below it */
#define IRBC_NEAR_D0 0.0625

struct DevRep {
    /* ---- geometry and packing ---- */
    int32_t Lx, Ly, Lz, ncell;
    int64_t coff;      /* first cell in packed cell arrays */
    int64_t coff1;     /* first entry of per-replica (ncell+1) offset arrays = coff + r */
    int64_t poff;      /* first particle slot */
    int32_t np;        /* particles */
    int32_t tile0;     /* first particle tile of this replica */
    int32_t ctile0;    /* first cell tile */
    int32_t voff, nv, eoff, ne, toff, nt;
    int32_t loff, nlig, nligx, nligy, lxoff, lyoff;
    int64_t llent_off; /* first entry of this replica's cell->triangle entries */
    int64_t llent_cap;
    int32_t inc_off;   /* vertex-major incidence offsets at voff + r */
    int32_t inc_ent_off; /* element term -> its position in the vertex-major order */
    int64_t scr_off;   /* element terms (double3): term k of vertex v at k * nv + v */
    int32_t rep_index;
    int32_t part_off;  /* first per-block area/volume partial of this replica */
    /* ---- parameters ---- */
    double navg, mass, Theta, dtmd, cosalpha, sinalpha, walls, ramp_speed;
    int64_t ramp_steps;
    int32_t csteps, uplist;
    double kappasa, kappad, kappav, kappab, pmass, ks, dthr, FD, Fg, el0, elap;
    double PRmass, paraR, onrate, offrate, Pi;
    double cos_t0, sin_t0;  /* cos, sin of theta0*Pi/180 (host libm) */
    double pon;             /* 1 - exp(-onrate*dtmd) (host libm) */
    double Sarea0, Volume0, theta0;
    /* ---- dynamic state advanced on the device ---- */
    int64_t step;           /* index of the step about to run */
    int32_t colsteps, upd;  /* counters after this step's increment */
    int32_t collide, llist; /* this step runs a collision / a list update */
    int32_t rotate_pending; /* collision done, its rotation applied by the next streaming pass: 1 the
                             * sorted state is in the records (rec), written back in their order;
                             * 2 the particles stayed in their slots, keys in skeya */
    int32_t sort_now;       /* this step's collision sorts the particles into cell order (sorts_at) */
    int32_t sort_every;     /* collisions per sort (1: every collision sorts) */
    int32_t key_src;        /* where the last collision's keys are: 1 the records, 2 skeya */
    double wall2[2];        /* top-wall speed of this and the next step, by step parity (the
                             * membrane part of step s writes that of step s+1) */
    double shift[3];
    double Sarea, Volume;
    double para[3], paraV[3], paraft[3], parafdt[3];
    int32_t paradx, parady, paraid;
    int32_t tapes;          /* 1: streams come from the tape buffers */
    uint64_t seed;
    int32_t status;
    int32_t bond_draws;     /* draws consumed by the last bond step */
    int32_t nbonds;         /* bonded knobs after the last bond step */
    int32_t near_count[2];  /* particles queued for bounce-back, by step parity */
    int32_t defer_count[2]; /* of those, particles that cross a triangle plane */
    int32_t wca_count;      /* vertices inside the parasite's cutoff this step */
    /* near-mask drift bookkeeping: float bits of the largest vertex speed after
     * the first half-kick (a vertex moves that speed x dt in the step) of
     * this step and of the coming one; by step parity, the sum of those
     * moves over the steps since the near masks were made, before this
     * step, and the margin this step's streaming subtracts from them */
    uint32_t vcur_bits, vnext_bits;
    double reach_acc[2], reach_drift[2];
    int32_t draw_total;     /* bond draws needed this step */
    /* the one-launch membrane step: workgroup tickets of the running launch
     * and the workgroups of each phase that have finished (all zero between
     * launches: the finishing workgroup clears them) */
    int32_t mem_ticket;
    int32_t mem_done[4];
    uint32_t scan_epoch;    /* advanced before each of the replica's cell scans (see k_scan_lb) */
    int32_t align_lists;    /* cell->triangle lists at collision steps (lists_due) */
    int32_t chain_xcc;      /* XCD id + 1 of the running chain's workgroups (0: none yet) */
    int32_t chain_xcc_last; /* XCD id of the last finished chain (-1: none) */
    int32_t nlisted;        /* cells that list a triangle (compacted at each list update) */
    int32_t mask_keep;      /* the coming mask launch keeps this replica's masks (decided by the
                             * launch before it, so every mask workgroup sees one decision) */
    /* time series: every `series_every` steps one record of IRBC_SERIES_REC
     * doubles (step, centroid x y z unwrapped, centroid velocity, bonds) */
    int32_t series_every, series_cap, series_n, series_pad;
    double *series;
    /* ---- particle buffers (packed arrays, this replica at poff) ---- */
    float *xa, *ya, *za, *vxa, *vya, *vza;
    uint32_t *ida;          /* particle id | inside-flag << 31 */
    /* The last collision's counting sort: per slot, in cell order, a
     * 16-byte velocity record in one plane (vx vy vz, the shifted-cell key
     * as bits; all k_cell reads) and a 16-byte position record in the
     * other (x y z, the id as bits).
     * Between a collision and the streaming pass after it the state lives
     * here; after that pass the SoA arrays hold it in the same order, and
     * the records' keys stay the particles' last cell keys. */
    float *rec, *rec1;
    /* A collision that does not sort leaves the particles in their slots
     * and writes each one's shifted-cell key here, slot by slot. */
    int32_t *skeya;
    /* Beside each bounce-back queue entry, by queue position, the particle as
     * the streaming pass saw it: two 16-byte words (x y z vx) (vy vz, slot,
     * cell | inside flag << 31), so the sphere pass reads them coalesced */
    float *nrec;
    int32_t skey_pad[2];
};

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

/* Does this step (upd_after steps since the last list update, after the
 * top-of-step increment) update the cell->triangle lists?  The reference's
 * cadence: every uplist steps (maincode.cpp:460, 501-505).  Aligned: also
 * at the last collision step before that deadline, so that list updates
 * (and the near masks they rebuild) fall on collision steps; a batch whose
 * replicas have different uplist then runs its list work only beside its
 * collisions.  A list built sooner holds every triangle the reference's
 * older list holds that a particle can still reach within the deadline, so
 * the hits are the same while the membrane moves less than the lists'
 * margin between updates (the reference's own assumption). */
__host__ __device__ inline bool lists_due(int32_t upd_after, bool collide, int32_t csteps, int32_t uplist, int align)
{
    return upd_after == uplist || (align && collide && upd_after + csteps > uplist);
}

/* Does the collision of step `step` (a collision step) sort the particles
 * into cell order?  Every sort_every-th collision does: consecutive
 * collisions lie csteps apart, so step / csteps advances by one between
 * them.  The results do not depend on it (the cell sums are fixed point and
 * every particle's arithmetic is its own); the sort keeps a tile's particles
 * in few cells, which the collision's LDS aggregation and the streaming
 * pass's cell-table reads rely on for their speed. */
__host__ __device__ inline bool sorts_at(int64_t step, int32_t csteps, int32_t every)
{
    return every <= 1 || (step / (csteps > 0 ? csteps : 1)) % every == 0;
}

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
struct u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        u4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

/* two 53-bit uniforms in [0,1) from one Philox block */
__device__ __forceinline__ void philox_u2(uint64_t seed, uint32_t stream, uint64_t index,
                                          uint32_t sub, double &u0, double &u1)
{
    u4 c = {(uint32_t)index, stream, (uint32_t)(index >> 32), sub};
    u4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)r.x << 32 | r.y) >> 11;
    uint64_t b = ((uint64_t)r.z << 32 | r.w) >> 11;
    u0 = (double)a * (1.0 / 9007199254740992.0);
    u1 = (double)b * (1.0 / 9007199254740992.0);
}

/* two standard normals (Box-Muller) from one Philox block */
__device__ __forceinline__ void philox_n2(uint64_t seed, uint32_t stream, uint64_t index,
                                          uint32_t sub, double &z0, double &z1)
{
    double u0, u1;
    philox_u2(seed, stream, index, sub, u0, u1);
    double r = sqrt(-2.0 * log(1.0 - u0)); /* 1-u0 in (0,1] */
    double t = 6.283185307179586 * u1;
    z0 = r * cos(t);
    z1 = r * sin(t);
}

/* the reference's bbound() while-loop wrap (Src/solfuncs.cpp:603-622); a
 * coordinate more than 64 periods away (or not finite) cannot come from a
 * valid state and maps to 0 so that the loop stays bounded */
__device__ __forceinline__ double wrap_loop(double x, double l)
{
    if (!(x >= -64.0 * l && x <= 64.0 * l)) return 0.0;
    if (x > l)
        while (x > l) x -= l;
    if (x < 0)
        while (x < 0) x += l;
    return x;
}

/* cindex[][][] lookup (maincode.cpp:112-136, define.h:330-334): the
 * never-written slab at i = Lx, j = Ly or k = Lz+1 reads as cell 0 */
__device__ __forceinline__ int32_t cell_of(int fx, int fy, int fz, int Lx, int Ly, int Lz)
{
    if (fx >= 0 && fx < Lx && fy >= 0 && fy < Ly && fz >= 0 && fz <= Lz)
        return (fx * Ly + fy) * (Lz + 1) + fz;
    return 0;
}

/* Workgroups are dealt to the 8 XCDs round-robin by launch index.  A grid
 * of xcd_grid(n) blocks mapped through xcd_tile hands each XCD one
 * contiguous eighth of the n logical tiles, so writes that land in the same
 * cache lines (neighbouring cells of a sort) come from one XCD's L2.
 * Blocks with xcd_tile(n) >= n do nothing.  With chunked = 0 the block
 * index is the tile (a grid of exactly n blocks). */
/* the membrane arrays as the kernels see them, for one parity of the
 * double-buffered kinematics */
struct MemDev {
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *poff, *pon;
    int32_t *ix, *iy, *pid, *bondi, *bondl, *knob;
    const int32_t *e1, *e2, *es1, *es2, *iedge, *t1, *t2, *t3;
    const double *l0, *kappap, *tarea;
    double *tri_area, *wca, *slot, *upos, *part, *tri_kin;
    double *tag;       /* per slot: NaN (a finished term), a triangle's area (the global-area term's vector
                        * C, to be scaled by beta / (4 area)) or -1 (the volume term's vector w, to be
                        * scaled as betav w / 6) */
    int64_t upos_base; /* vertex g's unwrapped position at upos[4 (g - upos_base)] */
    double *vkin;      /* per vertex: position, velocity after the coming first half-kick */
    float4 *tri_sph;  /* per triangle: bounding-sphere centre and reach for the coming step */
    const int32_t *inc_off, *inc_ent;
    int32_t *ncand, *cand, *draw_off, *flag, *list;
    int32_t *ligi;
    const double *ligx, *ligy;
    const double *bond_u;
};

#define IRBC_NXCD 8
__device__ __forceinline__ int xcd_tile_of(int n, int chunked, int b)
{
    if (!chunked) return b;
    const int per = (n + IRBC_NXCD - 1) / IRBC_NXCD;
    return (b % IRBC_NXCD) * per + b / IRBC_NXCD;
}

__device__ __forceinline__ int xcd_tile(int n, int chunked) { return xcd_tile_of(n, chunked, (int)blockIdx.x); }

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = (a[1] * b[2]) - (b[1] * a[2]);
    c[1] = (a[2] * b[0]) - (b[2] * a[0]);
    c[2] = (a[0] * b[1]) - (b[0] * a[1]);
}

#endif /* __HIPCC__ */

#endif /* IRBC_DEV_H */
