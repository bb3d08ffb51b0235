/*
 * irbc.hip — host side of the C ABI (include/irbc.h) and the kernels it
 * launches, built into libirbc_amd.so for gfx950.
 *
 * Layout in HBM (packed replica-major; each replica's particle slots start
 * on a tile boundary so every workgroup belongs to exactly one replica):
 *   solvent   x y z vx vy vz FP32 SoA, double-buffered for the counting sort,
 *             + id|inside-flag u32, shifted key and rank i32      ~40 B/slot
 *   cells     count/start i32, SRD table (vcm, axis, scale) 64 B, vertex list
 *             head, cell->triangle CSR                            ~80 B/cell
 *   membrane  FP64 vertex state, topology, per-element force slots
 *   ligands   occupancy i32 and grid coordinate tables
 */
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <string>
#include <map>
#include <numeric>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "../../include/irbc.h"
#include "irbc_dev.h"

#include "k_solvent.inc"
#include "k_membrane.inc"
#include "k_step.inc"

namespace {

thread_local std::string g_err;
/* Graph captures and stream-ordered allocations of any context on any
 * thread take turns: a pool allocation or free made while another thread
 * captures would invalidate that capture. */
static std::mutex g_capture_mu;
/* Library calls and graph captures: a capture (inside irbc_md_steps) holds
 * this exclusively and every other entry point that touches the device
 * holds it shared, so no call made on another thread (a request loop sets
 * up contexts on its own threads) runs during a capture.  Nested entry
 * points take it once per thread. */
static std::shared_mutex g_api_mu;
static thread_local int g_api_depth = 0;
struct ApiGuard {
    bool held;
    ApiGuard() : held(g_api_depth++ == 0)
    {
        if (held) g_api_mu.lock_shared();
    }
    ~ApiGuard()
    {
        if (held) g_api_mu.unlock_shared();
        --g_api_depth;
    }
};
#define API_GUARD ApiGuard api_guard_

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(IRBC_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

#define LAUNCHCHK()                                                                       \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return fail(IRBC_ERR_HIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

struct HostRep {
    irbc_replica_spec spec;
    int64_t step = 0;
    int32_t colsteps = -1, upd = 0; /* reference counters before the top-of-step increment */
    bool mesh = false, solvent = false;
    std::vector<int32_t> e1, e2, t1, t2, t3;
};

struct Sizes {
    int64_t slots = 0, cells = 0, cells1 = 0, llent = 0, scr = 0;
    int32_t nv = 0, ne = 0, nt = 0, nlig = 0, nligx = 0, nligy = 0, inc = 0, part = 0;
    int32_t maxnv = 0, maxne = 0, maxnt = 0;
    int32_t tiles = 0, ctiles = 0, scanblk = 0;
};

int32_t lig_count(double L, double sp)
{
    int32_t n = 0;
    double m = 0.0;
    while (m < L) {
        ++n;
        m += sp;
    }
    return n;
}

/* upper bound of vertex incidences: 2 per edge end, 3 per triangle corner, 1 per hinge corner */
int32_t inc_bound(const irbc_replica_spec &s) { return 4 * s.n_edges + 9 * s.n_triangles + 4 * s.n_edges; }
/* Element terms are stored vertex-interleaved: term k of vertex v at
 * k * nv + v, so that the lanes of a wave, one vertex each, read adjacent
 * terms together.  The planes run to the largest per-vertex count, held to
 * twice the mean (a valence up to about twice the mean valence). */
int64_t scr_bound(const irbc_replica_spec &s) { return 2 * (int64_t)inc_bound(s); }

Sizes plan(const irbc_replica_spec *specs, int n, std::vector<DevRep> *reps,
           std::vector<ScanBlk> *blks, std::vector<int32_t> *rep_first_blk,
           std::vector<int32_t> *tile_rep, std::vector<int32_t> *ctile_rep)
{
    Sizes z;
    for (int r = 0; r < n; ++r) {
        const irbc_replica_spec &s = specs[r];
        const irbc_params &p = s.params;
        DevRep R;
        std::memset(&R, 0, sizeof(R));
        R.Lx = p.Lx;
        R.Ly = p.Ly;
        R.Lz = p.Lz;
        R.ncell = (int32_t)irbc_ncells(&p);
        R.coff = z.cells;
        R.coff1 = z.cells1;
        R.tile0 = z.tiles;
        R.poff = (int64_t)z.tiles * IRBC_WG;
        R.np = (int32_t)s.n_solvent;
        R.ctile0 = z.ctiles;
        R.voff = z.nv;
        R.nv = s.n_vertices;
        R.eoff = z.ne;
        R.ne = s.n_edges;
        R.toff = z.nt;
        R.nt = s.n_triangles;
        R.nligx = lig_count(p.Lx, p.lig_spacing);
        R.nligy = lig_count(p.Ly, p.lig_spacing);
        R.nlig = R.nligx * R.nligy;
        R.loff = z.nlig;
        R.lxoff = z.nligx;
        R.lyoff = z.nligy;
        R.llent_off = z.llent;
        R.llent_cap = (int64_t)IRBC_LLIST_PER_TRI * s.n_triangles;
        R.inc_off = z.nv + r;
        R.inc_ent_off = z.inc;
        R.scr_off = z.scr;
        R.rep_index = r;
        R.chain_xcc_last = -1;
        R.part_off = z.part;
        const int32_t tiles = (int32_t)((s.n_solvent + IRBC_WG - 1) / IRBC_WG);
        const int32_t ctiles = (R.ncell + IRBC_WG - 1) / IRBC_WG;
        if (tile_rep) tile_rep->insert(tile_rep->end(), tiles, r);
        if (ctile_rep) ctile_rep->insert(ctile_rep->end(), ctiles, r);
        const int32_t nb = (R.ncell + IRBC_SCAN_BLOCK - 1) / IRBC_SCAN_BLOCK;
        if (rep_first_blk) rep_first_blk->push_back(z.scanblk);
        for (int b = 0; b < nb && blks; ++b) {
            ScanBlk B;
            B.rep = r;
            B.first = b * IRBC_SCAN_BLOCK;
            B.count = std::min(IRBC_SCAN_BLOCK, R.ncell - B.first);
            B.bidx = b;
            B.bfirst = z.scanblk;
            B.nblk = nb;
            blks->push_back(B);
        }
        z.scanblk += nb;
        z.tiles += tiles;
        z.ctiles += ctiles;
        z.slots = (int64_t)z.tiles * IRBC_WG;
        z.cells += R.ncell;
        z.cells1 += R.ncell + 1;
        z.nv += s.n_vertices;
        z.ne += s.n_edges;
        z.nt += s.n_triangles;
        z.nlig += R.nlig;
        z.nligx += R.nligx;
        z.nligy += R.nligy;
        z.llent += R.llent_cap;
        z.scr += scr_bound(s);
        z.inc += inc_bound(s);
        z.part += (std::max(s.n_edges, s.n_triangles) + IRBC_WG - 1) / IRBC_WG;
        z.maxnv = std::max(z.maxnv, s.n_vertices);
        z.maxne = std::max(z.maxne, s.n_edges);
        z.maxnt = std::max(z.maxnt, s.n_triangles);
        if (reps) reps->push_back(R);
    }
    if (rep_first_blk) rep_first_blk->push_back(z.scanblk);
    return z;
}

int64_t bytes_of(const Sizes &z, int n)
{
    int64_t b = 0;
    b += z.slots * (6 * 4 + 4 + 32 + 3 * 4 + 32);         /* particles, ids, sort records, keys, ranks, queue, its records */
    b += z.cells * (4 + 8 * 8 + 4 + 4 + 8 * 8 + 4 + 8 + 4 + 32); /* count, ctab, vhead, llcount, diag, cursor, near mask, listed, sums */
    b += z.cells * 8 * 8;                                 /* tapes */
    b += z.cells1 * (4 + 4);                              /* start, lloff */
    b += z.llent * (4 + 8);                                /* entries and their sub-cell masks */
    b += (int64_t)z.nv * (31 * 8 + 12 * 4 + IRBC_MAX_CAND * (4 + 8)); /* incl. vertex kinematics x2 */
    b += 2 * (int64_t)z.nt * (IRBC_TRI_KIN * 8 + 16) + (int64_t)z.part * 16 + z.slots * 8;
    b += (int64_t)z.ne * (5 * 4 + 2 * 8) + (int64_t)z.nt * (3 * 4 + 2 * 8);
    b += z.scr * 32 + (int64_t)z.inc * 4 + (int64_t)(z.nv + n) * 4;
    b += (int64_t)z.nlig * 4 + (int64_t)(z.nligx + z.nligy) * 8;
    b += (int64_t)z.scanblk * 8 + 4 + (int64_t)n * sizeof(DevRep);
    return b;
}

} // namespace

struct irbc_ctx {
    int device = 0;
    hipStream_t st = nullptr;   /* solvent chain */
    hipStream_t st2 = nullptr;  /* membrane chain, overlapped with the solvent's streaming */
    hipEvent_t ev_pre = nullptr, ev_mem2[2] = {nullptr, nullptr}, ev_str[2] = {nullptr, nullptr};
    int par = 0;                /* parity of the current step: picks the double-buffered state */
    int mem_fused = 0;          /* membrane step as one workgroup per replica (1) or five grid kernels (0) */
    int mem_lds_nv = 0;         /* largest mesh whose unwrapped positions the one-workgroup step keeps in LDS */
    int use_graphs = 1;         /* irbc_md_steps replays captured chunks of IRBC_GRAPH_STEPS steps */
    int near_merged = 0;        /* near-membrane streaming in one kernel (1) or sphere pass + hit pass (0); set by irbc_create */
    int xcd_chunk = 1;          /* sort and cell kernels hand each XCD a contiguous eighth of the tiles */
    int stream4 = 0;            /* bulk streaming with four particles per lane */
    int stream_grid = 0;        /* bulk streaming on a capped grid (IRBC_STREAM_WGS_PER_CU per CU) */
    int mem_after = 0;          /* the membrane of a step starts after that step's bulk streaming */
    int mem_chain = 1;          /* the membrane step as one launch whose workgroups take their phase by ticket */
    int chain_dev = 0;          /* its hand-offs at device scope (1) or XCD-local (0) */
    int near_masks = 1;         /* sub-cell near masks (1) or every particle of a listed cell queued (0) */
    int one_queue = 1;          /* with the chain: the step on one queue, the membrane inside the solvent's launches */
    int classify_all = 0;       /* classification tests every triangle (1) or the binned ones (0) */
    int align_lists = 0;        /* list updates at collision steps (lists_due) */
    int sort_every = IRBC_SORT_EVERY_DEFAULT; /* collisions per counting sort (sorts_at) */
    int bin_span = 1;           /* particle tiles per binning workgroup of a sorting collision (bin_span) */
    double span_ms[3] = {NAN, NAN, NAN}; /* the autotuner's ms per step with bin spans 1, 2, 4 */
    int ncu = 256;              /* compute units of the device */
    /* the device state is as the last irbc_md_steps left it (vertex and
     * triangle kinematics, near masks, drift margins, the next step's grid
     * shift), so the next call needs no prologue; any call that changes
     * state clears it */
    bool primed = false;
    std::map<std::pair<uint64_t, uint64_t>, hipGraphExec_t> graphs;
    int nrep = 0;
    std::vector<HostRep> hr;
    std::vector<DevRep> rep;
    Sizes z;
    int64_t bytes = 0;
    std::vector<void *> allocs;
    std::vector<size_t> alloc_bytes; /* size of each entry of allocs */

    DevRep *d_rep = nullptr;
    int32_t *d_tile_rep = nullptr, *d_ctile_rep = nullptr, *d_rep_first_blk = nullptr;
    ScanBlk *d_blk = nullptr;
    unsigned long long *scanflag = nullptr; /* published sums and prefixes of the look-back scan */
    int32_t *scanticket = nullptr;
    /* solvent */
    float *xa, *ya, *za, *vxa, *vya, *vza;
    uint32_t *ida;
    float *rec; /* the collision's sorted records, 8 floats per slot */
    int32_t *skey, *srank, *nearq, *deferq;
    float4 *nrec; /* queue records beside nearq, 2 per slot (DevRep::nrec) */
    unsigned long long *csum; /* per cell: fixed-point momentum x y z and dvv of an in-place collision */
    /* cells */
    int32_t *ccount, *cstart, *vhead, *lloff, *llcount, *llent;
    double *ctab;
    uint64_t *cmask;
    uint64_t *llmask; /* per cell->triangle entry: the cell's sub-cells that triangle can reach (k_cell_mask) */
    int32_t *listed;  /* per replica: the cells that list a triangle, compacted */
    CellDiag diag;
    double *cell_u, *wall_z, *mem_z;
    /* membrane */
    double *mx, *my, *mz, *mvx, *mvy, *mvz, *mfx, *mfy, *mfz, *poff, *pon;
    int32_t *mix, *miy, *mpid, *bondi, *bondl, *knob, *vnext;
    int32_t *e1, *e2, *es1, *es2, *iedge, *t1, *t2, *t3;
    double *l0, *kappap, *tarea, *tri_area, *wca, *slot, *tag, *upos, *part, *tri_kin, *vkin;
    float4 *tri_sph;
    int32_t *inc_off, *inc_ent;
    int32_t *ncand, *cand, *draw_off, *flag, *list;
    int32_t *ligi;
    double *ligx, *ligy, *bond_u;
    /* staging for downloads */
    float *stage_f;
    int32_t *stage_i;
    bool diag_on = false;

    /* profiling: HIP events around every prof_every-th step's launches */
    bool prof = false;
    int prof_every = 0;
    int64_t prof_step = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pend[4];
    double acc_ms[4] = {0, 0, 0, 0};
    int64_t acc_n[4] = {0, 0, 0, 0};

    template <class T> int alloc(T *&p, int64_t n)
    {
        p = nullptr;
        if (n <= 0) n = 1;
        void *q = nullptr;
        /* stream-ordered: a context freed while others run waits for its
         * own stream only (hipFree would wait for the whole device) */
        hipError_t e;
        {
            std::lock_guard<std::mutex> lock(g_capture_mu);
            e = hipMallocAsync(&q, (size_t)n * sizeof(T), st);
        }
        if (e != hipSuccess)
            return fail(IRBC_ERR_OOM, "hipMallocAsync of %lld bytes failed", (long long)(n * sizeof(T)));
        if (hipMemsetAsync(q, 0, (size_t)n * sizeof(T), st) != hipSuccess)
            return fail(IRBC_ERR_HIP, "hipMemsetAsync failed");
        allocs.push_back(q);
        alloc_bytes.push_back((size_t)n * sizeof(T));
        bytes += n * (int64_t)sizeof(T);
        p = (T *)q;
        return IRBC_OK;
    }

    int64_t tri_stride() const { return (int64_t)IRBC_TRI_KIN * z.nt; }
    MemDev mem(int tbuf) const
    {
        MemDev M;
        M.x = mx; M.y = my; M.z = mz; M.vx = mvx; M.vy = mvy; M.vz = mvz;
        M.fx = mfx; M.fy = mfy; M.fz = mfz; M.poff = poff; M.pon = pon;
        M.ix = mix; M.iy = miy; M.pid = mpid; M.bondi = bondi; M.bondl = bondl; M.knob = knob;
        M.e1 = e1; M.e2 = e2; M.es1 = es1; M.es2 = es2; M.iedge = iedge;
        M.t1 = t1; M.t2 = t2; M.t3 = t3;
        M.l0 = l0; M.kappap = kappap; M.tarea = tarea;
        M.tri_area = tri_area; M.wca = wca; M.slot = slot; M.tag = tag;
        M.upos = upos; M.upos_base = 0; M.part = part; M.tri_kin = tri_kin + (tbuf & 1) * tri_stride();
        M.tri_sph = tri_sph + (tbuf & 1) * z.nt;
        M.vkin = vkin + (tbuf & 1) * 6 * (int64_t)z.nv;
        M.inc_off = inc_off; M.inc_ent = inc_ent;
        M.ncand = ncand; M.cand = cand; M.draw_off = draw_off; M.flag = flag; M.list = list;
        M.ligi = ligi; M.ligx = ligx; M.ligy = ligy; M.bond_u = bond_u;
        return M;
    }
    MemView view(int tbuf) const
    {
        MemView V;
        V.x = mx; V.y = my; V.z = mz; V.vx = mvx; V.vy = mvy; V.vz = mvz;
        V.fx = mfx; V.fy = mfy; V.fz = mfz; V.t1 = t1; V.t2 = t2; V.t3 = t3;
        V.tri_kin = tri_kin + (tbuf & 1) * tri_stride();
        return V;
    }
    hipError_t sync_both() const
    {
        hipError_t e = hipStreamSynchronize(st);
        if (e == hipSuccess && st2) e = hipStreamSynchronize(st2);
        return e;
    }
    int wait_all()
    {
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipStreamSynchronize(st2));
        return IRBC_OK;
    }
    int pull(int r)
    {
        int rc_ = wait_all();
        if (rc_) return rc_;
        HIPCHK(hipMemcpy(&rep[r], d_rep + r, sizeof(DevRep), hipMemcpyDeviceToHost));
        return IRBC_OK;
    }
    int push(int r)
    {
        primed = false;
        HIPCHK(hipMemcpyAsync(d_rep + r, &rep[r], sizeof(DevRep), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return IRBC_OK;
    }
    hipEvent_t event()
    {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        /* timing only: no system-scope fence, so the event brackets the
         * kernel and not a cache flush as well */
        hipEvent_t e;
        hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
        return e;
    }
    int flush_prof()
    {
        int rc_ = wait_all();
        if (rc_) return rc_;
        for (int k = 0; k < 4; ++k) {
            for (auto &pr : pend[k]) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
                acc_ms[k] += ms;
                acc_n[k] += 1;
                ev_pool.push_back(pr.first);
                ev_pool.push_back(pr.second);
            }
            pend[k].clear();
        }
        return IRBC_OK;
    }
};

/* top-of-step increment of the time loop (maincode.cpp:463-484, 501-505) */
static void top_of_step(const irbc_params &p, int64_t step, int32_t colsteps_pre, int32_t upd_pre,
                        DevRep &R)
{
    R.step = step;
    R.wall2[0] = R.wall2[1] = irbc_wallspeed(&p, step);
    R.colsteps = colsteps_pre + 1;
    R.upd = upd_pre + 1;
    R.collide = R.colsteps == p.csteps;
    if (R.collide) R.colsteps = 0;
    R.sort_now = R.collide && sorts_at(step, p.csteps, R.sort_every);
    R.llist = lists_due(R.upd, R.collide != 0, p.csteps, p.uplist, R.align_lists);
    if (R.llist) R.upd = 0;
}

static void advance_host(const irbc_params &p, HostRep &h, int align)
{
    int32_t c = h.colsteps + 1, u = h.upd + 1;
    const bool col = c == p.csteps;
    if (col) c = 0;
    if (lists_due(u, col, p.csteps, p.uplist, align)) u = 0;
    h.colsteps = c;
    h.upd = u;
    h.step += 1;
}

/* the same parameters, field by field (the padding after the five
 * leading ints is not compared) */
static bool same_params(const irbc_params &a, const irbc_params &b)
{
    if (a.Lx != b.Lx || a.Ly != b.Ly || a.Lz != b.Lz || a.csteps != b.csteps || a.uplist != b.uplist) return false;
    const size_t o = offsetof(irbc_params, navg);
    return std::memcmp((const char *)&a + o, (const char *)&b + o, sizeof(irbc_params) - o) == 0;
}

/* time-loop counters as the loop can hold them before its top-of-step
 * increment */
static bool counters_valid(const irbc_params &p, int64_t step, int32_t colsteps, int32_t upd)
{
    return step >= 0 && colsteps >= -1 && colsteps < p.csteps && upd >= -1 && upd < p.uplist;
}

static bool host_collides(const irbc_params &p, const HostRep &h) { return h.colsteps + 1 == p.csteps; }
static bool host_llist(const irbc_params &p, const HostRep &h, int align)
{
    return lists_due(h.upd + 1, h.colsteps + 1 == p.csteps, p.csteps, p.uplist, align);
}

extern "C" {

const char *irbc_last_error(void) { return g_err.c_str(); }

int64_t irbc_estimate_bytes(const irbc_replica_spec *specs, int n)
{
    if (!specs || n <= 0) return -1;
    Sizes z = plan(specs, n, nullptr, nullptr, nullptr, nullptr, nullptr);
    return bytes_of(z, n);
}

int irbc_create(const irbc_replica_spec *specs, int n, int device, irbc_ctx **out)
{
    API_GUARD;
    if (!specs || n <= 0 || !out) return fail(IRBC_ERR_ARG, "irbc_create: bad arguments");
    if (n > IRBC_MAX_REPLICAS) return fail(IRBC_ERR_ARG, "irbc_create: at most %d replicas per context", IRBC_MAX_REPLICAS);
    *out = nullptr;
    for (int r = 0; r < n; ++r) {
        const irbc_replica_spec &s = specs[r];
        const irbc_params &p = s.params;
        if (p.Lx < 8 || p.Ly < 8 || p.Lz < 4 || s.n_solvent <= 0 || s.n_solvent >= (1ll << 31) ||
            s.n_vertices < 4 || s.n_edges < 6 || s.n_triangles < 4 || p.csteps < 1 || p.uplist < 1 ||
            p.dtmd <= 0 || p.lig_spacing <= 0)
            return fail(IRBC_ERR_ARG, "irbc_create: replica %d has invalid sizes or parameters", r);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(IRBC_ERR_NODEVICE, "irbc_create: no HIP device");
    if (device < 0 || device >= ndev) return fail(IRBC_ERR_ARG, "irbc_create: no device %d", device);
    HIPCHK(hipSetDevice(device));
    {
        /* contexts come and go in a request loop: the device's pool keeps
         * freed memory for the next one instead of returning it at each
         * synchronisation */
        std::lock_guard<std::mutex> lock(g_capture_mu);
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    irbc_ctx *c = new irbc_ctx();
    c->device = device;
    c->nrep = n;
    /* the one-pass bounce-back suits a few replicas (the single 64x64x48
     * replica: 60 against 71 us per step), the two passes a batch (64
     * replicas: 359 against 439) */
    c->near_merged = n < IRBC_NXCD ? 1 : 0;
    {
        /* dynamic LDS of the one-workgroup membrane step: 32 bytes a vertex,
         * within 64 KiB per workgroup less 16 KiB for its static arrays */
        int lim = 65536;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lim = 65536;
        c->mem_lds_nv = std::max(0, (std::min(lim, 65536) - 16384) / 32);
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
            c->ncu = ncu;
    }
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_pre, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_mem2[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_mem2[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_str[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_str[1], hipEventDisableTiming) != hipSuccess) {
        irbc_destroy(c);
        return fail(IRBC_ERR_HIP, "stream or event creation failed");
    }
    std::vector<ScanBlk> blks;
    std::vector<int32_t> rfb, trep, ctrep;
    c->z = plan(specs, n, &c->rep, &blks, &rfb, &trep, &ctrep);
    c->hr.resize(n);
    const Sizes &z = c->z;
    int rc = IRBC_OK;
#define A(p, n_)                                \
    if (rc == IRBC_OK) rc = c->alloc(p, (n_));
    A(c->d_rep, n) A(c->d_tile_rep, z.tiles) A(c->d_ctile_rep, z.ctiles)
    A(c->d_rep_first_blk, n + 1) A(c->d_blk, z.scanblk) A(c->scanflag, z.scanblk)
    A(c->scanticket, 1)
    A(c->xa, z.slots) A(c->ya, z.slots) A(c->za, z.slots) A(c->vxa, z.slots) A(c->vya, z.slots) A(c->vza, z.slots)
    A(c->rec, 8 * z.slots)
    A(c->ida, z.slots) A(c->skey, z.slots) A(c->srank, z.slots) A(c->nearq, z.slots)
    A(c->deferq, 2 * z.slots) A(c->nrec, 2 * z.slots)
    A(c->ccount, z.cells) A(c->csum, 4 * z.cells) A(c->cstart, z.cells1) A(c->vhead, z.cells) A(c->lloff, z.cells1)
    A(c->llcount, z.cells) A(c->llent, z.llent) A(c->llmask, z.llent) A(c->ctab, z.cells * 8) A(c->cmask, z.cells) A(c->listed, z.cells)
    A(c->diag.vcm, 3 * z.cells) A(c->diag.ncc, z.cells) A(c->diag.rand1, z.cells)
    A(c->diag.rand2, z.cells) A(c->diag.dvv, z.cells) A(c->diag.the, z.cells) A(c->diag.nc, z.cells)
    A(c->cell_u, 2 * z.cells) A(c->wall_z, 3 * z.cells) A(c->mem_z, 3 * z.cells)
    A(c->mx, z.nv) A(c->my, z.nv) A(c->mz, z.nv) A(c->mvx, z.nv) A(c->mvy, z.nv) A(c->mvz, z.nv)
    A(c->mfx, z.nv) A(c->mfy, z.nv) A(c->mfz, z.nv) A(c->poff, z.nv) A(c->pon, z.nv)
    A(c->mix, z.nv) A(c->miy, z.nv) A(c->mpid, z.nv) A(c->bondi, z.nv) A(c->bondl, z.nv)
    A(c->knob, z.nv) A(c->vnext, z.nv)
    A(c->e1, z.ne) A(c->e2, z.ne) A(c->es1, z.ne) A(c->es2, z.ne) A(c->iedge, z.ne)
    A(c->t1, z.nt) A(c->t2, z.nt) A(c->t3, z.nt)
    A(c->l0, z.ne) A(c->kappap, z.ne) A(c->tarea, z.nt) A(c->tri_area, z.nt) A(c->wca, 3 * z.nv)
    A(c->slot, 3 * z.scr) A(c->tag, z.scr) A(c->inc_off, z.nv + n) A(c->inc_ent, z.inc)
    A(c->upos, 4 * (int64_t)z.nv) A(c->part, 2 * (int64_t)z.part) A(c->tri_kin, 2 * (int64_t)IRBC_TRI_KIN * z.nt) A(c->tri_sph, 2 * (int64_t)z.nt) A(c->vkin, 12 * (int64_t)z.nv)
    A(c->ncand, z.nv) A(c->cand, (int64_t)z.nv * IRBC_MAX_CAND) A(c->draw_off, z.nv)
    A(c->flag, z.nv) A(c->list, z.nv)
    A(c->ligi, z.nlig) A(c->ligx, z.nligx) A(c->ligy, z.nligy)
    A(c->bond_u, (int64_t)z.nv * IRBC_MAX_CAND)
#undef A
    int64_t maxnp = 0;
    int32_t maxnv = 0;
    for (int r = 0; r < n; ++r) {
        maxnp = std::max<int64_t>(maxnp, specs[r].n_solvent);
        maxnv = std::max<int32_t>(maxnv, specs[r].n_vertices);
    }
    (void)maxnv;
    if (rc == IRBC_OK) rc = c->alloc(c->stage_f, 6 * maxnp);
    if (rc == IRBC_OK) rc = c->alloc(c->stage_i, 2 * maxnp);
    if (rc != IRBC_OK) {
        std::string keep = g_err;
        irbc_destroy(c);
        g_err = keep;
        return rc;
    }
    /* vertex-list heads start empty */
    hipMemsetAsync(c->vhead, 0xFF, (size_t)z.cells * 4, c->st);
    std::vector<double> lx, ly;
    for (int r = 0; r < n; ++r) {
        const irbc_replica_spec &s = specs[r];
        const irbc_params &p = s.params;
        HostRep &h = c->hr[r];
        h.spec = s;
        DevRep &R = c->rep[r];
        R.navg = p.navg; R.mass = p.mass; R.Theta = p.Theta; R.dtmd = p.dtmd;
        R.cosalpha = p.cosalpha; R.sinalpha = p.sinalpha; R.walls = p.walls;
        R.ramp_speed = p.ramp_speed; R.ramp_steps = p.ramp_steps;
        R.csteps = p.csteps; R.uplist = p.uplist;
        R.kappasa = p.kappasa; R.kappad = p.kappad; R.kappav = p.kappav; R.kappab = p.kappab;
        R.pmass = p.pmass; R.ks = p.ks; R.dthr = p.dthr; R.FD = p.FD; R.Fg = p.Fg; R.el0 = p.el0;
        R.elap = p.elap; R.PRmass = p.PRmass; R.paraR = p.paraR; R.onrate = p.onrate;
        R.offrate = p.offrate; R.Pi = p.Pi;
        R.pon = 1.0 - std::exp(-p.onrate * p.dtmd);
        R.seed = s.seed;
        R.xa = c->xa; R.ya = c->ya; R.za = c->za; R.vxa = c->vxa; R.vya = c->vya; R.vza = c->vza;
        R.ida = c->ida;
        R.rec = c->rec;
        R.rec1 = c->rec + 4 * c->z.slots;
        R.skeya = c->skey;
        R.nrec = (float *)c->nrec;
        R.sort_every = c->sort_every;
        R.key_src = 1;
        top_of_step(p, 0, -1, 0, R);
        /* ligand grid coordinates (maincode.cpp:77-93) */
        for (double m = 0.0; m < p.Lx; m += p.lig_spacing) lx.push_back(m);
        for (double m = 0.0; m < p.Ly; m += p.lig_spacing) ly.push_back(m);
    }
    HIPCHK(hipMemcpyAsync(c->d_rep, c->rep.data(), n * sizeof(DevRep), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_tile_rep, trep.data(), trep.size() * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_ctile_rep, ctrep.data(), ctrep.size() * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_rep_first_blk, rfb.data(), rfb.size() * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_blk, blks.data(), blks.size() * sizeof(ScanBlk), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->ligx, lx.data(), lx.size() * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->ligy, ly.data(), ly.size() * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(c->sync_both());
    *out = c;
    return IRBC_OK;
}

int irbc_destroy(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return IRBC_OK;
    if (c->st) hipStreamSynchronize(c->st);
    if (c->st2) hipStreamSynchronize(c->st2);
    for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    /* back to the device's pool in this context's stream order: nothing
     * else on the device is waited for */
    {
        std::lock_guard<std::mutex> lock(g_capture_mu);
        for (void *p : c->allocs) hipFreeAsync(p, c->st);
    }
    if (c->st) hipStreamSynchronize(c->st);
    for (int k = 0; k < 4; ++k)
        for (auto &pr : c->pend[k]) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_pre, c->ev_mem2[0], c->ev_mem2[1], c->ev_str[0], c->ev_str[1]})
        if (e) hipEventDestroy(e);
    if (c->st) hipStreamDestroy(c->st);
    if (c->st2) hipStreamDestroy(c->st2);
    delete c;
    return IRBC_OK;
}

int irbc_device_bytes(const irbc_ctx *c, int64_t *bytes)
{
    API_GUARD;
    if (!c || !bytes) return fail(IRBC_ERR_ARG, "irbc_device_bytes: bad arguments");
    *bytes = c->bytes;
    return IRBC_OK;
}

#define CHECK_R(c, r)                                                    \
    do {                                                                 \
        if (!(c) || (r) < 0 || (r) >= (c)->nrep)                         \
            return fail(IRBC_ERR_ARG, "%s: bad context or replica", __func__); \
    } while (0)

int irbc_sync(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "irbc_sync: null context");
    int rw = c->wait_all();
    if (rw) return rw;
    HIPCHK(hipMemcpy(c->rep.data(), c->d_rep, c->nrep * sizeof(DevRep), hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; ++r)
        if (c->rep[r].status & IRBC_ST_XCD) {
            /* the dispatch did not keep a replica's chain on one XCD: hand
             * off at device scope from now on */
            c->chain_dev = 1;
            return fail(IRBC_ERR_DEVICE,
                        "replica %d: membrane-chain workgroups ran on two XCDs (status 0x%x); the launched steps "
                        "are invalid, the context now hands off at device scope",
                        r, c->rep[r].status);
        }
    for (int r = 0; r < c->nrep; ++r)
        if (c->rep[r].status)
            return fail(IRBC_ERR_DEVICE, "replica %d raised device status 0x%x", r, c->rep[r].status);
    return IRBC_OK;
}

/* ---------------- scenario setup ---------------- */

int irbc_set_mesh(irbc_ctx *c, int r, const double *xyz, const int32_t *edges, const int32_t *tris)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!xyz || !edges || !tris) return fail(IRBC_ERR_ARG, "irbc_set_mesh: null array");
    HostRep &h = c->hr[r];
    const irbc_params &p = h.spec.params;
    const int nv = h.spec.n_vertices, ne = h.spec.n_edges, nt = h.spec.n_triangles;
    for (int i = 0; i < 2 * ne; ++i)
        if (edges[i] < 0 || edges[i] >= nv) return fail(IRBC_ERR_ARG, "irbc_set_mesh: edge index out of range");
    for (int i = 0; i < 3 * nt; ++i)
        if (tris[i] < 0 || tris[i] >= nv) return fail(IRBC_ERR_ARG, "irbc_set_mesh: triangle index out of range");
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    std::vector<double> x(nv), y(nv), zz(nv);
    for (int i = 0; i < nv; ++i) {
        x[i] = xyz[3 * i];
        y[i] = xyz[3 * i + 1];
        zz[i] = xyz[3 * i + 2];
    }
    h.e1.resize(ne); h.e2.resize(ne); h.t1.resize(nt); h.t2.resize(nt); h.t3.resize(nt);
    for (int i = 0; i < ne; ++i) { h.e1[i] = edges[2 * i]; h.e2[i] = edges[2 * i + 1]; }
    for (int i = 0; i < nt; ++i) { h.t1[i] = tris[3 * i]; h.t2[i] = tris[3 * i + 1]; h.t3[i] = tris[3 * i + 2]; }
    /* rest lengths (maincode.cpp:247-250) */
    std::vector<double> l0(ne), kp(ne), ta(nt);
    for (int i = 0; i < ne; ++i) {
        const int a = h.e1[i], b = h.e2[i];
        double dx = (x[a] - x[b]) * (x[a] - x[b]), dy = (y[a] - y[b]) * (y[a] - y[b]),
               dz = (zz[a] - zz[b]) * (zz[a] - zz[b]);
        l0[i] = std::sqrt(dx + dy + dz);
    }
    /* hinge partners: third vertex of the first and of the last triangle on the edge (maincode.cpp:276-342) */
    std::vector<int32_t> es1(ne, 0), es2(ne, 0);
    {
        std::vector<std::vector<int32_t>> vt(nv);
        for (int j = 0; j < nt; ++j) {
            vt[h.t1[j]].push_back(j);
            vt[h.t2[j]].push_back(j);
            vt[h.t3[j]].push_back(j);
        }
        for (auto &l : vt) { std::sort(l.begin(), l.end()); l.erase(std::unique(l.begin(), l.end()), l.end()); }
        for (int i = 0; i < ne; ++i) {
            const int a = h.e1[i], b = h.e2[i];
            std::vector<int32_t> both;
            std::set_intersection(vt[a].begin(), vt[a].end(), vt[b].begin(), vt[b].end(), std::back_inserter(both));
            auto third = [&](int j, int32_t &dst) {
                if (h.t1[j] != a && h.t1[j] != b) dst = h.t1[j];
                if (h.t2[j] != a && h.t2[j] != b) dst = h.t2[j];
                if (h.t3[j] != a && h.t3[j] != b) dst = h.t3[j];
            };
            if (!both.empty()) {
                third(both.front(), es1[i]);
                third(both.back(), es2[i]);
            }
        }
    }
    auto area = [&](int a, int b, int cc) {
        double A[3] = {x[b] - x[a], y[b] - y[a], zz[b] - zz[a]};
        double B[3] = {x[cc] - x[a], y[cc] - y[a], zz[cc] - zz[a]};
        double e0 = (A[1] * B[2]) - (B[1] * A[2]), e1_ = (A[2] * B[0]) - (B[2] * A[0]),
               e2_ = (A[0] * B[1]) - (B[0] * A[1]);
        return 0.5 * std::sqrt(e0 * e0 + e1_ * e1_ + e2_ * e2_);
    };
    auto vol = [&](int a, int b, int cc) {
        double Cn0 = (x[a] + x[b] + x[cc]) / 3.0, Cn1 = (y[a] + y[b] + y[cc]) / 3.0,
               Cn2 = (zz[a] + zz[b] + zz[cc]) / 3.0;
        double A[3] = {x[b] - x[a], y[b] - y[a], zz[b] - zz[a]};
        double B[3] = {x[cc] - x[a], y[cc] - y[a], zz[cc] - zz[a]};
        double D0 = (A[1] * B[2]) - (B[1] * A[2]), D1 = (A[2] * B[0]) - (B[2] * A[0]),
               D2 = (A[0] * B[1]) - (B[0] * A[1]);
        return (1.0 / 6.0) * ((D0 * Cn0) + (D1 * Cn1) + (D2 * Cn2));
    };
    /* A0, V0 (TotalVA0, Src/memfuncs.cpp:192-211) */
    double A0 = 0, V0 = 0;
    for (int i = 0; i < nt; ++i) {
        A0 += area(h.t1[i], h.t2[i], h.t3[i]);
        V0 += vol(h.t1[i], h.t2[i], h.t3[i]);
    }
    /* kappap (maincode.cpp:347-361) */
    for (int i = 0; i < ne; ++i) {
        double ll0 = l0[i], lmax = 2.5 * ll0, x0 = ll0 / lmax, om = 1.0 - x0;
        double s1 = 1.0 / (4.0 * om * om), s2 = x0 / (2.0 * om * om * om);
        double s4 = (1.732 / (4.0 * p.elap * lmax * x0));
        double sh = s4 * s2 - s4 * s1 + s4 * 0.25;
        kp[i] = (p.shear - sh) * 4.0 * ll0 * ll0 * ll0 / (3.0 * 1.732);
    }
    for (int i = 0; i < nt; ++i) ta[i] = area(h.t1[i], h.t2[i], h.t3[i]);
    /* theta0 (maincode.cpp:207-209) */
    double numer = 1.73205080757 * ((double)nv - 2.0) - 5.0 * p.Pi;
    double denom = 1.73205080757 * ((double)nv - 2.0) - 3.0 * p.Pi;
    double theta0 = 180.0 * std::acos(numer / denom) / p.Pi;
    /* centre of mass and hinge orientation (Src/memfuncs.cpp:55-74, 533-588) */
    double rr[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i) { rr[0] += x[i] + 0 * (double)p.Lx; rr[1] += y[i] + 0 * (double)p.Ly; rr[2] += zz[i]; }
    double cen[3] = {std::fmod(rr[0] / (double)nv, (double)p.Lx), std::fmod(rr[1] / (double)nv, (double)p.Ly),
                     std::fmod(rr[2] / (double)nv, (double)p.Lz)};
    std::vector<int32_t> ied(ne, 1);
    for (int i = 0; i < ne; ++i) {
        const int l2 = h.e1[i], l3 = h.e2[i], l1 = es1[i], l4 = es2[i];
        double P1[3] = {x[l1], y[l1], zz[l1]}, P2[3] = {x[l2], y[l2], zz[l2]};
        double P3[3] = {x[l3], y[l3], zz[l3]}, P4[3] = {x[l4], y[l4], zz[l4]};
        double a21[3], a31[3], a24[3], a34[3], tt1[3], tt2[3];
        for (int d = 0; d < 3; ++d) {
            a21[d] = P2[d] - P1[d]; a31[d] = P3[d] - P1[d];
            a24[d] = P2[d] - P4[d]; a34[d] = P3[d] - P4[d];
            tt1[d] = ((P1[d] + P2[d] + P3[d]) / 3.0) - cen[d];
            tt2[d] = ((P4[d] + P2[d] + P3[d]) / 3.0) - cen[d];
        }
        double eta[3] = {(a21[1] * a31[2]) - (a31[1] * a21[2]), (a21[2] * a31[0]) - (a31[2] * a21[0]),
                         (a21[0] * a31[1]) - (a31[0] * a21[1])};
        double zeta[3] = {(a34[1] * a24[2]) - (a24[1] * a34[2]), (a34[2] * a24[0]) - (a24[2] * a34[0]),
                          (a34[0] * a24[1]) - (a24[0] * a34[1])};
        double che1 = eta[0] * tt1[0] + eta[1] * tt1[1] + eta[2] * tt1[2];
        double che2 = zeta[0] * tt2[0] + zeta[1] * tt2[1] + zeta[2] * tt2[2];
        ied[i] = (che1 < 0 && che2 < 0) ? -1 : 1;
    }
    /* per-vertex force-slot incidence, in the reference's summation order */
    std::vector<std::vector<int32_t>> inc(nv);
    for (int e = 0; e < ne; ++e) {
        inc[h.e1[e]].push_back(4 * e + 0);
        inc[h.e1[e]].push_back(4 * e + 2);
        inc[h.e2[e]].push_back(4 * e + 1);
        inc[h.e2[e]].push_back(4 * e + 3);
    }
    const int ST = 4 * ne, SH = 4 * ne + 9 * nt;
    for (int t = 0; t < nt; ++t) {
        const int vv[3] = {h.t1[t], h.t2[t], h.t3[t]};
        for (int a = 0; a < 3; ++a)
            for (int m = 0; m < 3; ++m) inc[vv[a]].push_back(ST + 9 * t + 3 * a + m);
    }
    for (int e = 0; e < ne; ++e) {
        int l2 = h.e1[e], l3 = h.e2[e];
        if (ied[e] == -1) std::swap(l2, l3);
        inc[es1[e]].push_back(SH + 4 * e + 0);
        inc[l2].push_back(SH + 4 * e + 1);
        inc[l3].push_back(SH + 4 * e + 2);
        inc[es2[e]].push_back(SH + 4 * e + 3);
    }
    std::vector<int32_t> ioff(nv + 1, 0), order;
    for (int v = 0; v < nv; ++v) {
        ioff[v + 1] = ioff[v] + (int32_t)inc[v].size();
        order.insert(order.end(), inc[v].begin(), inc[v].end());
    }
    if ((int64_t)order.size() != inc_bound(h.spec)) return fail(IRBC_ERR_ARG, "irbc_set_mesh: every edge needs two hinge partners");
    /* element term -> its slot: term k of vertex v (in the reference's
     * summation order) at k * nv + v */
    size_t kmax = 0;
    for (int v = 0; v < nv; ++v) kmax = std::max(kmax, inc[v].size());
    if ((int64_t)kmax * nv > scr_bound(h.spec))
        return fail(IRBC_ERR_ARG, "irbc_set_mesh: a vertex has %zu element terms, more than twice the mean", kmax);
    std::vector<int32_t> ient(order.size());
    for (int v = 0; v < nv; ++v)
        for (size_t k = 0; k < inc[v].size(); ++k) ient[inc[v][k]] = (int32_t)(k * nv + v);
    R.Sarea0 = A0;
    R.Volume0 = V0;
    R.theta0 = theta0;
    R.cos_t0 = std::cos(theta0 * p.Pi / 180.0);
    R.sin_t0 = std::sin(theta0 * p.Pi / 180.0);
    hipStream_t s = c->st;
    std::vector<double> zero(nv, 0.0);
    std::vector<int32_t> izero(nv, 0), rno(nv, IRBC_RANDNO);
    HIPCHK(hipMemcpyAsync(c->mx + R.voff, x.data(), nv * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->my + R.voff, y.data(), nv * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->mz + R.voff, zz.data(), nv * 8, hipMemcpyHostToDevice, s));
    double *zs[] = {c->mvx, c->mvy, c->mvz, c->mfx, c->mfy, c->mfz, c->poff, c->pon};
    for (double *d : zs) HIPCHK(hipMemcpyAsync(d + R.voff, zero.data(), nv * 8, hipMemcpyHostToDevice, s));
    int32_t *is[] = {c->mix, c->miy, c->mpid, c->bondi, c->knob, c->ncand};
    for (int32_t *d : is) HIPCHK(hipMemcpyAsync(d + R.voff, izero.data(), nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->bondl + R.voff, rno.data(), nv * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->e1 + R.eoff, h.e1.data(), ne * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->e2 + R.eoff, h.e2.data(), ne * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->es1 + R.eoff, es1.data(), ne * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->es2 + R.eoff, es2.data(), ne * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->iedge + R.eoff, ied.data(), ne * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->l0 + R.eoff, l0.data(), ne * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->kappap + R.eoff, kp.data(), ne * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->t1 + R.toff, h.t1.data(), nt * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->t2 + R.toff, h.t2.data(), nt * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->t3 + R.toff, h.t3.data(), nt * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->tarea + R.toff, ta.data(), nt * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->inc_off + R.inc_off, ioff.data(), (nv + 1) * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->inc_ent + R.inc_ent_off, ient.data(), ient.size() * 4, hipMemcpyHostToDevice, s));
    /* every slot's tag NaN (all bytes 0xFF): a finished term, until the
     * triangle terms tag theirs */
    HIPCHK(hipMemsetAsync(c->tag + R.scr_off, 0xFF, (size_t)scr_bound(h.spec) * 8, s));
    h.mesh = true;
    return c->push(r);
}

int irbc_set_vertices(irbc_ctx *c, int r, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *fx,
                      const double *fy, const double *fz, const int32_t *ix, const int32_t *iy,
                      const int32_t *pid, const int32_t *bondi, const int32_t *bondl,
                      const int32_t *knob)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    const size_t n = c->hr[r].spec.n_vertices;
    const double *ds[] = {x, y, z, vx, vy, vz, fx, fy, fz};
    double *dd[] = {c->mx, c->my, c->mz, c->mvx, c->mvy, c->mvz, c->mfx, c->mfy, c->mfz};
    for (int k = 0; k < 9; ++k)
        if (ds[k]) HIPCHK(hipMemcpyAsync(dd[k] + R.voff, ds[k], n * 8, hipMemcpyHostToDevice, c->st));
    c->primed = false;
    const int32_t *is[] = {ix, iy, pid, bondi, bondl, knob};
    int32_t *id[] = {c->mix, c->miy, c->mpid, c->bondi, c->bondl, c->knob};
    for (int k = 0; k < 6; ++k)
        if (is[k]) HIPCHK(hipMemcpyAsync(id[k] + R.voff, is[k], n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(c->sync_both());
    return IRBC_OK;
}

int irbc_get_vertices(irbc_ctx *c, int r, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *fx, double *fy, double *fz, double *poff, double *pon,
                      int32_t *ix, int32_t *iy, int32_t *pid, int32_t *bondi, int32_t *bondl,
                      int32_t *knob)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    const size_t n = c->hr[r].spec.n_vertices;
    HIPCHK(c->sync_both());
    double *dd[] = {x, y, z, vx, vy, vz, fx, fy, fz, poff, pon};
    const double *ds[] = {c->mx, c->my, c->mz, c->mvx, c->mvy, c->mvz, c->mfx, c->mfy, c->mfz, c->poff, c->pon};
    for (int k = 0; k < 11; ++k)
        if (dd[k]) HIPCHK(hipMemcpy(dd[k], ds[k] + R.voff, n * 8, hipMemcpyDeviceToHost));
    int32_t *id[] = {ix, iy, pid, bondi, bondl, knob};
    const int32_t *is[] = {c->mix, c->miy, c->mpid, c->bondi, c->bondl, c->knob};
    for (int k = 0; k < 6; ++k)
        if (id[k]) HIPCHK(hipMemcpy(id[k], is[k] + R.voff, n * 4, hipMemcpyDeviceToHost));
    return IRBC_OK;
}

/* particle indices in order of their unshifted cell (a stable counting
 * sort; positions outside the box are folded in, for the order only) */
static std::vector<int64_t> cell_order(const DevRep &R, const double *x, const double *y, const double *z, int64_t n)
{
    const int Lx = R.Lx, Ly = R.Ly, Lz = R.Lz;
    const int64_t nc = (int64_t)Lx * Ly * (Lz + 1);
    std::vector<int32_t> key(n);
    std::vector<int64_t> start(nc + 1, 0), order(n);
    auto fold = [](double v, int L) {
        const double w = v - L * std::floor(v / L);
        return (w >= 0.0 && w < L) ? std::min(L - 1, (int)w) : 0; /* NaN and inf fold to 0 */
    };
    for (int64_t i = 0; i < n; ++i) {
        const double zz = z[i] + 1.0;
        const int fz = (zz >= 0.0 && zz < Lz + 1.0) ? std::min(Lz, (int)zz) : 0;
        key[i] = (int32_t)(((int64_t)fold(x[i], Lx) * Ly + fold(y[i], Ly)) * (Lz + 1) + fz);
        ++start[key[i] + 1];
    }
    for (int64_t k = 0; k < nc; ++k) start[k + 1] += start[k];
    for (int64_t i = 0; i < n; ++i) order[start[key[i]]++] = i;
    return order;
}

int irbc_set_solvent(irbc_ctx *c, int r, int64_t n, const double *x, const double *y,
                     const double *z, const double *vx, const double *vy, const double *vz,
                     const int32_t *indio)
{
    API_GUARD;
    CHECK_R(c, r);
    if (n != c->hr[r].spec.n_solvent || !x || !y || !z || !vx || !vy || !vz)
        return fail(IRBC_ERR_ARG, "irbc_set_solvent: expected %lld particles", (long long)c->hr[r].spec.n_solvent);
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    /* The particles go to the device in cell order (unshifted cells, key
     * layout of cell_of): the first streaming steps and the first
     * collision's sort then touch neighbouring cells together, as every step
     * does once a collision has sorted them (a random upload order made the
     * first ten steps twice as slow and the first sort seven times).  No
     * result depends on the storage order: cell sums are exact integer sums,
     * every other per-particle step is independent, downloads return
     * particle order. */
    const std::vector<int64_t> order = cell_order(R, x, y, z, n);
    std::vector<float> f(n);
    const double *src[] = {x, y, z, vx, vy, vz};
    float *dst[] = {c->xa, c->ya, c->za, c->vxa, c->vya, c->vza};
    for (int k = 0; k < 6; ++k) {
        for (int64_t i = 0; i < n; ++i) f[i] = (float)src[k][order[i]];
        HIPCHK(hipMemcpy(dst[k] + R.poff, f.data(), n * 4, hipMemcpyHostToDevice));
    }
    std::vector<uint32_t> id(n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t j = order[i];
        id[i] = (uint32_t)j | ((indio ? (uint32_t)(indio[j] != 0) : 1u) << 31);
    }
    HIPCHK(hipMemcpy(c->ida + R.poff, id.data(), n * 4, hipMemcpyHostToDevice));
    /* no collision yet: the records' keys (the particles' last cell keys) read 0 */
    HIPCHK(hipMemset(R.rec + 4 * R.poff, 0, (size_t)n * 16));
    HIPCHK(hipMemset(R.rec1 + 4 * R.poff, 0, (size_t)n * 16));
    R.rotate_pending = 0;
    R.key_src = 1;
    c->hr[r].solvent = true;
    return c->push(r);
}

static int launch_scan(irbc_ctx *c, int staged, int which, int32_t *in, int32_t *out, int only_rep = -1);

int irbc_init_solvent(irbc_ctx *c, int r, int maxwell)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    const unsigned nb = (unsigned)((R.np + IRBC_WG - 1) / IRBC_WG), T = (unsigned)c->z.tiles;
    k_init_solvent<<<nb, IRBC_WG, 0, c->st>>>(c->d_rep, r, maxwell ? 1 : 0);
    k_init_bin<<<T, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, (int)T, r, c->skey, c->srank, c->ccount);
    rc = launch_scan(c, 1, 0, c->ccount, c->cstart, r);
    if (rc) return rc;
    k_scatter<<<T, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, (int)T, 0, 1, c->skey, c->srank, c->cstart, r);
    k_unpack<<<nb, IRBC_WG, 0, c->st>>>(c->d_rep, r);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->st));
    rc = c->pull(r); /* the scan epoch the binning advanced */
    if (rc) return rc;
    c->rep[r].rotate_pending = 0;
    c->rep[r].key_src = 1;
    c->hr[r].solvent = true;
    return c->push(r);
}

int irbc_get_solvent(irbc_ctx *c, int r, double *x, double *y, double *z, double *vx, double *vy,
                     double *vz, int32_t *key, int32_t *indio)
{
    API_GUARD;
    CHECK_R(c, r);
    const int64_t n = c->hr[r].spec.n_solvent;
    HIPCHK(c->sync_both());
    k_unpermute<<<(unsigned)((n + IRBC_WG - 1) / IRBC_WG), IRBC_WG, 0, c->st>>>(c->d_rep, r,
                                                                               c->stage_f, c->stage_i);
    LAUNCHCHK();
    HIPCHK(c->sync_both());
    std::vector<float> f(6 * n);
    std::vector<int32_t> iv(2 * n);
    HIPCHK(hipMemcpy(f.data(), c->stage_f, 6 * n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(iv.data(), c->stage_i, 2 * n * 4, hipMemcpyDeviceToHost));
    double *dd[] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < 6; ++k)
        if (dd[k])
            for (int64_t i = 0; i < n; ++i) dd[k][i] = (double)f[k * n + i];
    if (key) std::memcpy(key, iv.data(), n * 4);
    if (indio) std::memcpy(indio, iv.data() + n, n * 4);
    return IRBC_OK;
}

int irbc_set_parasite(irbc_ctx *c, int r, const double *d, const int32_t *i)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!d || !i) return fail(IRBC_ERR_ARG, "irbc_set_parasite: null array");
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    for (int k = 0; k < 3; ++k) {
        R.para[k] = d[k];
        R.paraV[k] = d[3 + k];
        R.paraft[k] = d[6 + k];
        R.parafdt[k] = d[9 + k];
    }
    R.paradx = i[0];
    R.parady = i[1];
    R.paraid = i[2];
    return c->push(r);
}

int irbc_get_parasite(irbc_ctx *c, int r, double *d, int32_t *i)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    const DevRep &R = c->rep[r];
    for (int k = 0; k < 3; ++k) {
        d[k] = R.para[k];
        d[3 + k] = R.paraV[k];
        d[6 + k] = R.paraft[k];
        d[9 + k] = R.parafdt[k];
    }
    i[0] = R.paradx;
    i[1] = R.parady;
    i[2] = R.paraid;
    return IRBC_OK;
}

int64_t irbc_n_ligands(const irbc_ctx *c, int r)
{
    API_GUARD;
    if (!c || r < 0 || r >= c->nrep) return IRBC_ERR_ARG;
    return c->rep[r].nlig;
}

int irbc_set_ligands(irbc_ctx *c, int r, const int32_t *ligi)
{
    API_GUARD;
    CHECK_R(c, r);
    HIPCHK(hipMemcpy(c->ligi + c->rep[r].loff, ligi, (size_t)c->rep[r].nlig * 4, hipMemcpyHostToDevice));
    return IRBC_OK;
}

int irbc_get_ligands(irbc_ctx *c, int r, int32_t *ligi)
{
    API_GUARD;
    CHECK_R(c, r);
    HIPCHK(c->sync_both());
    HIPCHK(hipMemcpy(ligi, c->ligi + c->rep[r].loff, (size_t)c->rep[r].nlig * 4, hipMemcpyDeviceToHost));
    return IRBC_OK;
}

int irbc_get_spec(const irbc_ctx *c, int r, irbc_replica_spec *out)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!out) return fail(IRBC_ERR_ARG, "irbc_get_spec: null output");
    *out = c->hr[r].spec;
    return IRBC_OK;
}

int irbc_get_globals(irbc_ctx *c, int r, double *g)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    const DevRep &R = c->rep[r];
    g[0] = R.Sarea;
    g[1] = R.Volume;
    g[2] = R.Sarea0;
    g[3] = R.Volume0;
    g[4] = R.theta0;
    g[5] = R.wall2[c->par];
    g[6] = R.status;
    g[7] = R.bond_draws;
    g[8] = (double)R.step;
    g[9] = R.nbonds;
    g[10] = R.near_count[c->par ^ 1];
    g[11] = R.chain_xcc_last;
    g[12] = R.key_src;
    return IRBC_OK;
}

int irbc_set_counters(irbc_ctx *c, int r, int64_t step, int32_t colsteps, int32_t upd)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!counters_valid(c->hr[r].spec.params, step, colsteps, upd))
        return fail(IRBC_ERR_ARG, "irbc_set_counters: need step >= 0, -1 <= colsteps < csteps, -1 <= upd < uplist");
    int rc = c->pull(r);
    if (rc) return rc;
    HostRep &h = c->hr[r];
    h.step = step;
    h.colsteps = colsteps;
    h.upd = upd;
    top_of_step(h.spec.params, step, colsteps, upd, c->rep[r]);
    return c->push(r);
}

int irbc_get_counters(irbc_ctx *c, int r, int64_t *step, int32_t *colsteps, int32_t *upd)
{
    API_GUARD;
    CHECK_R(c, r);
    *step = c->hr[r].step;
    *colsteps = c->hr[r].colsteps;
    *upd = c->hr[r].upd;
    return IRBC_OK;
}

int irbc_set_wallspeed(irbc_ctx *c, int r, double w)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    c->rep[r].wall2[0] = c->rep[r].wall2[1] = w;
    return c->push(r);
}

int irbc_set_seed(irbc_ctx *c, int r, uint64_t seed)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    c->rep[r].seed = seed;
    c->hr[r].spec.seed = seed;
    return c->push(r);
}

int irbc_set_series(irbc_ctx *c, int r, int32_t every, int32_t cap)
{
    API_GUARD;
    CHECK_R(c, r);
    if (every < 0 || cap < 0 || (every > 0 && cap == 0))
        return fail(IRBC_ERR_ARG, "irbc_set_series: every >= 0 and cap >= 1 when on");
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    if (every > 0 && cap > R.series_cap) {
        double *buf = nullptr;
        rc = c->alloc(buf, (int64_t)cap * IRBC_SERIES_REC);
        if (rc) return rc;
        R.series = buf; /* freed with the context */
    }
    R.series_every = every;
    if (every > 0) R.series_cap = std::max(cap, R.series_cap);
    R.series_n = 0;
    return c->push(r);
}

int64_t irbc_get_series(irbc_ctx *c, int r, double *out, int64_t cap)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    const DevRep &R = c->rep[r];
    const int64_t n = R.series_n;
    if (out && n > 0 && R.series)
        HIPCHK(hipMemcpy(out, R.series, std::min(n, cap) * IRBC_SERIES_REC * sizeof(double),
                         hipMemcpyDeviceToHost));
    return n;
}

int irbc_query(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "irbc_query: null context");
    hipStream_t ss[2] = {c->st, c->st2};
    for (hipStream_t q : ss) {
        const hipError_t e = hipStreamQuery(q);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return fail(IRBC_ERR_HIP, "irbc_query: %s", hipGetErrorString(e));
    }
    return 1;
}

int irbc_use_tapes(irbc_ctx *c, int r, int on)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = c->pull(r);
    if (rc) return rc;
    c->rep[r].tapes = on ? 1 : 0;
    return c->push(r);
}

int irbc_set_collision_tape(irbc_ctx *c, int r, const double *shift_u, const double *cell_u,
                            const double *wall_z, const double *mem_z)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!shift_u || !cell_u || !wall_z || !mem_z) return fail(IRBC_ERR_ARG, "irbc_set_collision_tape: null tape");
    int rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    const size_t nc = R.ncell;
    HIPCHK(hipMemcpy(c->cell_u + 2 * R.coff, cell_u, 2 * nc * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wall_z + 3 * R.coff, wall_z, 3 * nc * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->mem_z + 3 * R.coff, mem_z, 3 * nc * 8, hipMemcpyHostToDevice));
    for (int k = 0; k < 3; ++k) R.shift[k] = 1.0 * shift_u[k];
    R.tapes = 1;
    return c->push(r);
}

int irbc_set_bond_tape(irbc_ctx *c, int r, const double *u, int64_t n)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    if (n < 0 || n > (int64_t)R.nv * IRBC_MAX_CAND) return fail(IRBC_ERR_ARG, "irbc_set_bond_tape: too many draws");
    if (n > 0) HIPCHK(hipMemcpy(c->bond_u + (int64_t)R.voff * IRBC_MAX_CAND, u, n * 8, hipMemcpyHostToDevice));
    return IRBC_OK;
}

/* ---------------- stages ---------------- */

static int launch_scan(irbc_ctx *c, int staged, int which, int32_t *in, int32_t *out, int only_rep)
{
    const unsigned nb = (unsigned)c->z.scanblk;
    k_scan_lb<<<nb, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_blk, (int)nb, staged, which, in, out, c->scanflag,
                                         c->scanticket, only_rep);
    LAUNCHCHK();
    return IRBC_OK;
}

#define IRBC_GRAPH_STEPS 10 /* steps per captured chunk: one collision period */
#define IRBC_TUNE_WARM_MAX 2100  /* most steps the autotuner runs to capture the chosen graphs */
#define IRBC_TUNE_CANDIDATES 512 /* index space: 9 option bits (irbc_autotune) */
#ifndef IRBC_STREAM_WGS_PER_CU
#define IRBC_STREAM_WGS_PER_CU 4 /* capped bulk-streaming grid: workgroups per CU */
#endif

/* grid for a kernel that maps blocks through xcd_tile */
static unsigned xcd_grid(unsigned n) { return IRBC_NXCD * ((n + IRBC_NXCD - 1) / IRBC_NXCD); }

/* Collision flags of a step: which kinds of collision its replicas run. */
#define COL_SORT 1    /* a replica's collision sorts (and every staged one) */
#define COL_INPLACE 2 /* a replica's collision leaves its particles in place */

static int launch_collision(irbc_ctx *c, int staged, int diag, int cf)
{
    if (staged) cf = COL_SORT;
    const unsigned T = (unsigned)c->z.tiles, CT = (unsigned)c->z.ctiles;
    if (staged) k_draw_shift<<<(unsigned)((c->nrep + 63) / 64), 64, 0, c->st>>>(c->d_rep, c->nrep, 0);
    const int xc = c->xcd_chunk;
    MemDev M = c->mem(c->par);
    const unsigned nvb = IRBC_NXCD * ((unsigned)(c->nrep + IRBC_NXCD - 1) / IRBC_NXCD);
    /* the cell-draw blocks keep the particle tiles' XCD chunking intact:
     * their count is a multiple of IRBC_NXCD */
    const unsigned ndb = xcd_grid(CT);
    /* the bin-sort tile shape: spans of bin_span tiles when every collision
     * of the step sorts (the span path gives no in-place sums) */
    const int span = (cf & COL_INPLACE) ? 1 : c->bin_span;
    const unsigned NS = (T + (unsigned)span - 1) / (unsigned)span;
    auto kb = span == 4 ? k_bin<4> : span == 2 ? k_bin<2> : k_bin<1>;
    kb<<<nvb + ndb + (xc ? xcd_grid(NS) : NS), IRBC_WG, 0, c->st>>>(
        c->d_rep, c->nrep, (int)nvb, c->d_tile_rep, (int)T, xc, staged, c->skey, c->srank, c->ccount, c->csum, M,
        c->vhead, c->vnext, c->d_ctile_rep, (int)CT, (int)ndb, c->cell_u, c->ctab, c->wall_z, diag, c->diag);
    if (cf & COL_SORT) {
        int rc = launch_scan(c, staged, 0, c->ccount, c->cstart);
        if (rc) return rc;
        k_scatter<<<xc ? xcd_grid(T) : T, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, (int)T, xc, staged, c->skey,
                                                               c->srank, c->cstart, -1);
    }
    /* without in-place replicas k_cell also rotates the vertices and parasites (k_vrot's work) */
    const bool inpl = (cf & COL_INPLACE) && !staged;
    auto kc = inpl ? k_cell<true> : k_cell<false>;
    kc<<<xc ? xcd_grid(CT) : CT, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_ctile_rep, (int)CT, xc, staged, c->cstart,
                                                         c->vhead, c->vnext, c->mvx, c->mvy, c->mvz, c->wall_z,
                                                         c->mem_z, c->ctab, diag, c->diag, c->par, c->ccount,
                                                         c->csum, M, inpl ? 0 : 1);
    if (cf & COL_INPLACE) {
        k_dvv<<<xc ? xcd_grid(T) : T, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, (int)T, xc, c->skey, c->ctab,
                                                           c->csum);
        k_cell_the<<<xc ? xcd_grid(CT) : CT, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_ctile_rep, (int)CT, xc, c->vhead,
                                                                 c->vnext, c->mvx, c->mvy, c->mvz, c->ctab, diag,
                                                                 c->diag, c->ccount, c->csum);
    }
    if (inpl) k_vrot<<<c->nrep, IRBC_WG, 0, c->st>>>(c->d_rep, staged, M, c->ctab);
    if (staged) k_rotate<<<T, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, c->ctab);
    LAUNCHCHK();
    return IRBC_OK;
}

/* the compacted list of cells that list a triangle, and cleared masks,
 * for the replicas that made lists (all when staged) */
static int launch_listed(irbc_ctx *c, int staged)
{
    k_listed_reset<<<(unsigned)((c->nrep + 63) / 64), 64, 0, c->st>>>(c->d_rep, c->nrep, staged);
    k_cell_listed<<<(unsigned)c->z.ctiles, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_ctile_rep, staged, c->lloff,
                                                               c->listed, c->cmask);
    LAUNCHCHK();
    return IRBC_OK;
}

static int launch_llist(irbc_ctx *c, int staged)
{
    int maxnt = 0;
    for (int r = 0; r < c->nrep; ++r) maxnt = std::max(maxnt, c->rep[r].nt);
    dim3 g((unsigned)((maxnt * 64 + IRBC_WG - 1) / IRBC_WG), (unsigned)c->nrep);
    MemView V = c->view(c->par);
    k_tri_count<<<g, IRBC_WG, 0, c->st>>>(c->d_rep, staged, V, c->mix, c->miy, c->llcount);
    int rc = launch_scan(c, staged, 1, c->llcount, c->lloff);
    if (rc) return rc;
    k_tri_fill<<<g, IRBC_WG, 0, c->st>>>(c->d_rep, staged, V, c->mix, c->miy, c->lloff, c->llcount, c->llent);
    k_tri_sort<<<(unsigned)c->z.ctiles, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_ctile_rep, staged, c->lloff,
                                                            c->llcount, c->llent);
    LAUNCHCHK();
    return launch_listed(c, staged);
}

/* Grids of the kernels that stride over a batch's work flattened across
 * its replicas (flat_prefix): per_rep workgroups per replica, at most
 * per_cu per CU in all; the prefix takes nrep + 1 ints of LDS. */
static unsigned flat_grid(const irbc_ctx *c, int64_t per_rep, int per_cu)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(per_rep * c->nrep, (int64_t)per_cu * c->ncu));
}
static size_t flat_lds(const irbc_ctx *c) { return (size_t)(c->nrep + 1) * sizeof(int32_t); }

/* the cells' near-membrane sub-cell masks from the current lists and
 * triangle kinematics;
 * staged: every replica, else those that collide or make lists this step */
#ifndef IRBC_MASK_PER_CU
#define IRBC_MASK_PER_CU 8
#endif
static int launch_cell_mask(irbc_ctx *c, int staged)
{
    if (staged) {
        /* the lists may have come from elsewhere (a checkpoint, a stage call) */
        const int rc = launch_listed(c, 1);
        if (rc) return rc;
    }
    /* one wave per listed cell, about two and a half cells per triangle,
     * flattened over the batch: a grid of that size for one replica, at
     * most eight workgroups per CU for a batch (they stride) */
    const unsigned G = flat_grid(c, std::max(16, c->z.maxnt * 5 / 8), IRBC_MASK_PER_CU);
    k_cell_mask<<<G, IRBC_WG, flat_lds(c), c->st>>>(
        c->d_rep, c->nrep, staged, c->par & 1, c->lloff, c->llent, c->listed,
        c->tri_kin + (c->par & 1) * c->tri_stride(), c->tri_sph + (c->par & 1) * c->z.nt, c->cmask, c->llmask,
        c->near_masks);
    LAUNCHCHK();
    return IRBC_OK;
}

/* Bounce-back grids per replica.  The queue holds the particles of marked
 * sub-cells, 0.2-1% of a replica (up to ~2% while the drift margin is
 * large), and every queued particle should find its lanes in one pass:
 * the one-pass kernel gives a particle 16 lanes (16 particles per block),
 * the sphere pass one lane (256 per block). */
static int64_t max_np(const irbc_ctx *c)
{
    int64_t maxnp = 0;
    for (const DevRep &R : c->rep) maxnp = std::max<int64_t>(maxnp, R.np);
    return maxnp;
}

static unsigned near_blocks(const irbc_ctx *c) { return (unsigned)std::max<int64_t>(16, (max_np(c) + 2047) / 2048); }

/* the sphere pass: IRBC_SPHERE_LANES lanes per queued particle, sized for a
 * queue of up to 2% of a replica's particles in one pass, at most 16
 * workgroups per CU for a batch (they stride over its flattened queue) */
#ifndef IRBC_SPHERE_PER_CU
#define IRBC_SPHERE_PER_CU 16
#endif
#ifndef IRBC_HIT_PER_CU
#define IRBC_HIT_PER_CU 4
#endif
static unsigned sphere_blocks(const irbc_ctx *c)
{
    const int64_t per_block = IRBC_WG / IRBC_SPHERE_LANES;
    return (unsigned)std::max<int64_t>(8, (max_np(c) / 50 + per_block - 1) / per_block);
}
static void launch_sphere(irbc_ctx *c, const double *tk, const float4 *ts)
{
    k_near_sphere<<<flat_grid(c, sphere_blocks(c), IRBC_SPHERE_PER_CU), IRBC_WG, flat_lds(c), c->st>>>(
        c->d_rep, c->nrep, c->lloff, c->llent, c->llmask, c->nearq, c->deferq, tk, ts, c->par);
}

static int launch_near(irbc_ctx *c)
{
    const double *tk = c->tri_kin + (c->par & 1) * c->tri_stride();
    const float4 *ts = c->tri_sph + (c->par & 1) * c->z.nt;
    if (c->near_merged) {
        k_near<<<dim3(near_blocks(c), (unsigned)c->nrep), IRBC_WG, 0, c->st>>>(c->d_rep, c->lloff, c->llent,
                                                                             c->nearq, tk, ts, c->par);
        LAUNCHCHK();
        return IRBC_OK;
    }
    launch_sphere(c, tk, ts);
    k_near_hit<<<flat_grid(c, std::max(16u, near_blocks(c) / 8u), IRBC_HIT_PER_CU), IRBC_WG, flat_lds(c), c->st>>>(
        c->d_rep, c->nrep, c->lloff, c->llent, c->deferq, tk, ts, c->par);
    LAUNCHCHK();
    return IRBC_OK;
}

/* the bulk streaming kernel in the chosen shape; e0/e1 (both or neither)
 * are events bound to the dispatch; with tri_kin its first blocks write
 * this step's triangle kinematics from the vertex kinematics */
static int launch_bulk(irbc_ctx *c, hipEvent_t e0, hipEvent_t e1, bool tri_kin)
{
    const unsigned T = (unsigned)c->z.tiles;
    const unsigned cap = (unsigned)(c->ncu * IRBC_STREAM_WGS_PER_CU);
    const MemDev M = c->mem(c->par);
    const TriKin tk{M.t1, M.t2, M.t3, M.vkin, M.tri_kin, M.tri_sph};
    const int tkb = std::max(1, (c->z.maxnt + IRBC_WG - 1) / IRBC_WG);
    const int ntk = tri_kin ? tkb * c->nrep : 0;
    const int32_t *tr = c->d_tile_rep, *lo = c->lloff;
    const double *ct = c->ctab;
    const uint64_t *cm = c->cmask;
    if (c->stream4) {
        unsigned G = (T + IRBC_WG / 64 - 1) / (IRBC_WG / 64);
        if (c->stream_grid) G = std::min(G, cap);
        G += (unsigned)ntk;
        if (e0)
            hipExtLaunchKernelGGL(k_stream4, dim3(G), dim3(IRBC_WG), 0, c->st, e0, e1, 0, c->d_rep, tr, (int)T, ct,
                                  lo, cm, c->nearq, c->par, tk, tkb, ntk);
        else
            k_stream4<<<G, IRBC_WG, 0, c->st>>>(c->d_rep, tr, (int)T, ct, lo, cm, c->nearq, c->par, tk, tkb, ntk);
    } else {
        const bool capped = c->stream_grid && T > cap;
        const unsigned G = (capped ? cap : T) + (unsigned)ntk;
        auto *kern = capped ? k_stream<true> : k_stream<false>;
        if (e0)
            hipExtLaunchKernelGGL(kern, dim3(G), dim3(IRBC_WG), 0, c->st, e0, e1, 0, c->d_rep, tr, (int)T, ct, lo,
                                  cm, c->nearq, c->par, tk, tkb, ntk);
        else
            kern<<<G, IRBC_WG, 0, c->st>>>(c->d_rep, tr, (int)T, ct, lo, cm, c->nearq, c->par, tk, tkb, ntk);
    }
    LAUNCHCHK();
    return IRBC_OK;
}

static int launch_stream(irbc_ctx *c)
{
    int rc = launch_bulk(c, nullptr, nullptr, false);
    if (rc) return rc;
    return launch_near(c);
}

/* this step's triangle kinematics and bounding spheres from its vertex
 * kinematics; refresh: first the vertex kinematics from the vertex arrays
 * (else the previous step's membrane gather or this step's collision wrote
 * them) */
static int launch_tri_kin(irbc_ctx *c, int refresh)
{
    const MemDev M = c->mem(c->par);
    if (refresh)
        k_vkin<<<dim3((unsigned)((c->z.maxnv + IRBC_WG - 1) / IRBC_WG), (unsigned)c->nrep), IRBC_WG, 0, c->st>>>(
            c->d_rep, M);
    k_tri_kin<<<dim3((unsigned)((c->z.maxnt + IRBC_WG - 1) / IRBC_WG), (unsigned)c->nrep), IRBC_WG, 0, c->st>>>(
        c->d_rep, M, c->par & 1);
    LAUNCHCHK();
    return IRBC_OK;
}

/* The membrane part of one step on stream `s`, as a list of launches so
 * that the time loop can interleave them with the solvent chain's on the
 * host (both chains then start at once).  The next step's vertex
 * kinematics go to the other parity's buffer.  Returns the number of
 * launches; launch k of them with membrane_launch(c, phases, s, k). */
static int membrane_launches(const irbc_ctx *c) { return (c->mem_chain || c->mem_fused) ? 1 : 5; }

/* working workgroups per replica of k_mem_chain: its phases for the
 * batch's largest mesh plus the finish */
static unsigned chain_blocks(const irbc_ctx *c)
{
    const unsigned vb = (unsigned)((c->z.maxnv + IRBC_WG - 1) / IRBC_WG);
    const unsigned eb = (unsigned)((c->z.maxne + IRBC_WG - 1) / IRBC_WG);
    const unsigned tb = (unsigned)((c->z.maxnt + IRBC_WG - 1) / IRBC_WG);
    const unsigned VPB = IRBC_WG / IRBC_GATHER_LANES; /* gather: vertices per workgroup */
    const unsigned gb = (unsigned)((c->z.maxnv + VPB - 1) / VPB);
    /* begin (none when the gather folds it in), elements, gather, finish */
    return (IRBC_CHAIN_FOLD_BEGIN ? 0u : vb) + (eb + tb + vb) + gb + 1;
}

static int membrane_launch(irbc_ctx *c, int phases, hipStream_t s, int k)
{
    const MemDev M = c->mem(c->par ^ 1);
    const unsigned nr = (unsigned)c->nrep;
    if (c->mem_chain) {
        auto *kern = c->chain_dev ? k_mem_chain<true> : k_mem_chain<false>;
        kern<<<dim3(IRBC_NXCD * chain_blocks(c), nr), IRBC_WG, 0, s>>>(c->d_rep, phases, M, c->par);
        LAUNCHCHK();
        return IRBC_OK;
    }
    if (c->mem_fused) {
        const int lds_nv = c->z.maxnv <= c->mem_lds_nv ? c->z.maxnv : 0;
        k_mem_step<<<nr, IRBC_MEM_WG, (size_t)32 * lds_nv, s>>>(c->d_rep, phases, M, c->par, lds_nv);
        LAUNCHCHK();
        return IRBC_OK;
    }
    const unsigned X = IRBC_MEM_XCD_AFFINE ? IRBC_NXCD : 1;
    const dim3 gv(X * (unsigned)((c->z.maxnv + IRBC_WG - 1) / IRBC_WG), nr);
    /* element blocks, then vertex blocks for the bond preparation */
    const dim3 ge(X * (unsigned)((std::max(c->z.maxne, c->z.maxnt) + IRBC_WG - 1) / IRBC_WG +
                                 (c->z.maxnv + IRBC_WG - 1) / IRBC_WG), nr);
    const dim3 gt(X * (unsigned)((c->z.maxnt + IRBC_WG - 1) / IRBC_WG), nr);
    switch (k) {
    case 0: k_mem_begin<<<gv, IRBC_WG, 0, s>>>(c->d_rep, phases, M); break;
    case 1:
        if (phases & (PH_VA | PH_FORCES | PH_BONDS)) k_mem_edges<<<ge, IRBC_WG, 0, s>>>(c->d_rep, phases, M);
        break;
    case 2:
        if (phases & (PH_VA | PH_FORCES)) k_mem_tris<<<gt, IRBC_WG, 0, s>>>(c->d_rep, phases, M);
        break;
    case 3:
        if (phases & (PH_FORCES | PH_EXTERNAL | PH_KICK2)) k_mem_gather<<<gv, IRBC_WG, 0, s>>>(c->d_rep, phases, M);
        break;
    default: k_mem_finish<<<nr, IRBC_FIN_WG, 0, s>>>(c->d_rep, phases, M, c->par); break;
    }
    LAUNCHCHK();
    return IRBC_OK;
}

static int launch_membrane(irbc_ctx *c, int phases, hipStream_t s)
{
    for (int k = 0; k < membrane_launches(c); ++k) {
        const int rc = membrane_launch(c, phases, s, k);
        if (rc) return rc;
    }
    return IRBC_OK;
}

static int ready(irbc_ctx *c, bool need_solvent)
{
    for (int r = 0; r < c->nrep; ++r) {
        if (!c->hr[r].mesh) return fail(IRBC_ERR_STATE, "replica %d has no mesh (irbc_set_mesh)", r);
        if (need_solvent && !c->hr[r].solvent)
            return fail(IRBC_ERR_STATE, "replica %d has no solvent (irbc_set_solvent)", r);
    }
    return IRBC_OK;
}

int irbc_collision_step(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->primed = false;
    int rc = ready(c, true);
    if (rc) return rc;
    rc = launch_collision(c, 1, 1, COL_SORT);
    if (rc) return rc;
    return irbc_sync(c);
}

int irbc_llist_update(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->primed = false;
    int rc = ready(c, false);
    if (rc) return rc;
    rc = launch_llist(c, 1);
    if (rc) return rc;
    return irbc_sync(c);
}

int irbc_stream(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->primed = false;
    int rc = ready(c, true);
    if (rc) return rc;
    int rc2 = launch_tri_kin(c, 1);
    if (rc2) return rc2;
    rc2 = launch_cell_mask(c, 1);
    if (rc2) return rc2;
    rc2 = launch_stream(c);
    if (rc2) return rc2;
    rc2 = launch_membrane(c, PH_RESET_NEAR, c->st);
    if (rc2) return rc2;
    return irbc_sync(c);
}

int irbc_classify_solvent(irbc_ctx *c)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->primed = false;
    int rc = ready(c, true);
    if (rc) return rc;
    if (c->classify_all) {
        k_classify<<<(unsigned)c->z.tiles, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, c->view(c->par), c->mix,
                                                               c->miy);
        LAUNCHCHK();
        return irbc_sync(c);
    }
    /* triangles binned by their (y, z) box: a set-up call, so its
     * temporaries are allocated here and freed at the end */
    const int n = c->nrep;
    std::vector<int64_t> soff(n), eoff(n);
    int64_t nbt = 0;
    for (int r = 0; r < n; ++r) {
        soff[r] = nbt;
        nbt += (int64_t)c->rep[r].Ly * (c->rep[r].Lz + 2);
    }
    int64_t *d_soff = nullptr, *d_eoff = nullptr;
    int32_t *d_cnt = nullptr, *d_start = nullptr, *d_ent = nullptr;
    /* stream-ordered, so that set-up beside a running context never waits
     * for that context's work */
    auto release = [&]() {
        {
            std::lock_guard<std::mutex> lock(g_capture_mu);
            for (void *p : {(void *)d_soff, (void *)d_eoff, (void *)d_cnt, (void *)d_start, (void *)d_ent})
                if (p) hipFreeAsync(p, c->st);
        }
        hipStreamSynchronize(c->st);
    };
    auto dmalloc = [&](void **p, size_t bytes) {
        std::lock_guard<std::mutex> lock(g_capture_mu);
        return hipMallocAsync(p, bytes, c->st);
    };
    if (dmalloc((void **)&d_soff, n * 8) != hipSuccess || dmalloc((void **)&d_eoff, n * 8) != hipSuccess ||
        dmalloc((void **)&d_cnt, nbt * 4) != hipSuccess || dmalloc((void **)&d_start, (nbt + n) * 4) != hipSuccess) {
        release();
        return fail(IRBC_ERR_OOM, "irbc_classify_solvent: no room for the triangle bins");
    }
    rc = IRBC_OK;
    std::vector<int32_t> start(nbt + n);
    const dim3 g((unsigned)((c->z.maxnt + IRBC_WG - 1) / IRBC_WG), (unsigned)n);
    const MemView V = c->view(c->par);
    if (hipMemcpyAsync(d_soff, soff.data(), n * 8, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemsetAsync(d_cnt, 0, nbt * 4, c->st) != hipSuccess)
        rc = fail(IRBC_ERR_HIP, "irbc_classify_solvent: set-up copies failed");
    if (!rc) {
        k_cls_bins<<<g, IRBC_WG, 0, c->st>>>(c->d_rep, V, 0, d_soff, d_cnt, nullptr, nullptr, nullptr);
        k_cls_scan<<<(unsigned)n, IRBC_WG, 0, c->st>>>(c->d_rep, d_soff, d_cnt, d_start);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->st) != hipSuccess ||
            hipMemcpyAsync(start.data(), d_start, (nbt + n) * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess)
            rc = fail(IRBC_ERR_HIP, "irbc_classify_solvent: counting the bins failed");
    }
    int64_t tot = 0;
    for (int r = 0; r < n && !rc; ++r) {
        eoff[r] = tot;
        tot += start[soff[r] + r + (int64_t)c->rep[r].Ly * (c->rep[r].Lz + 2)];
    }
    if (!rc && (dmalloc((void **)&d_ent, std::max<int64_t>(tot, 1) * 4) != hipSuccess))
        rc = fail(IRBC_ERR_OOM, "irbc_classify_solvent: no room for the bin entries");
    if (!rc && hipMemcpyAsync(d_eoff, eoff.data(), n * 8, hipMemcpyHostToDevice, c->st) != hipSuccess)
        rc = fail(IRBC_ERR_HIP, "irbc_classify_solvent: set-up copies failed");
    if (!rc) {
        k_cls_bins<<<g, IRBC_WG, 0, c->st>>>(c->d_rep, V, 1, d_soff, d_cnt, d_start, d_eoff, d_ent);
        k_classify_binned<<<(unsigned)c->z.tiles, IRBC_WG, 0, c->st>>>(c->d_rep, c->d_tile_rep, V, c->mix, c->miy,
                                                                      d_soff, d_start, d_eoff, d_ent);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->st) != hipSuccess)
            rc = fail(IRBC_ERR_HIP, "irbc_classify_solvent: binned classification failed");
    }
    release();
    if (rc) return rc;
    return irbc_sync(c);
}

int irbc_membrane_phases(irbc_ctx *c, int phases)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    if (phases & ~(255 | IRBC_PH_PARASITE | IRBC_PH_MEMBRANE)) return fail(IRBC_ERR_ARG, "unknown phase bits 0x%x", phases);
    /* the public force bits onto the kernels' two force phases */
    int dev = phases & 255 & ~PH_FORCES;
    if (phases & (IRBC_PH_FORCES | IRBC_PH_MEMBRANE)) dev |= PH_FORCES;
    if (phases & (IRBC_PH_FORCES | IRBC_PH_PARASITE)) dev |= PH_WCA;
    c->primed = false;
    int rc = ready(c, false);
    if (rc) return rc;
    rc = launch_membrane(c, dev, c->st);
    if (rc) return rc;
    if (phases & PH_COUNTERS) {
        for (int r = 0; r < c->nrep; ++r) advance_host(c->hr[r].spec.params, c->hr[r], c->align_lists);
        c->par ^= 1;
    }
    return irbc_sync(c);
}

int irbc_membrane_va(irbc_ctx *c, int r, double *area, double *volume)
{
    API_GUARD;
    CHECK_R(c, r);
    int rc = irbc_membrane_phases(c, PH_VA);
    if (rc) return rc;
    if (area) *area = c->rep[r].Sarea;
    if (volume) *volume = c->rep[r].Volume;
    return IRBC_OK;
}

int irbc_membrane_forces(irbc_ctx *c) { return irbc_membrane_phases(c, PH_VA | PH_FORCES | PH_EXTERNAL); }

int irbc_bond_step(irbc_ctx *c) { return irbc_membrane_phases(c, PH_BONDS); }

int irbc_get_cells(irbc_ctx *c, int r, double *vcm, int32_t *nc, double *ncc, double *rand1,
                   double *rand2, double *dvv, double *the)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    const size_t n = R.ncell;
    HIPCHK(c->sync_both());
    if (vcm) HIPCHK(hipMemcpy(vcm, c->diag.vcm + 3 * R.coff, 3 * n * 8, hipMemcpyDeviceToHost));
    if (nc) HIPCHK(hipMemcpy(nc, c->diag.nc + R.coff, n * 4, hipMemcpyDeviceToHost));
    double *dd[] = {ncc, rand1, rand2, dvv, the};
    double *ds[] = {c->diag.ncc, c->diag.rand1, c->diag.rand2, c->diag.dvv, c->diag.the};
    for (int k = 0; k < 5; ++k)
        if (dd[k]) HIPCHK(hipMemcpy(dd[k], ds[k] + R.coff, n * 8, hipMemcpyDeviceToHost));
    return IRBC_OK;
}

int64_t irbc_get_llist(irbc_ctx *c, int r, int64_t *off, int32_t *ent, int64_t cap)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    const size_t n = R.ncell;
    HIPCHK(c->sync_both());
    std::vector<int32_t> o(n + 1);
    HIPCHK(hipMemcpy(o.data(), c->lloff + R.coff1, (n + 1) * 4, hipMemcpyDeviceToHost));
    const int64_t tot = o[n];
    if (off)
        for (size_t i = 0; i <= n; ++i) off[i] = o[i];
    if (ent && tot > 0) HIPCHK(hipMemcpy(ent, c->llent + R.llent_off, std::min(tot, cap) * 4, hipMemcpyDeviceToHost));
    return tot;
}

int64_t irbc_get_pairs(irbc_ctx *c, int r, int32_t *cnt, int32_t *lig, int64_t cap)
{
    API_GUARD;
    CHECK_R(c, r);
    const DevRep &R = c->rep[r];
    const int nv = R.nv;
    HIPCHK(c->sync_both());
    std::vector<int32_t> nc(nv), cd((size_t)nv * IRBC_MAX_CAND);
    HIPCHK(hipMemcpy(nc.data(), c->ncand + R.voff, nv * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cd.data(), c->cand + (int64_t)R.voff * IRBC_MAX_CAND, cd.size() * 4, hipMemcpyDeviceToHost));
    int64_t t = 0;
    for (int v = 0; v < nv; ++v) {
        if (cnt) cnt[v] = nc[v];
        for (int k = 0; k < nc[v]; ++k) {
            if (lig && t < cap) lig[t] = cd[(size_t)v * IRBC_MAX_CAND + k];
            ++t;
        }
    }
    return t;
}

/* ---------------- launch-shape autotuner ---------------- */

/* The whole device state of a context (every allocation) and the host
 * mirrors, copied on the device so that trial steps can be undone. */
struct Snapshot {
    std::vector<void *> buf;
    std::vector<HostRep> hr;
    std::vector<DevRep> rep;
    int par = 0;
};

static int snap_take(irbc_ctx *c, Snapshot &S)
{
    HIPCHK(c->sync_both());
    if (S.buf.empty()) {
        for (size_t i = 0; i < c->allocs.size(); ++i) {
            void *q = nullptr;
            if (hipMalloc(&q, c->alloc_bytes[i]) != hipSuccess) {
                for (void *b : S.buf) hipFree(b);
                S.buf.clear();
                return fail(IRBC_ERR_OOM, "autotune: no room for a state snapshot");
            }
            S.buf.push_back(q);
        }
    }
    for (size_t i = 0; i < c->allocs.size(); ++i)
        HIPCHK(hipMemcpyAsync(S.buf[i], c->allocs[i], c->alloc_bytes[i], hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    S.hr = c->hr;
    S.rep = c->rep;
    S.par = c->par;
    return IRBC_OK;
}

static int snap_restore(irbc_ctx *c, const Snapshot &S)
{
    HIPCHK(c->sync_both());
    for (size_t i = 0; i < c->allocs.size(); ++i)
        HIPCHK(hipMemcpyAsync(c->allocs[i], S.buf[i], c->alloc_bytes[i], hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->hr = S.hr;
    c->rep = S.rep;
    c->par = S.par;
    c->primed = false;
    return IRBC_OK;
}

/* launch options from an index: bit 0 one-workgroup membrane, 1 merged
 * bounce-back, 2 graphs, 3 XCD chunking, 4 four particles per lane, 5 capped
 * streaming grid, 6 membrane after this step's streaming, 7 membrane chain,
 * 8 one queue */
static void apply_options(irbc_ctx *c, int k)
{
    c->mem_fused = (k >> 0) & 1;
    c->near_merged = (k >> 1) & 1;
    c->use_graphs = (k >> 2) & 1;
    c->xcd_chunk = (k >> 3) & 1;
    c->stream4 = (k >> 4) & 1;
    c->stream_grid = (k >> 5) & 1;
    c->mem_after = (k >> 6) & 1;
    c->mem_chain = (k >> 7) & 1;
    c->one_queue = (k >> 8) & 1;
}

/* The indices the autotuner times: the membrane as the chain on one queue,
 * the chain on its own stream, or the five grid kernels on its own stream,
 * times bounce-back shape, graphs, XCD chunking and stream shape (48).  The
 * one-workgroup membrane, the capped streaming grid and the late membrane
 * start never won and are left to irbc_set_option. */
static bool tune_candidate(int k)
{
    if (k & ((1 << 0) | (1 << 5) | (1 << 6))) return false;
    const bool chain = (k >> 7) & 1, oneq = (k >> 8) & 1;
    return chain || !oneq;
}

#ifndef IRBC_TUNE_REFINE
#define IRBC_TUNE_REFINE 4 /* candidates the autotuner times again over longer runs */
#endif
int irbc_autotune(irbc_ctx *c, int32_t steps, double *report)
{
    std::vector<double> own(IRBC_TUNE_CANDIDATES);
    if (!report) report = own.data();
    if (!c || steps < 1) return fail(IRBC_ERR_ARG, "irbc_autotune: bad arguments");
    int rc = ready(c, true);
    if (rc) return rc;
    if (c->prof) return fail(IRBC_ERR_STATE, "irbc_autotune: switch event profiling off first");
    Snapshot S;
    rc = snap_take(c, S);
    if (rc) return rc;
    double best = 1e300;
    int best_k = 0;
    for (int k = 0; k < IRBC_TUNE_CANDIDATES; ++k) {
        if (!tune_candidate(k)) {
            if (report) report[k] = std::nan("");
            continue;
        }
        apply_options(c, k);
        if (!rc) rc = snap_restore(c, S);
        if (!rc) rc = irbc_md_steps(c, steps); /* warm-up; also captures graphs */
        double ms = 0.0;
        if (!rc) {
            rc = snap_restore(c, S);
            hipEvent_t e0 = c->event(), e1 = c->event();
            if (!rc) rc = hipEventRecord(e0, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
            if (!rc) rc = irbc_md_steps(c, steps);
            if (!rc) rc = hipEventRecord(e1, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
            if (!rc) rc = c->wait_all();
            float f = 0.f;
            if (!rc && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(IRBC_ERR_HIP, "autotune timing");
            ms = f;
            c->ev_pool.push_back(e0);
            c->ev_pool.push_back(e1);
        }
        if (rc) break;
        if (report) report[k] = ms / steps;
        if (ms < best) {
            best = ms;
            best_k = k;
        }
    }
    int64_t n2 = steps; /* the later stages' runs: at least four collisions */
    for (const HostRep &h : c->hr) n2 = std::max<int64_t>(n2, 4 * (int64_t)h.spec.params.csteps);
    n2 = std::min<int64_t>(n2, IRBC_TUNE_WARM_MAX);
    /* the fastest few again, over n2 steps, twice in turn: one short run
     * each leaves shapes within a few per cent of each other to chance
     * (on the 64x64x48 replica the XCD-chunked collision lost to noise) */
    if (!rc && report) {
        std::vector<std::pair<double, int>> order;
        for (int k = 0; k < IRBC_TUNE_CANDIDATES; ++k)
            if (tune_candidate(k) && report[k] == report[k]) order.push_back({report[k], k});
        std::sort(order.begin(), order.end());
        if (order.size() > (size_t)IRBC_TUNE_REFINE) order.resize(IRBC_TUNE_REFINE);
        std::vector<double> tot(order.size(), 0.0);
        for (int round = 0; round < 2 && !rc && order.size() > 1; ++round)
            for (size_t j = 0; j < order.size() && !rc; ++j) {
                apply_options(c, order[j].second);
                rc = snap_restore(c, S);
                if (!rc) rc = irbc_md_steps(c, n2); /* warm-up; also captures graphs */
                if (!rc) rc = snap_restore(c, S);
                hipEvent_t e0 = c->event(), e1 = c->event();
                if (!rc) rc = hipEventRecord(e0, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
                if (!rc) rc = irbc_md_steps(c, n2);
                if (!rc) rc = hipEventRecord(e1, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
                if (!rc) rc = c->wait_all();
                float f = 0.f;
                if (!rc && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(IRBC_ERR_HIP, "autotune timing");
                c->ev_pool.push_back(e0);
                c->ev_pool.push_back(e1);
                tot[j] += f;
            }
        if (!rc && order.size() > 1) {
            size_t bj = 0;
            for (size_t j = 1; j < order.size(); ++j)
                if (tot[j] < tot[bj]) bj = j;
            best_k = order[bj].second;
        }
    }
    apply_options(c, best_k);
    /* then, on the chosen launch shape: the bin-sort tile shape (tiles per
     * binning workgroup), timed over n2 steps */
    if (!rc && c->sort_every == 1) {
        double best2 = 1e300;
        int best_span = 1;
        for (int sp : {1, 2, 4}) {
            c->bin_span = sp;
            double ms = 0.0;
            rc = snap_restore(c, S);
            if (!rc) rc = irbc_md_steps(c, n2); /* warm-up; also captures graphs */
            if (!rc) rc = snap_restore(c, S);
            hipEvent_t e0 = c->event(), e1 = c->event();
            if (!rc) rc = hipEventRecord(e0, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
            if (!rc) rc = irbc_md_steps(c, n2);
            if (!rc) rc = hipEventRecord(e1, c->st) == hipSuccess ? IRBC_OK : IRBC_ERR_HIP;
            if (!rc) rc = c->wait_all();
            float f = 0.f;
            if (!rc && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(IRBC_ERR_HIP, "autotune timing");
            ms = f;
            c->ev_pool.push_back(e0);
            c->ev_pool.push_back(e1);
            if (rc) break;
            c->span_ms[sp == 1 ? 0 : sp == 2 ? 1 : 2] = ms / n2;
            if (ms < best2) {
                best2 = ms;
                best_span = sp;
            }
        }
        c->bin_span = best_span;
    }
    if (!rc && c->use_graphs) {
        /* capture the chosen shape's graphs for one whole collide/list-update
         * cycle now, so that later runs replay instead of capturing */
        int64_t cyc = 20;
        for (const HostRep &h : c->hr)
            for (int64_t p : {(int64_t)h.spec.params.csteps, (int64_t)h.spec.params.uplist})
                if (p > 0 && cyc <= IRBC_TUNE_WARM_MAX) cyc = cyc / std::gcd(cyc, p) * p;
        rc = snap_restore(c, S);
        if (!rc) rc = irbc_md_steps(c, std::min<int64_t>(cyc, IRBC_TUNE_WARM_MAX) + IRBC_GRAPH_STEPS);
        if (!rc) rc = c->wait_all();
    }
    int rc2 = snap_restore(c, S);
    for (void *b : S.buf) hipFree(b);
    if (rc) return rc;
    if (rc2) return rc2;
    return best_k;
}

/* ---------------- checkpoint / resume ---------------- */

/* A replica's resumable state: what the next step reads and cannot derive.
 * Triangle kinematics, near queues and the grid shift are recomputed at the
 * start of irbc_md_steps; the cell->triangle lists are saved because they
 * belong to the last list update, not to the current positions. */
struct CkptHeader {
    char magic[8];
    int32_t version, tapes;
    irbc_params params;
    int64_t ns, ncell, llen;
    int32_t nv, ne, nt, nlig;
    int64_t step;
    int32_t colsteps, upd;
    uint64_t seed;
    double wallspeed;
};

static int64_t ckpt_bytes(const irbc_ctx *c, int r, int64_t llen)
{
    const HostRep &h = c->hr[r];
    const int64_t ns = h.spec.n_solvent, nv = h.spec.n_vertices;
    const int64_t ncell = c->rep[r].ncell, nlig = c->rep[r].nlig;
    return (int64_t)sizeof(CkptHeader) + ns * (6 * 4 + 4) + nv * (9 * 8 + 6 * 4) + nlig * 4 + 12 * 8 + 4 * 4 +
           (ncell + 1) * 4 + llen * 4;
}

static int64_t llist_len(irbc_ctx *c, int r)
{
    const DevRep &R = c->rep[r];
    int32_t a = 0, b = 0;
    if (hipMemcpy(&a, c->lloff + R.coff1, 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&b, c->lloff + R.coff1 + R.ncell, 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(IRBC_ERR_HIP, "checkpoint: reading the cell lists failed");
    return (int64_t)b - a;
}

int64_t irbc_checkpoint_bytes(irbc_ctx *c, int r)
{
    API_GUARD;
    CHECK_R(c, r);
    HIPCHK(c->sync_both());
    const int64_t llen = llist_len(c, r);
    if (llen < 0) return llen;
    return ckpt_bytes(c, r, llen);
}

int64_t irbc_checkpoint_save(irbc_ctx *c, int r, void *buf, int64_t cap)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!buf) return fail(IRBC_ERR_ARG, "irbc_checkpoint_save: null buffer");
    int rc = c->pull(r);
    if (rc) return rc;
    const DevRep &R = c->rep[r];
    if (R.rotate_pending) return fail(IRBC_ERR_STATE, "irbc_checkpoint_save: a collision's rotation is pending");
    const int64_t llen = llist_len(c, r);
    if (llen < 0) return llen;
    const int64_t need = ckpt_bytes(c, r, llen);
    if (cap < need) return fail(IRBC_ERR_ARG, "irbc_checkpoint_save: needs %lld bytes", (long long)need);
    const HostRep &h = c->hr[r];
    CkptHeader H;
    std::memset(&H, 0, sizeof H);
    std::memcpy(H.magic, "IRBCCKP1", 8);
    H.version = 1;
    H.tapes = R.tapes;
    H.params = h.spec.params;
    H.ns = h.spec.n_solvent;
    H.ncell = R.ncell;
    H.llen = llen;
    H.nv = R.nv;
    H.ne = R.ne;
    H.nt = R.nt;
    H.nlig = R.nlig;
    H.step = h.step;
    H.colsteps = h.colsteps;
    H.upd = h.upd;
    H.seed = R.seed;
    H.wallspeed = R.wall2[c->par & 1];
    char *p = (char *)buf;
    std::memcpy(p, &H, sizeof H);
    p += sizeof H;
    const float *sf[] = {c->xa, c->ya, c->za, c->vxa, c->vya, c->vza};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpy(p, sf[k] + R.poff, H.ns * 4, hipMemcpyDeviceToHost));
        p += H.ns * 4;
    }
    HIPCHK(hipMemcpy(p, c->ida + R.poff, H.ns * 4, hipMemcpyDeviceToHost));
    p += H.ns * 4;
    const double *vd[] = {c->mx, c->my, c->mz, c->mvx, c->mvy, c->mvz, c->mfx, c->mfy, c->mfz};
    for (int k = 0; k < 9; ++k) {
        HIPCHK(hipMemcpy(p, vd[k] + R.voff, (size_t)H.nv * 8, hipMemcpyDeviceToHost));
        p += (size_t)H.nv * 8;
    }
    const int32_t *vi[] = {c->mix, c->miy, c->mpid, c->bondi, c->bondl, c->knob};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpy(p, vi[k] + R.voff, (size_t)H.nv * 4, hipMemcpyDeviceToHost));
        p += (size_t)H.nv * 4;
    }
    HIPCHK(hipMemcpy(p, c->ligi + R.loff, (size_t)H.nlig * 4, hipMemcpyDeviceToHost));
    p += (size_t)H.nlig * 4;
    double pd[12];
    int32_t pi[4] = {0, 0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        pd[k] = R.para[k];
        pd[3 + k] = R.paraV[k];
        pd[6 + k] = R.paraft[k];
        pd[9 + k] = R.parafdt[k];
    }
    pi[0] = R.paradx;
    pi[1] = R.parady;
    pi[2] = R.paraid;
    std::memcpy(p, pd, sizeof pd);
    p += sizeof pd;
    std::memcpy(p, pi, sizeof pi);
    p += sizeof pi;
    std::vector<int32_t> off(H.ncell + 1);
    HIPCHK(hipMemcpy(off.data(), c->lloff + R.coff1, (H.ncell + 1) * 4, hipMemcpyDeviceToHost));
    const int32_t o0 = off[0];
    for (int32_t &o : off) o -= o0;
    std::memcpy(p, off.data(), (H.ncell + 1) * 4);
    p += (H.ncell + 1) * 4;
    if (llen > 0) HIPCHK(hipMemcpy(p, c->llent + R.llent_off + o0, llen * 4, hipMemcpyDeviceToHost));
    p += llen * 4;
    return (int64_t)(p - (char *)buf);
}

/* Everything in a checkpoint that device code uses as an index is checked
 * on the host before any of it is copied: particle ids (a permutation),
 * vertex cells, bonds, knobs, ligand occupancy, the parasite's cell and the
 * cell->triangle lists. */
static int ckpt_check(const irbc_ctx *c, int r, const CkptHeader &H, const char *p)
{
    const DevRep &R = c->rep[r];
    const HostRep &h = c->hr[r];
    if (!same_params(H.params, h.spec.params))
        return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: replica %d has other parameters than the checkpoint", r);
    if (!counters_valid(h.spec.params, H.step, H.colsteps, H.upd) || (H.tapes != 0 && H.tapes != 1))
        return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: counters out of range");
    const int64_t ns = H.ns, nv = H.nv, nlig = H.nlig, ncell = H.ncell;
    const char *q = p + 6 * ns * 4;
    std::vector<uint8_t> seen(ns, 0);
    for (int64_t i = 0; i < ns; ++i) {
        uint32_t w;
        std::memcpy(&w, q + 4 * i, 4);
        const uint32_t id = w & 0x7FFFFFFFu;
        if ((int64_t)id >= ns || seen[id]) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: particle ids are not a permutation");
        seen[id] = 1;
    }
    q += ns * 4 + 9 * nv * 8;
    std::vector<int32_t> vi(6 * nv);
    std::memcpy(vi.data(), q, (size_t)6 * nv * 4);
    q += (size_t)6 * nv * 4;
    for (int64_t v = 0; v < nv; ++v) {
        const int32_t pid = vi[2 * nv + v], bi = vi[3 * nv + v], bl = vi[4 * nv + v], kb = vi[5 * nv + v];
        if (pid < 0 || pid >= ncell || (bi != 0 && bi != 1) || (kb != 0 && kb != 1) || (bi == 1 && (bl < 0 || bl >= nlig)))
            return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: vertex %lld: cell, bond or knob out of range", (long long)v);
    }
    for (int64_t l = 0; l < nlig; ++l) {
        int32_t o;
        std::memcpy(&o, q + 4 * l, 4);
        if (o != 0 && o != 1) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: ligand occupancy out of range");
    }
    q += nlig * 4 + 12 * 8;
    int32_t pi[4];
    std::memcpy(pi, q, sizeof pi);
    if (pi[2] < 0 || pi[2] >= ncell) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: parasite cell out of range");
    q += sizeof pi;
    std::vector<int32_t> off(ncell + 1);
    std::memcpy(off.data(), q, (ncell + 1) * 4);
    q += (ncell + 1) * 4;
    if (off[0] != 0 || off[ncell] != H.llen) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: cell lists do not add up");
    for (int64_t k = 0; k < ncell; ++k)
        if (off[k + 1] < off[k]) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: cell list offsets decrease");
    for (int64_t e = 0; e < H.llen; ++e) {
        int32_t t;
        std::memcpy(&t, q + 4 * e, 4);
        if (t < 0 || t >= H.nt) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: cell list entry out of range");
    }
    (void)R;
    return IRBC_OK;
}

int irbc_checkpoint_load(irbc_ctx *c, int r, const void *buf, int64_t n)
{
    API_GUARD;
    CHECK_R(c, r);
    if (!buf || n < (int64_t)sizeof(CkptHeader)) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: short buffer");
    CkptHeader H;
    std::memcpy(&H, buf, sizeof H);
    const HostRep &h = c->hr[r];
    if (std::memcmp(H.magic, "IRBCCKP1", 8) != 0 || H.version != 1)
        return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: not a checkpoint of this format");
    if (!h.mesh) return fail(IRBC_ERR_STATE, "irbc_checkpoint_load: load the replica's mesh first");
    if (H.ns != h.spec.n_solvent || H.nv != h.spec.n_vertices || H.ne != h.spec.n_edges ||
        H.nt != h.spec.n_triangles || H.ncell != c->rep[r].ncell || H.nlig != c->rep[r].nlig)
        return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: replica %d has other sizes than the checkpoint", r);
    if (n < ckpt_bytes(c, r, H.llen)) return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: truncated buffer");
    if (H.llen < 0 || H.llen > c->rep[r].llent_cap)
        return fail(IRBC_ERR_ARG, "irbc_checkpoint_load: cell lists exceed capacity");
    int rc = ckpt_check(c, r, H, (const char *)buf + sizeof H);
    if (rc) return rc;
    rc = c->pull(r);
    if (rc) return rc;
    DevRep &R = c->rep[r];
    const char *p = (const char *)buf + sizeof H;
    float *sf[] = {c->xa, c->ya, c->za, c->vxa, c->vya, c->vza};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpy(sf[k] + R.poff, p, H.ns * 4, hipMemcpyHostToDevice));
        p += H.ns * 4;
    }
    HIPCHK(hipMemcpy(c->ida + R.poff, p, H.ns * 4, hipMemcpyHostToDevice));
    p += H.ns * 4;
    double *vd[] = {c->mx, c->my, c->mz, c->mvx, c->mvy, c->mvz, c->mfx, c->mfy, c->mfz};
    for (int k = 0; k < 9; ++k) {
        HIPCHK(hipMemcpy(vd[k] + R.voff, p, (size_t)H.nv * 8, hipMemcpyHostToDevice));
        p += (size_t)H.nv * 8;
    }
    int32_t *vi[] = {c->mix, c->miy, c->mpid, c->bondi, c->bondl, c->knob};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpy(vi[k] + R.voff, p, (size_t)H.nv * 4, hipMemcpyHostToDevice));
        p += (size_t)H.nv * 4;
    }
    HIPCHK(hipMemcpy(c->ligi + R.loff, p, (size_t)H.nlig * 4, hipMemcpyHostToDevice));
    p += (size_t)H.nlig * 4;
    double pd[12];
    int32_t pi[4];
    std::memcpy(pd, p, sizeof pd);
    p += sizeof pd;
    std::memcpy(pi, p, sizeof pi);
    p += sizeof pi;
    for (int k = 0; k < 3; ++k) {
        R.para[k] = pd[k];
        R.paraV[k] = pd[3 + k];
        R.paraft[k] = pd[6 + k];
        R.parafdt[k] = pd[9 + k];
    }
    R.paradx = pi[0];
    R.parady = pi[1];
    R.paraid = pi[2];
    std::vector<int32_t> off(H.ncell + 1);
    std::memcpy(off.data(), p, (H.ncell + 1) * 4);
    p += (H.ncell + 1) * 4;
    HIPCHK(hipMemcpy(c->lloff + R.coff1, off.data(), (H.ncell + 1) * 4, hipMemcpyHostToDevice));
    if (H.llen > 0) HIPCHK(hipMemcpy(c->llent + R.llent_off, p, H.llen * 4, hipMemcpyHostToDevice));
    /* the particle order of the checkpoint, no pending rotation */
    HIPCHK(hipMemset(R.rec + 4 * R.poff, 0, (size_t)H.ns * 16));
    HIPCHK(hipMemset(R.rec1 + 4 * R.poff, 0, (size_t)H.ns * 16));
    R.rotate_pending = 0;
    R.key_src = 1;
    R.seed = H.seed;
    R.tapes = H.tapes;
    R.status = 0;
    R.near_count[0] = R.near_count[1] = R.defer_count[0] = R.defer_count[1] = 0;
    HostRep &hm = c->hr[r];
    hm.spec.seed = H.seed;
    hm.step = H.step;
    hm.colsteps = H.colsteps;
    hm.upd = H.upd;
    hm.solvent = true;
    top_of_step(hm.spec.params, H.step, H.colsteps, H.upd, R);
    R.wall2[0] = R.wall2[1] = H.wallspeed;
    return c->push(r);
}

/* A replica moved device to device: src's replica rs into dst's slot rd,
 * with everything a checkpoint holds, its mesh (topology, rest lengths,
 * rest areas, hinge orientation, force-slot incidence, A0, V0, theta0) and
 * its time series (cadence, capacity, records).  Same sizes and parameters
 * as a checkpoint load requires; dst needs no irbc_set_mesh first.  dst
 * then continues as src would have.  A request loop uses it to re-pack the
 * running replicas into a new context when requests arrive or finish;
 * nothing crosses the host but the descriptors and two list offsets. */
int irbc_copy_replica(irbc_ctx *dst, int rd, irbc_ctx *src, int rs)
{
    API_GUARD;
    CHECK_R(dst, rd);
    CHECK_R(src, rs);
    if (dst == src && rd == rs) return IRBC_OK;
    if (dst->device != src->device) return fail(IRBC_ERR_ARG, "irbc_copy_replica: contexts on different devices");
    const HostRep &hs = src->hr[rs];
    HostRep &hd = dst->hr[rd];
    if (!hs.mesh || !hs.solvent) return fail(IRBC_ERR_STATE, "irbc_copy_replica: the source replica is not set up");
    const irbc_replica_spec &a = hs.spec, &b = hd.spec;
    if (a.n_solvent != b.n_solvent || a.n_vertices != b.n_vertices || a.n_edges != b.n_edges ||
        a.n_triangles != b.n_triangles || src->rep[rs].ncell != dst->rep[rd].ncell ||
        src->rep[rs].nlig != dst->rep[rd].nlig)
        return fail(IRBC_ERR_ARG, "irbc_copy_replica: the replicas have other sizes");
    if (!same_params(a.params, b.params))
        return fail(IRBC_ERR_ARG, "irbc_copy_replica: the replicas have other parameters");
    int rc = src->pull(rs);
    if (rc) return rc;
    rc = dst->pull(rd);
    if (rc) return rc;
    const DevRep &S = src->rep[rs];
    DevRep &R = dst->rep[rd];
    if (S.rotate_pending) return fail(IRBC_ERR_STATE, "irbc_copy_replica: a collision's rotation is pending");
    const int64_t ns = a.n_solvent, nv = a.n_vertices, ne = a.n_edges, nt = a.n_triangles;
    const int64_t nlig = S.nlig, ncell = S.ncell;
    if (S.series_every > 0 && R.series_cap < S.series_cap) {
        double *buf = nullptr;
        rc = dst->alloc(buf, (int64_t)S.series_cap * IRBC_SERIES_REC);
        if (rc) return rc;
        R.series = buf; /* freed with the context */
        R.series_cap = S.series_cap;
    }
    int32_t ends[2] = {0, 0};
    HIPCHK(hipMemcpy(&ends[0], src->lloff + S.coff1, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&ends[1], src->lloff + S.coff1 + ncell, 4, hipMemcpyDeviceToHost));
    const int64_t llen = (int64_t)ends[1] - ends[0];
    if (ends[0] < 0 || llen < 0 || llen > R.llent_cap)
        return fail(IRBC_ERR_ARG, "irbc_copy_replica: cell lists exceed the destination's capacity");
    hipStream_t q = dst->st;
#define D2D(d_, s_, bytes_) HIPCHK(hipMemcpyAsync((d_), (s_), (size_t)(bytes_), hipMemcpyDeviceToDevice, q))
    /* solvent */
    float *const df[] = {dst->xa, dst->ya, dst->za, dst->vxa, dst->vya, dst->vza};
    const float *const sf[] = {src->xa, src->ya, src->za, src->vxa, src->vya, src->vza};
    for (int k = 0; k < 6; ++k) D2D(df[k] + R.poff, sf[k] + S.poff, ns * 4);
    D2D(dst->ida + R.poff, src->ida + S.poff, ns * 4);
    /* vertices, bonds and knobs */
    double *const dd[] = {dst->mx, dst->my, dst->mz, dst->mvx, dst->mvy, dst->mvz,
                          dst->mfx, dst->mfy, dst->mfz, dst->poff, dst->pon};
    const double *const sd[] = {src->mx, src->my, src->mz, src->mvx, src->mvy, src->mvz,
                                src->mfx, src->mfy, src->mfz, src->poff, src->pon};
    for (int k = 0; k < 11; ++k) D2D(dd[k] + R.voff, sd[k] + S.voff, nv * 8);
    int32_t *const di[] = {dst->mix, dst->miy, dst->mpid, dst->bondi, dst->bondl, dst->knob, dst->ncand};
    const int32_t *const si[] = {src->mix, src->miy, src->mpid, src->bondi, src->bondl, src->knob, src->ncand};
    for (int k = 0; k < 7; ++k) D2D(di[k] + R.voff, si[k] + S.voff, nv * 4);
    if (nlig > 0) D2D(dst->ligi + R.loff, src->ligi + S.loff, nlig * 4);
    /* the mesh: every index in these arrays is the replica's own */
    int32_t *const die[] = {dst->e1, dst->e2, dst->es1, dst->es2, dst->iedge};
    const int32_t *const sie[] = {src->e1, src->e2, src->es1, src->es2, src->iedge};
    for (int k = 0; k < 5; ++k) D2D(die[k] + R.eoff, sie[k] + S.eoff, ne * 4);
    D2D(dst->l0 + R.eoff, src->l0 + S.eoff, ne * 8);
    D2D(dst->kappap + R.eoff, src->kappap + S.eoff, ne * 8);
    int32_t *const dit[] = {dst->t1, dst->t2, dst->t3};
    const int32_t *const sit[] = {src->t1, src->t2, src->t3};
    for (int k = 0; k < 3; ++k) D2D(dit[k] + R.toff, sit[k] + S.toff, nt * 4);
    D2D(dst->tarea + R.toff, src->tarea + S.toff, nt * 8);
    D2D(dst->inc_off + R.inc_off, src->inc_off + S.inc_off, (nv + 1) * 4);
    D2D(dst->inc_ent + R.inc_ent_off, src->inc_ent + S.inc_ent_off, inc_bound(a) * 4);
    D2D(dst->tag + R.scr_off, src->tag + S.scr_off, scr_bound(a) * 8);
    /* the cell->triangle lists of the last update */
    if (ends[0] == 0) {
        D2D(dst->lloff + R.coff1, src->lloff + S.coff1, (ncell + 1) * 4);
    } else {
        /* offsets relative to the replica's first entry */
        std::vector<int32_t> off(ncell + 1);
        HIPCHK(hipMemcpy(off.data(), src->lloff + S.coff1, (ncell + 1) * 4, hipMemcpyDeviceToHost));
        for (int32_t &o : off) o -= ends[0];
        HIPCHK(hipMemcpyAsync(dst->lloff + R.coff1, off.data(), (ncell + 1) * 4, hipMemcpyHostToDevice, q));
        HIPCHK(hipStreamSynchronize(q));
    }
    if (llen > 0) D2D(dst->llent + R.llent_off, src->llent + S.llent_off + ends[0], llen * 4);
    if (S.series_every > 0 && S.series_n > 0)
        D2D(R.series, S.series, (int64_t)std::min(S.series_n, S.series_cap) * IRBC_SERIES_REC * 8);
    /* the particle order of the source, no pending rotation */
    HIPCHK(hipMemsetAsync(R.rec + 4 * R.poff, 0, (size_t)ns * 16, q));
    HIPCHK(hipMemsetAsync(R.rec1 + 4 * R.poff, 0, (size_t)ns * 16, q));
#undef D2D
    R.Sarea0 = S.Sarea0;
    R.Volume0 = S.Volume0;
    R.theta0 = S.theta0;
    R.cos_t0 = S.cos_t0;
    R.sin_t0 = S.sin_t0;
    for (int k = 0; k < 3; ++k) {
        R.para[k] = S.para[k];
        R.paraV[k] = S.paraV[k];
        R.paraft[k] = S.paraft[k];
        R.parafdt[k] = S.parafdt[k];
    }
    R.paradx = S.paradx;
    R.parady = S.parady;
    R.paraid = S.paraid;
    R.rotate_pending = 0;
    R.key_src = 1;
    R.seed = S.seed;
    R.tapes = S.tapes;
    R.status = 0;
    R.near_count[0] = R.near_count[1] = R.defer_count[0] = R.defer_count[1] = 0;
    R.series_every = S.series_every;
    R.series_n = S.series_every > 0 ? S.series_n : 0;
    hd.e1 = hs.e1;
    hd.e2 = hs.e2;
    hd.t1 = hs.t1;
    hd.t2 = hs.t2;
    hd.t3 = hs.t3;
    hd.mesh = true;
    hd.spec.seed = hs.spec.seed;
    hd.step = hs.step;
    hd.colsteps = hs.colsteps;
    hd.upd = hs.upd;
    hd.solvent = true;
    top_of_step(hd.spec.params, hs.step, hs.colsteps, hs.upd, R);
    R.wall2[0] = R.wall2[1] = S.wall2[src->par & 1];
    return dst->push(rd);
}

/* ---------------- the time loop ---------------- */

/* Launches of one MD step s (parity p).  Two chains: the solvent chain
 * (collision, cell lists, triangle kinematics, streaming, bounce-back) on
 * c->st and the membrane chain on c->st2.
 *  - The solvent chain of step s waits for the membrane of step s-1: it
 *    reads the vertex kinematics and the drift margin that membrane wrote
 *    (parity p), and a collision or list update reads the vertices.
 *  - The membrane of step s writes parity p^1 (vertex kinematics, drift
 *    margin, next wall speed), last read by the solvent chain of step s-1
 *    up to its bulk streaming.  So on a step without a collision or list
 *    update it waits only for that streaming (ev_str of step s-1) and runs
 *    beside the bounce-back of step s-1 and the solvent chain of step s.
 *  - On a collision or list-update step, and on the first step of a run or
 *    captured chunk, the membrane waits for this step's collision, lists and
 *    near masks (ev_pre).
 * ev_mem and ev_str alternate by parity, so each wait binds to the step it
 * means.  `first` skips the wait for the previous step's membrane (the
 * caller has joined the chains), `rec` brackets phases with timing events.
 * join_chains() makes the solvent stream wait for the last membrane. */
static int launch_step(irbc_ctx *c, int cf, bool ll, bool rec, bool first)
{
    int rc;
    const bool col = cf != 0;
    const int p = c->par & 1;
    const int mem_phases = PH_ALL & ~PH_RESET_NEAR; /* k_stream clears the other parity's queue */
    const bool lock = col || ll || first;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!first) HIPCHK(hipStreamWaitEvent(c->st, c->ev_mem2[p ^ 1], 0));
    if (col) {
        if (rec) { e0 = c->event(); hipEventRecord(e0, c->st); }
        rc = launch_collision(c, 0, c->diag_on ? 1 : 0, cf);
        if (rc) return rc;
        if (rec) { e1 = c->event(); hipEventRecord(e1, c->st); c->pend[1].push_back({e0, e1}); }
    }
    if (ll) {
        rc = launch_llist(c, 0);
        if (rc) return rc;
    }
    /* replicas that collided or made lists get new near masks
     * (from this step's triangle spheres), on which the membrane finish's
     * drift margin builds */
    if (col || ll) {
        rc = launch_tri_kin(c, 0);
        if (rc) return rc;
        rc = launch_cell_mask(c, 0);
        if (rc) return rc;
    }
    /* mem_after: the membrane waits for this step's bulk streaming instead
     * (after the collision and lists too) and overlaps only the bounce-back,
     * so it never shares HBM with the streaming sweep */
    const bool after = c->mem_after != 0;
    hipEvent_t m0 = nullptr, s0 = nullptr;
    const int nm = membrane_launches(c);
    int km = 0;
    if (!after) {
        if (lock) {
            HIPCHK(hipEventRecord(c->ev_pre, c->st));
            HIPCHK(hipStreamWaitEvent(c->st2, c->ev_pre, 0));
        } else {
            HIPCHK(hipStreamWaitEvent(c->st2, c->ev_str[p ^ 1], 0));
        }
        /* the two chains' launches alternate on the host, membrane first:
         * the host's launch latency then delays neither chain's start */
        if (rec) { m0 = c->event(); hipEventRecord(m0, c->st2); }
        rc = membrane_launch(c, mem_phases, c->st2, km++);
        if (rc) return rc;
    }
    /* the triangle kinematics of a step without collision or lists come
     * with the streaming launch (its first blocks) */
    const bool tk = !col && !ll;
    if (rec) {
        /* events bound to the dispatch itself: the kernel's own duration */
        e0 = c->event();
        e1 = c->event();
        rc = launch_bulk(c, e0, e1, tk);
        if (rc) return rc;
        c->pend[0].push_back({e0, e1});
        s0 = c->event();
        hipEventRecord(s0, c->st);
    } else {
        rc = launch_bulk(c, nullptr, nullptr, tk);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(c->ev_str[p], c->st));
    if (after) {
        HIPCHK(hipStreamWaitEvent(c->st2, c->ev_str[p], 0));
        if (rec) { m0 = c->event(); hipEventRecord(m0, c->st2); }
    }
    if (km < nm) {
        rc = membrane_launch(c, mem_phases, c->st2, km++);
        if (rc) return rc;
    }
    rc = launch_near(c);
    if (rc) return rc;
    if (rec) { e1 = c->event(); hipEventRecord(e1, c->st); c->pend[3].push_back({s0, e1}); }
    while (km < nm) {
        rc = membrane_launch(c, mem_phases, c->st2, km++);
        if (rc) return rc;
    }
    if (rec) { e1 = c->event(); hipEventRecord(e1, c->st2); c->pend[2].push_back({m0, e1}); }
    HIPCHK(hipEventRecord(c->ev_mem2[p], c->st2));
    LAUNCHCHK();
    return IRBC_OK;
}

/* Is the step run on one queue (k_step_stream + k_step_near)?  Needs the
 * one-launch membrane chain. */
static bool one_queue(const irbc_ctx *c) { return c->one_queue && c->mem_chain; }

/* The membrane chain (phases up to the gather) of this step, this step's
 * triangle kinematics (tri_kin) and its bulk streaming in one launch;
 * e0/e1 (both or neither) are events bound to the dispatch. */
#ifndef IRBC_CHAIN_RESIDENCY
#define IRBC_CHAIN_RESIDENCY 1.0 /* share of the launch a chain workgroup holds its wave slots */
#endif
#ifndef IRBC_CHAIN_RESIDENCY_SOLO
#define IRBC_CHAIN_RESIDENCY_SOLO 5.3 /* the same for a single replica: its XCD keeps ~9% of a share */
#endif
/* The bulk units' split over the XCDs for the fused streaming launch, a
 * contiguous range per XCD.  With fewer replicas than XCDs some XCDs host
 * no chain, and XCD x's weight is 1 - IRBC_CHAIN_RESIDENCY x (the share of
 * its wave slots its replicas' chain workgroups hold; replica r's chain runs
 * on XCD r % 8).  With more, every XCD hosts chains and the shares are
 * equal.  Measured (us per step, one box): the 64x64x48 replica 59.3 with
 * the tiles dealt round-robin, 58.8 with equal ranges, 56.8 weighted; a
 * mixed 16-request group 293 round-robin, 257 equal, 307 weighted; the
 * 64-replica batch unchanged.
 * A single replica's chain is a string of dependent round trips through its
 * XCD's L2, each slower while that XCD also streams: leaving it little of
 * the bulk (IRBC_CHAIN_RESIDENCY_SOLO) shortens the chain more than the
 * other seven XCDs' larger shares lengthen the sweep.  On the 64x64x48
 * replica (one box, us per step, one and four particles per lane): 54.0 and
 * 54.0 with the slot-share weight, 51.2 and 51.6 at 4.0, 50.5 and 51.0 at
 * 5.3, 51.0 and 50.3 with no bulk at all on that XCD. */
static XcdSplit xcd_split(const irbc_ctx *c, int units)
{
    const double res = c->nrep == 1 ? IRBC_CHAIN_RESIDENCY_SOLO
                                    : (c->nrep < IRBC_NXCD ? IRBC_CHAIN_RESIDENCY : 0.0);
    const double slots = 32.0 * 4.0 * 4.0 / 4.0; /* 32 CUs x 4 SIMDs x 4 waves, in 4-wave workgroups */
    double w[IRBC_NXCD], ws = 0.0;
    for (int x = 0; x < IRBC_NXCD; ++x) w[x] = 0.0;
    for (int r = 0; r < c->nrep; ++r) {
        const DevRep &R = c->rep[r];
        const int vb = (R.nv + IRBC_WG - 1) / IRBC_WG, eb = (R.ne + IRBC_WG - 1) / IRBC_WG,
                  tb = (R.nt + IRBC_WG - 1) / IRBC_WG;
        const int VPB = IRBC_WG / IRBC_GATHER_LANES;
        w[r % IRBC_NXCD] += vb + (eb + tb + vb) + (R.nv + VPB - 1) / VPB; /* the chain's workgroups */
    }
    for (int x = 0; x < IRBC_NXCD; ++x) {
        w[x] = std::max(0.05, 1.0 - res * std::min(0.9, w[x] / slots));
        ws += w[x];
    }
    XcdSplit xs;
    int given = 0;
    for (int x = 0; x < IRBC_NXCD; ++x) {
        xs.q[x] = (int)std::floor(units * w[x] / ws);
        given += xs.q[x];
    }
    /* the rounding remainder to the XCDs with the most weight */
    for (int k = 0; given < units; ++k, ++given) {
        int best = 0;
        for (int x = 1; x < IRBC_NXCD; ++x)
            if (w[x] - (double)xs.q[x] / units * ws > w[best] - (double)xs.q[best] / units * ws) best = x;
        xs.q[best] += 1;
    }
    int o = 0;
    for (int x = 0; x < IRBC_NXCD; ++x) {
        xs.o[x] = o;
        o += xs.q[x];
    }
    return xs;
}

static int launch_step_stream(irbc_ctx *c, hipEvent_t e0, hipEvent_t e1, bool tri_kin, int phases)
{
    const MemDev Mm = c->mem(c->par ^ 1); /* the membrane writes the next step's parity */
    const MemDev Ms = c->mem(c->par);
    const TriKin tk{Ms.t1, Ms.t2, Ms.t3, Ms.vkin, Ms.tri_kin, Ms.tri_sph};
    const int NB = (int)chain_blocks(c) - 1; /* no finish ticket: the finish runs in k_step_near */
    const int CB = IRBC_NXCD * NB * ((c->nrep + IRBC_NXCD - 1) / IRBC_NXCD);
    const int tkb = std::max(1, (c->z.maxnt + IRBC_WG - 1) / IRBC_WG);
    const int ntk = tri_kin ? tkb * c->nrep : 0;
    const int ntkp = IRBC_NXCD * ((ntk + IRBC_NXCD - 1) / IRBC_NXCD); /* the bulk starts on XCD 0 */
    const int T = (int)c->z.tiles;
    const int32_t *tr = c->d_tile_rep, *lo = c->lloff;
    const double *ct = c->ctab;
    const uint64_t *cm = c->cmask;
    const bool four = c->stream4 != 0;
    const int units = four ? (T + IRBC_WG / 64 - 1) / (IRBC_WG / 64) : T;
    const XcdSplit xs = xcd_split(c, units);
    int qmax = 0;
    for (int x = 0; x < IRBC_NXCD; ++x) qmax = std::max(qmax, xs.q[x]);
    const unsigned G = (unsigned)(CB + ntkp + IRBC_NXCD * qmax);
    auto *kern = four ? k_step_stream<true> : k_step_stream<false>;
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(G), dim3(IRBC_WG), 0, c->st, e0, e1, 0, c->d_rep, c->nrep, phases, Mm, NB,
                              CB, tr, T, ct, lo, cm, c->nearq, c->par, tk, tkb, ntk, ntkp, xs);
    else
        kern<<<G, IRBC_WG, 0, c->st>>>(c->d_rep, c->nrep, phases, Mm, NB, CB, tr, T, ct, lo, cm, c->nearq, c->par,
                                       tk, tkb, ntk, ntkp, xs);
    LAUNCHCHK();
    return IRBC_OK;
}

/* this step's bounce-back and the membrane finish (one block per replica) */
static int launch_step_near(irbc_ctx *c, int phases)
{
    const double *tk = c->tri_kin + (c->par & 1) * c->tri_stride();
    const float4 *ts = c->tri_sph + (c->par & 1) * c->z.nt;
    const MemDev Mm = c->mem(c->par ^ 1);
    const unsigned nr = (unsigned)c->nrep;
    if (c->near_merged) {
        k_step_near<true><<<dim3(near_blocks(c) + 1, nr), IRBC_WG, 0, c->st>>>(c->d_rep, c->lloff, c->llent,
                                                                           c->nearq, tk, ts, c->par, phases, Mm);
    } else {
        launch_sphere(c, tk, ts);
        k_step_near<false><<<dim3(std::max(16u, near_blocks(c) / 8u) + 1, nr), IRBC_WG, 0, c->st>>>(
            c->d_rep, c->lloff, c->llent, c->deferq, tk, ts, c->par, phases, Mm);
    }
    LAUNCHCHK();
    return IRBC_OK;
}

/* One MD step on one queue: [collision] [lists] [triangle kinematics and
 * near masks after either], the fused streaming launch, the fused
 * bounce-back launch.  rec brackets phases with timing events (the
 * membrane runs inside the streaming launch and has no figure of its own). */
static int launch_step_one(irbc_ctx *c, int cf, bool ll, bool rec)
{
    int rc;
    const bool col = cf != 0;
    const int mem_phases = PH_ALL & ~PH_RESET_NEAR; /* k_stream clears the other parity's queue */
    hipEvent_t e0 = nullptr, e1 = nullptr, s0 = nullptr;
    if (col) {
        if (rec) { e0 = c->event(); hipEventRecord(e0, c->st); }
        rc = launch_collision(c, 0, c->diag_on ? 1 : 0, cf);
        if (rc) return rc;
        if (rec) { e1 = c->event(); hipEventRecord(e1, c->st); c->pend[1].push_back({e0, e1}); }
    }
    if (ll) {
        rc = launch_llist(c, 0);
        if (rc) return rc;
    }
    if (col || ll) {
        rc = launch_tri_kin(c, 0);
        if (rc) return rc;
        rc = launch_cell_mask(c, 0);
        if (rc) return rc;
    }
    const bool tk = !col && !ll;
    if (rec) {
        e0 = c->event();
        e1 = c->event();
        rc = launch_step_stream(c, e0, e1, tk, mem_phases);
        if (rc) return rc;
        c->pend[0].push_back({e0, e1});
        s0 = c->event();
        hipEventRecord(s0, c->st);
    } else {
        rc = launch_step_stream(c, nullptr, nullptr, tk, mem_phases);
        if (rc) return rc;
    }
    rc = launch_step_near(c, mem_phases);
    if (rc) return rc;
    if (rec) { e1 = c->event(); hipEventRecord(e1, c->st); c->pend[3].push_back({s0, e1}); }
    return IRBC_OK;
}

/* the solvent stream waits for the membrane of the last launched step
 * (parity par ^ 1): after it, work on c->st sees the whole step */
static int join_chains(irbc_ctx *c)
{
    if (one_queue(c)) return IRBC_OK;
    HIPCHK(hipStreamWaitEvent(c->st, c->ev_mem2[(c->par ^ 1) & 1], 0));
    return IRBC_OK;
}

/* collide / list-update flags of the next step of every replica, then
 * the host mirror advanced by one step */
static void next_flags(irbc_ctx *c, std::vector<HostRep> &hr, int &cf, bool &ll)
{
    cf = 0;
    ll = false;
    for (int r = 0; r < c->nrep; ++r) {
        const irbc_params &p = hr[r].spec.params;
        if (host_collides(p, hr[r])) cf |= sorts_at(hr[r].step, p.csteps, c->sort_every) ? COL_SORT : COL_INPLACE;
        ll |= host_llist(hr[r].spec.params, hr[r], c->align_lists);
        advance_host(hr[r].spec.params, hr[r], c->align_lists);
    }
}

/* A chunk of steps captured once as a HIP graph and replayed whenever the
 * same chunk shape comes again: key = the collide and list-update flags of
 * each step, the starting parity and the launch options. */
static int run_graph_chunk(irbc_ctx *c, int n, uint64_t flags)
{
    /* first: 4 flag bits per step; second: parity, chunk length and every
     * option that changes the launches, each in a field of its own */
    const uint64_t shape = (uint64_t)(c->par & 1) | ((uint64_t)n << 1) | ((uint64_t)(c->mem_fused & 1) << 8) |
                           ((uint64_t)(c->diag_on & 1) << 9) | ((uint64_t)(c->near_merged & 1) << 10) |
                           ((uint64_t)(c->xcd_chunk & 1) << 11) | ((uint64_t)(c->stream4 & 1) << 12) |
                           ((uint64_t)(c->stream_grid & 1) << 13) | ((uint64_t)(c->mem_after & 1) << 14) |
                           ((uint64_t)(c->mem_chain & 1) << 15) | ((uint64_t)(c->chain_dev & 1) << 16) |
                           ((uint64_t)(c->near_masks & 1) << 17) | ((uint64_t)(c->one_queue & 1) << 18) |
                           ((uint64_t)(c->bin_span & 7) << 19);
    const std::pair<uint64_t, uint64_t> key(flags, shape);
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        const int par0 = c->par;
        if (g_api_depth > 0) return fail(IRBC_ERR_STATE, "graph capture inside another library call");
        std::unique_lock<std::shared_mutex> api_lock(g_api_mu);
        std::lock_guard<std::mutex> lock(g_capture_mu);
        HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
        int rc = IRBC_OK;
        for (int k = 0; k < n && rc == IRBC_OK; ++k) {
            const int cf = (int)((flags >> (4 * k)) & 3);
            const bool ll = (flags >> (4 * k + 2)) & 1;
            rc = one_queue(c) ? launch_step_one(c, cf, ll, false) : launch_step(c, cf, ll, false, k == 0);
            c->par ^= 1;
        }
        if (rc == IRBC_OK) rc = join_chains(c); /* the capture ends with both streams joined */
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->st, &g);
        c->par = par0;
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(IRBC_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(IRBC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        if (c->graphs.size() >= 64) {
            /* bounded cache: drop everything, once nothing launched from it
             * is still queued or running (destroying an executable frees
             * its kernel-argument memory, which a launch in flight still
             * reads) */
            HIPCHK(c->sync_both());
            for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
            c->graphs.clear();
        }
        it = c->graphs.emplace(key, ex).first;
    }
    HIPCHK(hipGraphLaunch(it->second, c->st));
    if (n & 1) c->par ^= 1;
    return IRBC_OK;
}

int irbc_md_steps(irbc_ctx *c, int64_t nsteps)
{
    if (!c || nsteps < 0) return fail(IRBC_ERR_ARG, "irbc_md_steps: bad arguments");
    int rc = ready(c, true);
    if (rc) return rc;
    if (nsteps == 0) return IRBC_OK;
    if (!c->primed) {
        /* the state came from elsewhere (set calls, stage calls, a
         * checkpoint): this step's shift, vertex and triangle kinematics and
         * fresh near masks with zero drift */
        k_draw_shift<<<(unsigned)((c->nrep + 63) / 64), 64, 0, c->st>>>(c->d_rep, c->nrep, 1);
        rc = launch_tri_kin(c, 1);
        if (rc) return rc;
        rc = launch_cell_mask(c, 1);
        if (rc) return rc;
    }
    c->primed = false; /* until this call's launches are all issued */
    if (c->use_graphs && !c->prof) {
        const int CH = IRBC_GRAPH_STEPS;
        for (int64_t s = 0; s < nsteps;) {
            const int n = (int)std::min<int64_t>(CH, nsteps - s);
            uint64_t flags = 0;
            for (int k = 0; k < n; ++k) {
                int cf;
                bool ll;
                next_flags(c, c->hr, cf, ll);
                flags |= (uint64_t)cf << (4 * k) | (uint64_t)ll << (4 * k + 2);
            }
            rc = run_graph_chunk(c, n, flags);
            if (rc) return rc;
            s += n;
        }
        c->primed = true;
        return IRBC_OK;
    }
    for (int64_t s = 0; s < nsteps; ++s) {
        int cf;
        bool ll;
        next_flags(c, c->hr, cf, ll);
        const bool rec = c->prof && (c->prof_step++ % c->prof_every) == 0;
        rc = one_queue(c) ? launch_step_one(c, cf, ll, rec) : launch_step(c, cf, ll, rec, s == 0);
        if (rc) return rc;
        c->par ^= 1;
        if (c->prof && c->pend[0].size() >= 2048) {
            rc = c->flush_prof();
            if (rc) return rc;
        }
    }
    rc = join_chains(c);
    if (rc == IRBC_OK) c->primed = true;
    return rc;
}

#ifdef IRBC_CHAIN_CLOCK
/* diagnostic builds only: the chain's clock records (3 per ticket, 64
 * tickets) of the last launch */
extern "C" int irbc_debug_chain_clock(irbc_ctx *c, long long *out)
{
    if (!c || !out) return fail(IRBC_ERR_ARG, "irbc_debug_chain_clock: bad arguments");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpyFromSymbol(out, HIP_SYMBOL(chain_clock_buf), sizeof(long long) * 6 * 256));
    return IRBC_OK;
}
#endif

int irbc_set_option(irbc_ctx *c, int option, int value)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->primed = false; /* the masks' and queues' use differs between shapes */
    switch (option) {
    case IRBC_OPT_MEMBRANE_FUSED:
        c->mem_fused = value != 0;
        return IRBC_OK;
    case IRBC_OPT_GRAPHS:
        c->use_graphs = value != 0;
        return IRBC_OK;
    case IRBC_OPT_NEAR_MERGED:
        c->near_merged = value != 0;
        return IRBC_OK;
    case IRBC_OPT_XCD_CHUNK:
        c->xcd_chunk = value != 0;
        return IRBC_OK;
    case IRBC_OPT_STREAM4:
        c->stream4 = value != 0;
        return IRBC_OK;
    case IRBC_OPT_STREAM_GRID:
        c->stream_grid = value != 0;
        return IRBC_OK;
    case IRBC_OPT_MEM_AFTER_STREAM:
        c->mem_after = value != 0;
        return IRBC_OK;
    case IRBC_OPT_MEMBRANE_CHAIN:
        c->mem_chain = value != 0;
        return IRBC_OK;
    case IRBC_OPT_CHAIN_DEVICE_SCOPE:
        c->chain_dev = value != 0;
        return IRBC_OK;
    case IRBC_OPT_NEAR_MASKS:
        c->near_masks = value != 0;
        return IRBC_OK;
    case IRBC_OPT_ONE_QUEUE:
        c->one_queue = value != 0;
        return IRBC_OK;
    case IRBC_OPT_CLASSIFY_ALL:
        c->classify_all = value != 0;
        return IRBC_OK;
    case IRBC_OPT_BIN_SPAN:
        if (value != 1 && value != 2 && value != 4)
            return fail(IRBC_ERR_ARG, "irbc_set_option: IRBC_OPT_BIN_SPAN is 1, 2 or 4");
        c->bin_span = value;
        return IRBC_OK;
    case IRBC_OPT_SORT_EVERY: {
        if (value < 1) return fail(IRBC_ERR_ARG, "irbc_set_option: IRBC_OPT_SORT_EVERY needs a value >= 1");
        HIPCHK(c->sync_both());
        HIPCHK(hipMemcpy(c->rep.data(), c->d_rep, c->nrep * sizeof(DevRep), hipMemcpyDeviceToHost));
        c->sort_every = value;
        for (int r = 0; r < c->nrep; ++r) {
            DevRep &R = c->rep[r];
            const HostRep &h = c->hr[r];
            R.sort_every = value;
            top_of_step(h.spec.params, h.step, h.colsteps, h.upd, R);
        }
        HIPCHK(hipMemcpy(c->d_rep, c->rep.data(), c->nrep * sizeof(DevRep), hipMemcpyHostToDevice));
        return IRBC_OK;
    }
    case IRBC_OPT_ALIGN_LISTS: {
        /* the device's copy of the rule for every replica, and the step
         * about to run re-decided under it, as the host mirror decides */
        HIPCHK(c->sync_both());
        HIPCHK(hipMemcpy(c->rep.data(), c->d_rep, c->nrep * sizeof(DevRep), hipMemcpyDeviceToHost));
        c->align_lists = value != 0;
        for (int r = 0; r < c->nrep; ++r) {
            DevRep &R = c->rep[r];
            const HostRep &h = c->hr[r];
            R.align_lists = c->align_lists;
            top_of_step(h.spec.params, h.step, h.colsteps, h.upd, R);
        }
        HIPCHK(hipMemcpy(c->d_rep, c->rep.data(), c->nrep * sizeof(DevRep), hipMemcpyHostToDevice));
        return IRBC_OK;
    }
    default:
        return fail(IRBC_ERR_ARG, "irbc_set_option: unknown option %d", option);
    }
}

int irbc_get_option(const irbc_ctx *c, int option, int *value)
{
    if (!c || !value) return fail(IRBC_ERR_ARG, "irbc_get_option: bad arguments");
    switch (option) {
    case IRBC_OPT_MEMBRANE_FUSED: *value = c->mem_fused; break;
    case IRBC_OPT_GRAPHS: *value = c->use_graphs; break;
    case IRBC_OPT_NEAR_MERGED: *value = c->near_merged; break;
    case IRBC_OPT_XCD_CHUNK: *value = c->xcd_chunk; break;
    case IRBC_OPT_STREAM4: *value = c->stream4; break;
    case IRBC_OPT_STREAM_GRID: *value = c->stream_grid; break;
    case IRBC_OPT_MEM_AFTER_STREAM: *value = c->mem_after; break;
    case IRBC_OPT_MEMBRANE_CHAIN: *value = c->mem_chain; break;
    case IRBC_OPT_CHAIN_DEVICE_SCOPE: *value = c->chain_dev; break;
    case IRBC_OPT_NEAR_MASKS: *value = c->near_masks; break;
    case IRBC_OPT_ONE_QUEUE: *value = c->one_queue; break;
    case IRBC_OPT_CLASSIFY_ALL: *value = c->classify_all; break;
    case IRBC_OPT_ALIGN_LISTS: *value = c->align_lists; break;
    case IRBC_OPT_SORT_EVERY: *value = c->sort_every; break;
    case IRBC_OPT_BIN_SPAN: *value = c->bin_span; break;
    default: return fail(IRBC_ERR_ARG, "irbc_get_option: no option %d", option);
    }
    return IRBC_OK;
}

int irbc_autotune_spans(const irbc_ctx *c, double *ms)
{
    if (!c || !ms) return fail(IRBC_ERR_ARG, "irbc_autotune_spans: bad arguments");
    for (int k = 0; k < 3; ++k) ms[k] = c->span_ms[k];
    return IRBC_OK;
}

int irbc_set_diagnostics(irbc_ctx *c, int on)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    c->diag_on = on != 0;
    return IRBC_OK;
}

int irbc_profile(irbc_ctx *c, int on)
{
    API_GUARD;
    if (!c) return fail(IRBC_ERR_ARG, "null context");
    int rc = c->flush_prof();
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
        c->acc_ms[k] = 0;
        c->acc_n[k] = 0;
    }
    c->prof = on > 0;
    c->prof_every = on > 0 ? on : 0;
    c->prof_step = 0;
    return IRBC_OK;
}

int irbc_profile_read(irbc_ctx *c, double *t)
{
    API_GUARD;
    if (!c || !t) return fail(IRBC_ERR_ARG, "irbc_profile_read: bad arguments");
    int rc = c->flush_prof();
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
        t[2 * k] = c->acc_ms[k];
        t[2 * k + 1] = (double)c->acc_n[k];
    }
    return IRBC_OK;
}

} /* extern "C" */

/* ---------------- TotalVA with no allocation ---------------- */

/* Per-block area/volume partials of irbc_membrane_va_static: one fixed
 * array per device, held until the process ends.  Host code never uses
 * this name as an address (it is a host-side stand-in there): the device
 * address comes from hipGetSymbolAddress, once per device. */
#define IRBC_VA_MAX_BLOCKS 1024 /* blocks of IRBC_WG triangles: up to 262144 triangles */
static __device__ __align__(256) double scratch[2 * IRBC_VA_MAX_BLOCKS];

#define IRBC_VA_MAX_DEVICES 64
static double *g_scratch_ptr[IRBC_VA_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* per triangle block: the unwrapped corners as the membrane step forms
 * them (x + ix Lx, y + iy Ly, z), Areatri/Voltri and the block's partials,
 * with the same arithmetic and the same blocks of IRBC_WG triangles as the
 * context's membrane kernels */
__global__ __launch_bounds__(IRBC_WG) void k_va_blocks(const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ z, const int32_t *__restrict__ ix,
                                                      const int32_t *__restrict__ iy,
                                                      const int32_t *__restrict__ tris, int nv, int nt, double Lx,
                                                      double Ly, double *__restrict__ part)
{
    __shared__ double red[2][IRBC_WG / 64];
    const int e = (int)blockIdx.x * IRBC_WG + (int)threadIdx.x;
    double sa = 0.0, sv = 0.0;
    if (e < nt) {
        double P[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int g = tris[3 * e + k];
            if (g < 0 || g >= nv) { /* a vertex index out of range: the result is NaN */
                P[k][0] = P[k][1] = P[k][2] = __builtin_nan("");
                continue;
            }
            P[k][0] = x[g] + ix[g] * Lx;
            P[k][1] = y[g] + iy[g] * Ly;
            P[k][2] = z[g];
        }
        tri_area_volume(P[0], P[1], P[2], sa, sv);
    }
    av_block_partials(sa, sv, red, part + 2 * (int64_t)blockIdx.x);
}

/* the blocks' partials summed in block order, as the membrane step sums them */
__global__ __launch_bounds__(64) void k_va_sum(const double *__restrict__ part, int nblk, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double a = 0.0, v = 0.0;
    for (int b = 0; b < nblk; ++b) {
        a += part[2 * b];
        v += part[2 * b + 1];
    }
    out[0] = a;
    out[1] = v;
}

extern "C" {

int irbc_membrane_va_static(const double *x, const double *y, const double *z, const int32_t *ix,
                            const int32_t *iy, int32_t nv, const int32_t *tris, int32_t nt, double Lx, double Ly,
                            double *out, void *stream)
{
    API_GUARD;
    if (!x || !y || !z || !ix || !iy || !tris || !out || nv < 3 || nt < 1)
        return fail(IRBC_ERR_ARG, "irbc_membrane_va_static: null array or empty mesh");
    const int nblk = (nt + IRBC_WG - 1) / IRBC_WG;
    if (nblk > IRBC_VA_MAX_BLOCKS)
        return fail(IRBC_ERR_ARG, "irbc_membrane_va_static: %d triangles, at most %d", nt,
                    IRBC_VA_MAX_BLOCKS * IRBC_WG);
    if (!(Lx > 0.0) || !(Ly > 0.0)) return fail(IRBC_ERR_ARG, "irbc_membrane_va_static: box lengths must be > 0");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= IRBC_VA_MAX_DEVICES)
        return fail(IRBC_ERR_ARG, "irbc_membrane_va_static: device %d beyond %d", dev, IRBC_VA_MAX_DEVICES);
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            /* the device's own copy of the array, looked up with that
             * device current (the calling thread's) */
            double *p = nullptr;
            const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail(IRBC_ERR_HIP, "irbc_membrane_va_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
            g_scratch_ptr[dev] = p;
        }
        part = g_scratch_ptr[dev];
    }
    /* every block writes its partials before the sum reads them: nothing
     * in the array needs clearing between calls */
    hipStream_t s = (hipStream_t)stream;
    k_va_blocks<<<(unsigned)nblk, IRBC_WG, 0, s>>>(x, y, z, ix, iy, tris, nv, nt, Lx, Ly, part);
    k_va_sum<<<1, 64, 0, s>>>(part, nblk, out);
    LAUNCHCHK();
    return IRBC_OK;
}

} /* extern "C" */
