/*
 * ref_hooks.cpp — harness around a build of the reference's own sources.
 *
 * TEST INFRASTRUCTURE ONLY (builds oracle/_ref/ref_run; used to write the
 * golden fixtures in this container, never by the product).
 *
 * The reference's main() (maincode.cpp, compiled with -Dmain=irbc_ref_main)
 * runs unmodified.  Link-time wrapping gives this harness three hooks:
 *   time()               returns IRBC_REF_SEED, so the run is reproducible
 *                        (maincode.cpp:58 seeds with time(0));
 *   TotalVA()            call k (k=1 is init, maincode.cpp:385; k=n+2 is step
 *                        n, maincode.cpp:602) — dump the state before it;
 *   force_calculate_dt() same counting — dump the forces after it, then stop
 *                        once the last requested step is written.
 * With IRBC_REF_TIMING=<seconds> set nothing is dumped: the harness takes the
 * time at TotalVA() of step 0 and again at each later step, and at step N or
 * at the first step past that many seconds it prints "ref_timing n seconds"
 * (n whole steps of the time loop, the set-up excluded) and stops.
 * Dumps are raw little-endian arrays in a simple tagged format read by
 * tests/golden/make_golden.py.
 */
#include "define.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

/* the wrapped C++ symbols are _Z7TotalVAv and _Z18force_calculate_dtv */
extern "C" time_t __real_time(time_t *);
extern "C" void __real__Z7TotalVAv();
extern "C" void __real__Z18force_calculate_dtv();
int irbc_ref_main(int argc, char *argv[]);

static int g_totalva_calls = 0;
static int g_force_calls = 0;
static long g_dump_step = -1;
static std::string g_dump_dir;
static bool g_timing = false;
static double g_timing_budget = 0.0;
static std::chrono::steady_clock::time_point g_t0;

extern "C" time_t __wrap_time(time_t *t)
{
    const char *s = std::getenv("IRBC_REF_SEED");
    time_t v = s ? (time_t)std::atol(s) : __real_time(nullptr);
    if (t) *t = v;
    return v;
}

static FILE *g_out = nullptr;

static void put(const char *tag, const void *data, size_t elem, size_t n, char kind)
{
    /* record: 16-byte tag, kind ('d' f64, 'i' i32), elem size, count, bytes */
    char t[16] = {0};
    std::strncpy(t, tag, 15);
    uint64_t hdr[2] = {(uint64_t)elem, (uint64_t)n};
    std::fwrite(t, 1, 16, g_out);
    std::fwrite(&kind, 1, 1, g_out);
    std::fwrite(hdr, 8, 2, g_out);
    std::fwrite(data, elem, n, g_out);
}

static void dump_state(const char *phase, long step)
{
    char path[512];
    std::snprintf(path, sizeof(path), "%s/%s_%ld.bin", g_dump_dir.c_str(), phase, step);
    g_out = std::fopen(path, "wb");
    if (!g_out) { std::perror(path); std::exit(2); }
    size_t ns = solvent.size();
    std::vector<double> d(ns);
    std::vector<int> iv(ns);
#define SOLD(field, tag) for (size_t i = 0; i < ns; ++i) d[i] = solvent[i].field; put(tag, d.data(), 8, ns, 'd');
    SOLD(xpos, "sx") SOLD(ypos, "sy") SOLD(zpos, "sz") SOLD(vx, "svx") SOLD(vy, "svy") SOLD(vz, "svz")
#undef SOLD
    for (size_t i = 0; i < ns; ++i) iv[i] = solvent[i].pid;
    put("spid", iv.data(), 4, ns, 'i');
    for (size_t i = 0; i < ns; ++i) iv[i] = solvent[i].indio;
    put("sindio", iv.data(), 4, ns, 'i');
    size_t nv = vts.size();
    std::vector<double> m(nv);
    std::vector<int> mi(nv);
#define VD(expr, tag) for (size_t i = 0; i < nv; ++i) m[i] = (expr); put(tag, m.data(), 8, nv, 'd');
#define VI(expr, tag) for (size_t i = 0; i < nv; ++i) mi[i] = (expr); put(tag, mi.data(), 4, nv, 'i');
    VD(vts[i].x, "mx") VD(vts[i].y, "my") VD(vts[i].z, "mz")
    VD(vts[i].vx, "mvx") VD(vts[i].vy, "mvy") VD(vts[i].vz, "mvz")
    VD(fmapdt[i].fx, "fx") VD(fmapdt[i].fy, "fy") VD(fmapdt[i].fz, "fz")
    VD(vts[i].poff, "poff") VD(vts[i].pon, "pon")
    VI(vts[i].ix, "ix") VI(vts[i].iy, "iy") VI(vts[i].pid, "mpid")
    VI(vts[i].bondi, "bondi") VI(vts[i].bondl, "bondl") VI(vts[i].knob, "knob")
#undef VD
#undef VI
    double par[12];
    for (int k = 0; k < 3; ++k) {
        par[k] = para[k];
        par[3 + k] = paraV[k];
        par[6 + k] = paraft[k];
        par[9 + k] = parafdt[k];
    }
    put("para", par, 8, 12, 'd');
    int pari[3] = {paradx, parady, paraid};
    put("parai", pari, 4, 3, 'i');
    double g[7] = {Sarea, Volume, Sarea0, Volume0, theta0, wallspeed, (double)nsolb};
    put("glob", g, 8, 7, 'd');
    put("ligi", ligi.data(), 4, ligi.size(), 'i');
    size_t nc = Cells.size();
    std::vector<double> cd(3 * nc);
    for (size_t c = 0; c < nc; ++c)
        for (int k = 0; k < 3; ++k) cd[3 * c + k] = Cells[c].vcm[k];
    put("vcm", cd.data(), 8, 3 * nc, 'd');
    std::vector<int> ci(nc);
    for (size_t c = 0; c < nc; ++c) ci[c] = Cells[c].nc;
    put("nc", ci.data(), 4, nc, 'i');
#define CD(field, tag) for (size_t c = 0; c < nc; ++c) cd[c] = Cells[c].field; put(tag, cd.data(), 8, nc, 'd');
    CD(ncc, "ncc") CD(dvv, "dvv") CD(ThE, "the") CD(rand1, "rand1") CD(rand2, "rand2")
#undef CD
    /* cell -> triangle lists as stored, CSR */
    std::vector<int> off(nc + 1, 0), ent;
    for (size_t c = 0; c < nc; ++c) {
        off[c + 1] = off[c] + (int)Cells[c].llist.size();
        for (int t : Cells[c].llist) ent.push_back(t);
    }
    put("lloff", off.data(), 4, nc + 1, 'i');
    put("llent", ent.data(), 4, ent.size(), 'i');
    std::fclose(g_out);
    g_out = nullptr;
}

extern "C" void __wrap__Z7TotalVAv()
{
    ++g_totalva_calls;
    long step = (long)g_totalva_calls - 2; /* call 1 is the init call */
    if (g_timing) {
        const auto now = std::chrono::steady_clock::now();
        if (step == 0) g_t0 = now;
        const double el = std::chrono::duration<double>(now - g_t0).count();
        if (step > 0 && (step == g_dump_step || el >= g_timing_budget)) {
            std::printf("ref_timing %ld %.9f\n", step, el);
            std::fflush(stdout);
            std::_Exit(0);
        }
    } else if (step >= 0 && step == g_dump_step) {
        dump_state("pre_va", step);
    }
    __real__Z7TotalVAv();
}

extern "C" void __wrap__Z18force_calculate_dtv()
{
    __real__Z18force_calculate_dtv();
    ++g_force_calls;
    long step = (long)g_force_calls - 2;
    if (g_timing) return;
    if (step == -1) dump_state("init_forces", 0);
    if (step >= 0 && step == g_dump_step) {
        dump_state("post_forces", step);
        std::fflush(stdout);
        std::_Exit(0);
    }
}

int main(int argc, char **argv)
{
    /* usage: ref_run DUMPDIR STEP  <the reference's own 8 arguments> */
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s dumpdir step kon koff V E T nknobs out stl\n", argv[0]);
        return 2;
    }
    g_dump_dir = argv[1];
    g_dump_step = std::atol(argv[2]);
    if (const char *tb = std::getenv("IRBC_REF_TIMING")) {
        g_timing = true;
        g_timing_budget = std::atof(tb);
    }
    std::vector<char *> args;
    args.push_back(argv[0]);
    for (int i = 3; i < argc; ++i) args.push_back(argv[i]);
    return irbc_ref_main((int)args.size(), args.data());
}
