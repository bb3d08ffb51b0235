/*
 * scenario.cpp — mesh files, the README run's initial state, and the
 * reference's output formats, over the C ABI.
 */
#include "scenario.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace irbc_host {

static bool read_table(const char *path, int cols, std::vector<double> &out, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = std::string("cannot open ") + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        double v[3];
        int k = 0;
        while (k < cols && (ss >> v[k])) ++k;
        if (k == 0) continue; /* blank line */
        if (k != cols) {
            err = std::string("malformed line in ") + path + ": " + line;
            return false;
        }
        for (int j = 0; j < cols; ++j) out.push_back(v[j]);
    }
    return true;
}

bool read_mesh(const char *vfile, const char *efile, const char *tfile, Mesh &m, std::string &err)
{
    std::vector<double> v, e, t;
    if (!read_table(vfile, 3, v, err) || !read_table(efile, 2, e, err) || !read_table(tfile, 3, t, err))
        return false;
    m.xyz = v;
    const int nv = (int)(v.size() / 3);
    m.edges.resize(e.size());
    m.tris.resize(t.size());
    for (size_t i = 0; i < e.size(); ++i) m.edges[i] = (int32_t)e[i] - 1;
    for (size_t i = 0; i < t.size(); ++i) m.tris[i] = (int32_t)t[i] - 1;
    for (int32_t a : m.edges)
        if (a < 0 || a >= nv) return err = "edge index out of range", false;
    for (int32_t a : m.tris)
        if (a < 0 || a >= nv) return err = "triangle index out of range", false;
    return true;
}

/* Src/memfuncs.cpp:144-157 with all images zero */
static double tri_area(const std::vector<double> &x, int p1, int p2, int p3)
{
    double A[3], B[3], e[3];
    for (int k = 0; k < 3; ++k) {
        A[k] = x[3 * p2 + k] - x[3 * p1 + k];
        B[k] = x[3 * p3 + k] - x[3 * p1 + k];
    }
    e[0] = (A[1] * B[2]) - (B[1] * A[2]);
    e[1] = (A[2] * B[0]) - (B[2] * A[0]);
    e[2] = (A[0] * B[1]) - (B[0] * A[1]);
    return 0.5 * std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
}

/* Src/memfuncs.cpp:124-142 with all images zero */
static double tri_volume(const std::vector<double> &x, int p1, int p2, int p3)
{
    double A[3], B[3], D[3], Cn[3];
    for (int k = 0; k < 3; ++k) {
        Cn[k] = (x[3 * p1 + k] + x[3 * p2 + k] + x[3 * p3 + k]) / 3.0;
        A[k] = x[3 * p2 + k] - x[3 * p1 + k];
        B[k] = x[3 * p3 + k] - x[3 * p1 + k];
    }
    D[0] = (A[1] * B[2]) - (B[1] * A[2]);
    D[1] = (A[2] * B[0]) - (B[2] * A[0]);
    D[2] = (A[0] * B[1]) - (B[0] * A[1]);
    return (1.0 / 6.0) * ((D[0] * Cn[0]) + (D[1] * Cn[1]) + (D[2] * Cn[2]));
}

void readme_scenario(const irbc_params &p, const Mesh &raw, int64_t ns, int nknobs, uint32_t seed, Scenario &s)
{
    RefStream rg(seed);
    s.p = p;
    s.mesh = raw;
    /* solvent: maincode.cpp:96-109 */
    const double vmax = std::sqrt(3.0 * p.Theta / p.mass);
    s.sx.resize(ns); s.sy.resize(ns); s.sz.resize(ns);
    s.svx.resize(ns); s.svy.resize(ns); s.svz.resize(ns);
    for (int64_t j = 0; j < ns; ++j) {
        s.sx[j] = p.Lx * rg.random();
        s.sy[j] = p.Ly * rg.random();
        s.sz[j] = p.Lz * rg.random();
        s.svx[j] = -vmax + 2.0 * vmax * rg.random();
        s.svy[j] = -vmax + 2.0 * vmax * rg.random();
        s.svz[j] = -vmax + 2.0 * vmax * rg.random();
    }
    /* vertices: velocities drawn as the file is read (maincode.cpp:175-195) */
    const int nv = raw.nv();
    const double vm = std::sqrt(3.0 * p.Theta / p.pmass);
    s.vvx.resize(nv); s.vvy.resize(nv); s.vvz.resize(nv);
    double zmin = 0.0;
    for (int i = 0; i < nv; ++i) {
        s.vvx[i] = -vm + 2.0 * vm * rg.random();
        s.vvy[i] = -vm + 2.0 * vm * rg.random();
        s.vvz[i] = -vm + 2.0 * vm * rg.random();
        if (i == 0 || raw.xyz[3 * i + 2] < zmin) zmin = raw.xyz[3 * i + 2];
    }
    /* placement (maincode.cpp:196-206) */
    const double C[3] = {p.Lx / 2.0 + 10.0, p.Ly / 2.0, -zmin + 0.3};
    for (int i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) s.mesh.xyz[3 * i + k] += C[k];
    /* parasite (maincode.cpp:210-225) */
    for (double &d : s.para) d = 0.0;
    s.para[0] = C[0] - 2.0;
    s.para[1] = C[1];
    s.para[2] = C[2];
    const double vp = std::sqrt(3.0 * p.Theta / p.PRmass);
    for (int k = 0; k < 3; ++k) s.para[3 + k] = -vp + 2.0 * vp * rg.random();
    s.parai[0] = s.parai[1] = s.parai[2] = 0;
    /* edge statistics (maincode.cpp:236-257) */
    const int ne = raw.ne();
    double me = 0.0, me2 = 0.0;
    for (int i = 0; i < ne; ++i) {
        const int a = raw.edges[2 * i], b = raw.edges[2 * i + 1];
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = s.mesh.xyz[3 * a + k] - s.mesh.xyz[3 * b + k];
            d2 += d * d;
        }
        const double l0 = std::sqrt(d2);
        me += l0;
        me2 += l0 * l0;
    }
    s.mean_edge = ne ? me / ne : 0.0;
    s.sd_edge = ne ? std::sqrt(me2 / ne - s.mean_edge * s.mean_edge) : 0.0;
    /* rest area and volume, per-triangle areas (maincode.cpp:343,370-374) */
    const int nt = raw.nt();
    std::vector<double> tarea(nt);
    s.sarea0 = s.volume0 = 0.0;
    for (int t = 0; t < nt; ++t) {
        const int *l = &raw.tris[3 * t];
        tarea[t] = tri_area(s.mesh.xyz, l[0], l[1], l[2]);
        s.sarea0 += tarea[t];
        s.volume0 += tri_volume(s.mesh.xyz, l[0], l[1], l[2]);
    }
    /* knobs: a triangle by area, then one of its corners; when that corner
     * already holds a knob the scan goes on to the next triangle with a new
     * corner draw (maincode.cpp:393-433) */
    s.knob.assign(nv, 0);
    std::vector<double> parea(nt);
    for (int j = 0; j < nt; ++j) parea[j] = tarea[j] / s.sarea0;
    int placed = 0;
    const int want = nknobs < nv ? nknobs : nv;
    while (placed < want) {
        const double rad = rg.random();
        double acc = 0.0;
        for (int j = 0; j < nt; ++j) {
            acc += parea[j];
            if (acc > rad) {
                const int pick = rg.irandom(1, 3);
                const int v = raw.tris[3 * j + pick - 1];
                if (s.knob[v] == 0) {
                    s.knob[v] = 1;
                    ++placed;
                    break;
                }
            }
        }
    }
}

int load_scenario(irbc_ctx *c, int r, const Scenario &s)
{
    int rc = irbc_set_mesh(c, r, s.mesh.xyz.data(), s.mesh.edges.data(), s.mesh.tris.data());
    if (rc) return rc;
    rc = irbc_set_solvent(c, r, (int64_t)s.sx.size(), s.sx.data(), s.sy.data(), s.sz.data(), s.svx.data(),
                          s.svy.data(), s.svz.data(), nullptr);
    if (rc) return rc;
    rc = irbc_set_vertices(c, r, nullptr, nullptr, nullptr, s.vvx.data(), s.vvy.data(), s.vvz.data(), nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, s.knob.data());
    if (rc) return rc;
    rc = irbc_set_parasite(c, r, s.para, s.parai);
    if (rc) return rc;
    rc = irbc_set_counters(c, r, 0, -1, 0);
    if (rc) return rc;
    rc = irbc_membrane_phases(c, IRBC_PH_VA | IRBC_PH_FORCES);
    if (rc) return rc;
    rc = irbc_classify_solvent(c);
    if (rc) return rc;
    return irbc_llist_update(c);
}

int write_frame(irbc_ctx *c, int r, int nv, int k, FILE *f)
{
    std::vector<double> x(nv), y(nv), z(nv);
    std::vector<int32_t> ix(nv), iy(nv), bondi(nv);
    int rc = irbc_get_vertices(c, r, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, ix.data(), iy.data(), nullptr, bondi.data(), nullptr,
                               nullptr);
    if (rc) return rc;
    std::fprintf(f, "#%d\n", k);
    for (int i = 0; i < nv; ++i)
        std::fprintf(f, "%lf\t%lf\t%lf\t%d\t%d\t%d\n", x[i], y[i], z[i], ix[i], iy[i], bondi[i]);
    std::fflush(f);
    return IRBC_OK;
}

int write_stl(irbc_ctx *c, int r, const Mesh &m, const irbc_params &p, const char *path)
{
    const int nv = m.nv();
    std::vector<double> x(nv), y(nv), z(nv);
    std::vector<int32_t> ix(nv), iy(nv);
    int rc = irbc_get_vertices(c, r, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, ix.data(), iy.data(), nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::ofstream f(path);
    if (!f) return IRBC_ERR_ARG;
    f << "solid ascii\n";
    for (int t = 0; t < m.nt(); ++t) {
        double P[3][3];
        for (int k = 0; k < 3; ++k) {
            const int v = m.tris[3 * t + k];
            P[k][0] = x[v] + ix[v] * p.Lx;
            P[k][1] = y[v] + iy[v] * p.Ly;
            P[k][2] = z[v];
        }
        double A[3], B[3];
        for (int k = 0; k < 3; ++k) {
            A[k] = P[1][k] - P[0][k];
            B[k] = P[2][k] - P[0][k];
        }
        const double N[3] = {(A[1] * B[2]) - (B[1] * A[2]), (A[2] * B[0]) - (B[2] * A[0]),
                             (A[0] * B[1]) - (B[0] * A[1])};
        f << "facet normal\t" << N[0] << "\t" << N[1] << "\t" << N[2] << "\n" << "outer loop\n";
        for (int k = 0; k < 3; ++k) f << "vertex\t" << P[k][0] << "\t" << P[k][1] << "\t" << P[k][2] << "\n";
        f << "endloop\nendfacet\n";
    }
    f << "endsolid\n";
    return f.good() ? IRBC_OK : IRBC_ERR_ARG;
}

int write_knobs(const Scenario &s, const char *path)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return -1;
    int n = 0;
    for (int i = 0; i < s.mesh.nv(); ++i)
        if (s.knob[i] == 1) {
            std::fprintf(f, "%lf\t%lf\t%lf\n", s.mesh.xyz[3 * i], s.mesh.xyz[3 * i + 1], s.mesh.xyz[3 * i + 2]);
            ++n;
        }
    std::fclose(f);
    return n;
}

} /* namespace irbc_host */
