/*
 * host_capi.cpp — C entry points of the host library for callers that are
 * not C++ (ctypes, cgo): the README run's initial state as plain arrays.
 */
#include <cstring>
#include <string>

#include "scenario.h"

extern "C" {

/* Draws the README run's set-up (maincode.cpp:76-446) for the mesh files
 * given.  Outputs (caller-allocated): solvent[6*ns] as x,y,z,vx,vy,vz
 * blocks; placed vertices xyz[3*nv]; vertex velocities vvel[3*nv] as
 * vx,vy,vz blocks; knob[nv]; para[12]; stats[4] = mean edge, edge SD,
 * Sarea0, Volume0.  Returns nv, or -1 on a file or size error. */
int irbc_host_readme_scenario(const irbc_params *p, const char *vfile, const char *efile, const char *tfile,
                              int64_t ns, int nknobs, uint32_t seed, int nv_cap, double *solvent, double *xyz,
                              double *vvel, int32_t *knob, double *para, double *stats)
{
    irbc_host::Mesh m;
    std::string err;
    if (!p || !irbc_host::read_mesh(vfile, efile, tfile, m, err) || m.nv() > nv_cap || ns < 0) return -1;
    irbc_host::Scenario s;
    irbc_host::readme_scenario(*p, m, ns, nknobs, seed, s);
    const double *blk[6] = {s.sx.data(), s.sy.data(), s.sz.data(), s.svx.data(), s.svy.data(), s.svz.data()};
    for (int k = 0; k < 6; ++k)
        if (solvent) std::memcpy(solvent + k * ns, blk[k], ns * sizeof(double));
    const int nv = m.nv();
    if (xyz) std::memcpy(xyz, s.mesh.xyz.data(), 3 * nv * sizeof(double));
    if (vvel) {
        std::memcpy(vvel, s.vvx.data(), nv * sizeof(double));
        std::memcpy(vvel + nv, s.vvy.data(), nv * sizeof(double));
        std::memcpy(vvel + 2 * nv, s.vvz.data(), nv * sizeof(double));
    }
    if (knob) std::memcpy(knob, s.knob.data(), nv * sizeof(int32_t));
    if (para) std::memcpy(para, s.para, 12 * sizeof(double));
    if (stats) {
        stats[0] = s.mean_edge;
        stats[1] = s.sd_edge;
        stats[2] = s.sarea0;
        stats[3] = s.volume0;
    }
    return nv;
}

} /* extern "C" */
