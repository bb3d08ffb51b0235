/*
 * scenario.h — host-side scenario setup for the MI355X iRBC library.
 *
 * Reads the reference's mesh files, draws the initial state of the README
 * run in the reference's draw order, and loads it into a replica of a
 * context through the C ABI.  Plain C++17; no device code.
 */
#ifndef IRBC_HOST_SCENARIO_H
#define IRBC_HOST_SCENARIO_H

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "irbc.h"

namespace irbc_host {

/* The reference's CRandomMersenne stream (Src/mersenne.cpp:27-125): the
 * standard MT19937 initialised with the seed, with 37 outputs discarded;
 * Random() = bits * 2^-32; IRandom(lo, hi) multiplies and truncates. */
class RefStream {
  public:
    explicit RefStream(uint32_t seed) : g_(seed) { g_.discard(37); }
    double random() { return (double)g_() * (1.0 / 4294967296.0); }
    int irandom(int lo, int hi)
    {
        if (hi <= lo) return hi == lo ? lo : (int)0x80000000;
        int r = (int)((double)(uint32_t)(hi - lo + 1) * random() + lo);
        return r > hi ? hi : r;
    }

  private:
    std::mt19937 g_;
};

struct Mesh {
    std::vector<double> xyz;     /* nv x 3 */
    std::vector<int32_t> edges;  /* ne x 2, 0-based */
    std::vector<int32_t> tris;   /* nt x 3, 0-based */
    int nv() const { return (int)(xyz.size() / 3); }
    int ne() const { return (int)(edges.size() / 2); }
    int nt() const { return (int)(tris.size() / 3); }
};

/* The three text files of the reference (maincode.cpp:175-272): "x y z"
 * per vertex, "l1 l2" per edge and "l1 l2 l3" per triangle, 1-based. */
bool read_mesh(const char *vfile, const char *efile, const char *tfile, Mesh &m, std::string &err);

/* Initial state of one replica, in the reference's units. */
struct Scenario {
    irbc_params p;
    Mesh mesh; /* placed in the box */
    std::vector<double> sx, sy, sz, svx, svy, svz;
    std::vector<double> vvx, vvy, vvz;
    std::vector<int32_t> knob;
    double para[12];
    int32_t parai[3];
    double sarea0 = 0.0, volume0 = 0.0;
    double mean_edge = 0.0, sd_edge = 0.0;
};

/* The README run's set-up (maincode.cpp:76-446), draw for draw from one
 * RefStream: solvent positions and velocities (uniform, +-sqrt(3 Theta/m)),
 * vertex velocities, placement (centre at Lx/2+10, Ly/2, lowest vertex at
 * z = 0.3), parasite bead, and knobs placed on triangles chosen by area. */
void readme_scenario(const irbc_params &p, const Mesh &raw, int64_t n_solvent, int nknobs, uint32_t seed,
                     Scenario &s);

/* Loads a scenario into replica r and runs the start of the time loop
 * (maincode.cpp:383-391): forces at rest, inside flags, cell lists. */
int load_scenario(irbc_ctx *c, int r, const Scenario &s);

/* Frame of the vertex data file: "#k" then "x y z ix iy bondi" per vertex
 * (maincode.cpp:746-754). */
int write_frame(irbc_ctx *c, int r, int nv, int k, FILE *f);
/* ASCII STL of the membrane (Src/memfuncs.cpp:590-626). */
int write_stl(irbc_ctx *c, int r, const Mesh &m, const irbc_params &p, const char *path);
/* knobs.dat: "x y z" of every knob (maincode.cpp:435-446); returns the count */
int write_knobs(const Scenario &s, const char *path);

} /* namespace irbc_host */

#endif
