/*
 * irbc_main.cpp — command-line driver with the reference's interface
 * (README.md usage, maincode.cpp:56-770), running the time loop on one
 * MI355X through the C ABI.
 *
 *   irbc_main onrate offrate vertices edges triangles knobs rbc_data_file stl_flag
 *             [--steps N] [--seed S] [--box Lx Ly Lz] [--walls W] [--frame-every F]
 *
 * Prints the reference's set-up lines, writes knobs.dat, and every F steps
 * (5000 in the reference) appends a frame "x y z ix iy bondi" to
 * rbc_data_file and, with stl_flag 1, writes rbcK.stl.  Without --steps it
 * runs the reference's 4,000,000 steps.  Exits non-zero with a message when
 * a call of the library fails.
 */
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "irbc.h"
#include "scenario.h"

using irbc_host::Mesh;
using irbc_host::Scenario;

static int die(const char *what, int rc)
{
    std::fprintf(stderr, "irbc_main: %s failed (%d): %s\n", what, rc, irbc_last_error());
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 9) {
        std::fprintf(stderr,
                     "usage: %s onrate offrate vertices edges triangles knobs rbc_data_file stl_flag\n"
                     "          [--steps N] [--seed S] [--box Lx Ly Lz] [--walls W] [--frame-every F]\n",
                     argv[0]);
        return 2;
    }
    irbc_params p;
    irbc_params_default(&p);
    p.onrate = std::atof(argv[1]);
    p.offrate = std::atof(argv[2]);
    const int nknobs = std::atoi(argv[6]);
    const int stl_flag = std::atoi(argv[8]);
    int64_t steps = 4000000;
    int64_t frame_every = 5000;
    uint32_t seed = (uint32_t)std::time(nullptr);
    for (int i = 9; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atoll(argv[++i]);
        else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        else if (!std::strcmp(argv[i], "--walls") && i + 1 < argc) p.walls = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "--frame-every") && i + 1 < argc) frame_every = std::atoll(argv[++i]);
        else if (!std::strcmp(argv[i], "--box") && i + 3 < argc) {
            p.Lx = std::atoi(argv[++i]);
            p.Ly = std::atoi(argv[++i]);
            p.Lz = std::atoi(argv[++i]);
        } else {
            std::fprintf(stderr, "irbc_main: unknown option %s\n", argv[i]);
            return 2;
        }
    }
    if (frame_every < 1 || steps < 0) {
        std::fprintf(stderr, "irbc_main: --steps must be >= 0 and --frame-every >= 1\n");
        return 2;
    }
    irbc_params_derive(&p);

    Mesh raw;
    std::string err;
    if (!irbc_host::read_mesh(argv[3], argv[4], argv[5], raw, err)) {
        std::fprintf(stderr, "irbc_main: %s\n", err.c_str());
        return 1;
    }
    if (nknobs > raw.nv()) {
        std::fprintf(stderr, "irbc_main: %d knobs exceed the %d vertices\n", nknobs, raw.nv());
        return 2;
    }
    const int64_t ns = (int64_t)std::llround(p.navg * p.Lx * p.Ly * p.Lz);
    Scenario s;
    irbc_host::readme_scenario(p, raw, ns, nknobs, seed, s);
    std::printf("Mean edge is:\t%g\n", s.mean_edge);
    std::printf("Standard deviation is:\t%g\n", s.sd_edge);
    std::printf("Sarea:\t%g\n", s.sarea0);
    std::printf("Volume:\t%g\n", s.volume0);
    const double shear = p.shear;
    const double ac = 2.0 * shear + p.kappasa + p.kappad;
    std::printf("Shear modulus::\t%g\n", shear);
    std::printf("Young's modulus::\t%g\n", (4.0 * ac * shear) / (ac + shear));
    std::printf("Poisson's ratio::\t%g\n", (ac - shear) / (ac + shear));

    irbc_replica_spec spec;
    std::memset(&spec, 0, sizeof spec);
    spec.params = p;
    spec.n_solvent = ns;
    spec.n_vertices = raw.nv();
    spec.n_edges = raw.ne();
    spec.n_triangles = raw.nt();
    spec.seed = seed;
    irbc_ctx *c = nullptr;
    int rc = irbc_create(&spec, 1, 0, &c);
    if (rc) return die("irbc_create", rc);
    rc = irbc_host::load_scenario(c, 0, s);
    if (rc) return die("loading the scenario", rc);
    const int nk = irbc_host::write_knobs(s, "knobs.dat");
    std::printf("no of vertices selected:\t%d\n", nk);
    std::fflush(stdout);

    FILE *fdat = std::fopen(argv[7], "w");
    if (!fdat) {
        std::fprintf(stderr, "irbc_main: cannot open %s\n", argv[7]);
        irbc_destroy(c);
        return 1;
    }
    /* frames after steps F, 2F, ... counted from step 0 (maincode.cpp:474-477,746-761) */
    const auto t0 = std::chrono::steady_clock::now();
    int64_t done = 0;
    int frame = 0;
    while (done < steps) {
        const int64_t next = (int64_t)(frame + 1) * frame_every + 1;
        const int64_t n = std::min(next, steps) - done;
        rc = irbc_md_steps(c, n);
        if (rc) return die("irbc_md_steps", rc);
        rc = irbc_sync(c);
        if (rc) return die("irbc_sync", rc);
        done += n;
        if (done == next) {
            ++frame;
            rc = irbc_host::write_frame(c, 0, raw.nv(), frame, fdat);
            if (rc) return die("writing a frame", rc);
            if (stl_flag == 1) {
                char name[64];
                std::snprintf(name, sizeof name, "rbc%d.stl", frame);
                rc = irbc_host::write_stl(c, 0, s.mesh, p, name);
                if (rc) return die("writing an STL file", rc);
            }
        }
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double g[13];
    irbc_get_globals(c, 0, g);
    std::fprintf(stderr, "irbc_main: %lld steps in %.3f s (%.1f us/step, %.3g particle-steps/s); "
                         "area %.6f volume %.6f bonds %d\n",
                 (long long)steps, sec, steps ? sec * 1e6 / steps : 0.0, steps ? ns * steps / sec : 0.0, g[0], g[1],
                 (int)g[9]);
    std::fclose(fdat);
    irbc_destroy(c);
    return 0;
}
