/*
 * refnames.cpp — reference-named functions over the C ABI (see refnames.h).
 */
#include "refnames.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

double Sarea = 0.0, Volume = 0.0, Sarea0 = 0.0, Volume0 = 0.0;
double center[3] = {0.0, 0.0, 0.0};

static irbc_ctx *g_ctx = nullptr;
static int g_rep = 0;

void irbc_ref_bind(irbc_ctx *ctx, int replica)
{
    g_ctx = ctx;
    g_rep = replica;
}

/* FatalError's convention: message, then exit */
static void check(int rc, const char *what)
{
    if (rc == IRBC_OK) return;
    std::fprintf(stderr, "%s: %s (%d)\n", what, g_ctx ? irbc_last_error() : "no context bound", rc);
    std::exit(1);
}

static irbc_ctx *ctx()
{
    if (!g_ctx) check(IRBC_ERR_STATE, "irbc_ref_bind not called");
    return g_ctx;
}

void CollisionStep() { check(irbc_collision_step(ctx()), "CollisionStep"); }
void llistupdate() { check(irbc_llist_update(ctx()), "llistupdate"); }
void sortsolvent_initial1() { check(irbc_classify_solvent(ctx()), "sortsolvent_initial1"); }
void streamstep() { check(irbc_stream(ctx()), "streamstep"); }

void TotalVA() { check(irbc_membrane_va(ctx(), g_rep, &Sarea, &Volume), "TotalVA"); }

/* the membrane's forces with the area/volume of the last TotalVA(), added
 * to the vertex forces, as in the reference */
void force_calculate_dt() { check(irbc_membrane_phases(ctx(), IRBC_PH_MEMBRANE), "force_calculate_dt"); }

/* the bead-membrane repulsion: vertex forces and the bead's */
void pfdt() { check(irbc_membrane_phases(ctx(), IRBC_PH_PARASITE), "pfdt"); }

void externalforces() { check(irbc_membrane_phases(ctx(), IRBC_PH_EXTERNAL), "externalforces"); }

void bond_dynamics() { check(irbc_bond_step(ctx()), "bond_dynamics"); }

void mdsteps(long n)
{
    check(irbc_md_steps(ctx(), n), "mdsteps");
    check(irbc_sync(ctx()), "mdsteps");
}

/* the current shape's area and volume as the rest values: the reference
 * calls it once, on the set-up mesh (maincode.cpp:344), which is where
 * irbc_set_mesh takes the device's own rest values from */
void TotalVA0() { check(irbc_membrane_va(ctx(), g_rep, &Sarea0, &Volume0), "TotalVA0"); }

/* the membrane centroid, folded into the box (Src/memfuncs.cpp:55-74) */
void centerofmass()
{
    irbc_replica_spec s;
    check(irbc_get_spec(ctx(), g_rep, &s), "centerofmass");
    const int nv = s.n_vertices;
    std::vector<double> x(nv), y(nv), z(nv);
    std::vector<int32_t> ix(nv), iy(nv);
    check(irbc_get_vertices(ctx(), g_rep, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr, nullptr, ix.data(), iy.data(), nullptr, nullptr, nullptr, nullptr),
          "centerofmass");
    const double L[3] = {(double)s.params.Lx, (double)s.params.Ly, (double)s.params.Lz};
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nv; ++i) {
        a[0] += x[i] + ix[i] * L[0];
        a[1] += y[i] + iy[i] * L[1];
        a[2] += z[i];
    }
    for (int k = 0; k < 3; ++k) center[k] = std::fmod(a[k] / (double)nv, L[k]);
}

/* a solvent coordinate wrapped into [0, lx] by whole box lengths
 * (Src/solfuncs.cpp:603-622; the device wraps inside its streaming) */
double bbound(double x, double lx)
{
    while (x > lx) x -= lx;
    while (x < 0) x += lx;
    return x;
}
