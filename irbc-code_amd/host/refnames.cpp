/*
 * refnames.cpp — reference-named functions over the C ABI (see refnames.h).
 */
#include "refnames.h"

#include <cstdio>
#include <cstdlib>

double Sarea = 0.0, Volume = 0.0;

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
