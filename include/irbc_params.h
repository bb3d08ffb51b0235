/*
 * irbc_params.h — runtime parameters of one iRBC adhesive-dynamics replica.
 *
 * Every field is a compile-time macro or a hard-coded literal in the
 * reference; here they are runtime values so that replicas of different
 * sizes and rates can share one batch.  irbc_params_default() fills in the
 * reference values:
 *   define.h:35-73        box, solvent, membrane, adhesion and parasite constants
 *   maincode.cpp:77-78    ligand grid spacing 0.5
 *   maincode.cpp:347-348  elap = 0.0005, shear modulus = 1200
 *   maincode.cpp:455,460  csteps = (int)(dtcd/dtmd), uplist = (int)(0.5/(walls*dtmd))
 *   maincode.cpp:463-470  wall-speed ramp: 0.1 for the first 250000 steps, then walls
 *
 * Plain C, no device types: included by the C-ABI header, the host driver and
 * the CPU oracle alike.
 */
#ifndef IRBC_PARAMS_H
#define IRBC_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irbc_params {
    /* box in units of the cell size a = 1 (define.h:42-45) */
    int32_t Lx, Ly, Lz;
    /* cadences (maincode.cpp:455,460); irbc_params_derive() recomputes them */
    int32_t csteps, uplist;
    /* solvent (define.h:46-54) */
    double navg;       /* particles per cell */
    double mass;       /* solvent particle mass */
    double Theta;      /* kT */
    double dtcd;       /* collision interval */
    double dtmd;       /* MD time step */
    double cosalpha;   /* SRD rotation angle, cos and sin (135 deg, truncated) */
    double sinalpha;
    double walls;      /* top-wall speed after the ramp; also sets uplist */
    double ramp_speed; /* top-wall speed during the ramp */
    int64_t ramp_steps;
    /* membrane (define.h:57-69, maincode.cpp:347-348) */
    double kappasa, kappad, kappav, kappab;
    double pmass;      /* vertex mass = navg * mass */
    double ks, dthr, FD, Fg, el0;
    double elap;       /* persistence length of the WLC springs */
    double shear;      /* target shear modulus used for kappap */
    /* parasite bead (define.h:72-73) */
    double PRmass, paraR;
    /* adhesion (argv[1], argv[2]; maincode.cpp:73-74) */
    double onrate, offrate;
    double lig_spacing;
    /* the reference's truncated constant (define.h:35) */
    double Pi;
} irbc_params;

/* Reference defaults (define.h, maincode.cpp literals); onrate/offrate are
 * the README run's 0.005 / 0.015. */
static inline void irbc_params_default(irbc_params *p)
{
    p->Lx = 60; p->Ly = 40; p->Lz = 40;
    p->navg = 10.0; p->mass = 1.0; p->Theta = 1.0;
    p->dtcd = 0.02; p->dtmd = 0.002;
    p->cosalpha = -0.70710678118; p->sinalpha = 0.70710678118;
    p->walls = 3.0; p->ramp_speed = 0.1; p->ramp_steps = 250000;
    p->kappasa = 60000.0; p->kappad = 1200.0; p->kappav = 6000.0; p->kappab = 100.0;
    p->pmass = p->navg * p->mass;
    p->ks = 6000.0; p->dthr = 0.50; p->FD = 1200.0; p->Fg = 40.0; p->el0 = 0.1;
    p->elap = 0.0005; p->shear = 1200.0;
    p->PRmass = 10.0; p->paraR = 1.7;
    p->onrate = 0.005; p->offrate = 0.015;
    p->lig_spacing = 0.5;
    p->Pi = 3.141592653;
    p->csteps = (int32_t)(p->dtcd / p->dtmd);
    p->uplist = (int32_t)(0.5 / (p->walls * p->dtmd));
}

/* Recompute the derived cadences after editing dtcd, dtmd or walls. */
static inline void irbc_params_derive(irbc_params *p)
{
    p->csteps = (int32_t)(p->dtcd / p->dtmd);
    p->uplist = (int32_t)(0.5 / (p->walls * p->dtmd));
    if (p->uplist < 1) p->uplist = 1;
}

/* Top-wall speed at a given step (maincode.cpp:463-470). */
static inline double irbc_wallspeed(const irbc_params *p, int64_t step)
{
    return step < p->ramp_steps ? p->ramp_speed : p->walls;
}

/* Number of cells: an extra z layer below the box holds the shifted grid
 * (maincode.cpp:113-136): Lx * Ly * (Lz + 1). */
static inline int64_t irbc_ncells(const irbc_params *p)
{
    return (int64_t)p->Lx * p->Ly * (p->Lz + 1);
}

#ifdef __cplusplus
}
#endif

#endif /* IRBC_PARAMS_H */
