/*
 * oracle.h — FP64 CPU oracle of the iRBC adhesive-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is
 * compared against; nothing in the product may link, load or call it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It restates the reference's algorithm (reference files cited per function
 * in oracle.cpp) with the reference's operation order, so that at one
 * thread its FP64 results equal the reference's bit for bit.  OpenMP is used
 * only over loops whose results do not depend on iteration order (per
 * particle, per cell), so any thread count gives the same bits.
 *
 * Random streams have two modes:
 *   MT mode   — two MT19937 generators (RanGen[0], sto[0] of the reference)
 *               drawn in the reference's order at one thread; every draw is
 *               also recorded into the tape arrays below;
 *   tape mode — the draws are read from caller-supplied tapes.
 * Tape layout (the stream contract shared with the HIP path):
 *   shift_u[3]        uniforms of the grid shift       (solfuncs.cpp:116-118)
 *   cell_u[2*ncell]   u1, u2 per cell for the axis      (solfuncs.cpp:262-271)
 *   wall_z[3*ncell]   standard normals, wall ghosts     (solfuncs.cpp:178-206)
 *   mem_z[3*ncell]    standard normals, membrane ghosts (solfuncs.cpp:230-245)
 *   bond_u[n]         uniforms of bond_dynamics in draw order (maincode.cpp:16-54)
 */
#ifndef IRBC_ORACLE_H
#define IRBC_ORACLE_H

#include <stdint.h>
#include "../include/irbc_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc *orc_create(const irbc_params *p);
void orc_destroy(orc *o);
void orc_set_threads(int n);
int orc_get_threads(void);
int64_t orc_ncells(const orc *o);

/* ---- streams ---- */
void orc_seed_mt(orc *o, int32_t seed_ran, int32_t seed_sto);
void orc_use_tapes(orc *o, int on);
/* counter-based streams: every draw from Philox4x32-10 keyed by (seed, step,
 * cell or vertex) exactly as the device's production streams are */
void orc_use_philox(orc *o, uint64_t seed);
void orc_philox_u2(uint64_t seed, uint32_t stream, uint64_t index, uint32_t sub, double *u);
void orc_philox_block(uint32_t *ctr4, uint32_t k0, uint32_t k1);
/* FP32 storage: round the solvent's positions and velocities to float where
 * the device stores them (now, on upload, after the solvent rotation and
 * after each streaming step); arithmetic stays FP64 */
void orc_set_fp32(orc *o, int on);
void orc_set_collision_tape(orc *o, const double *shift_u, const double *cell_u,
                            const double *wall_z, const double *mem_z);
void orc_get_collision_tape(const orc *o, double *shift_u, double *cell_u,
                            double *wall_z, double *mem_z);
void orc_set_bond_tape(orc *o, const double *u, int64_t n);
int64_t orc_get_bond_tape(const orc *o, double *u, int64_t cap);
/* raw MT draws for knob placement and tests */
double orc_mt_uniform(orc *o);
uint32_t orc_mt_bits(orc *o);
double orc_sto_normal(orc *o, double m, double s);

/* ---- mesh and state ---- */
int orc_set_mesh(orc *o, int nv, const double *xyz, int ne, const int32_t *edges,
                 int nt, const int32_t *tris);
void orc_set_solvent(orc *o, int64_t n, const double *x, const double *y, const double *z,
                     const double *vx, const double *vy, const double *vz, const int32_t *indio);
int64_t orc_n_solvent(const orc *o);
void orc_get_solvent(const orc *o, double *x, double *y, double *z, double *vx, double *vy,
                     double *vz, int32_t *pid, int32_t *indio);
int orc_n_vertices(const orc *o);
/* dbl: 15 arrays of nv: x y z vx vy vz fx fy fz poff pon (11 used) ...
 * int: ix iy pid bondi bondl knob  */
void orc_set_vertices(orc *o, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz,
                      const double *fx, const double *fy, const double *fz,
                      const int32_t *ix, const int32_t *iy, const int32_t *pid,
                      const int32_t *bondi, const int32_t *bondl, const int32_t *knob);
void orc_get_vertices(const orc *o, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *fx, double *fy, double *fz, double *poff, double *pon,
                      int32_t *ix, int32_t *iy, int32_t *pid, int32_t *bondi, int32_t *bondl,
                      int32_t *knob);
/* parasite: d[12] = para[3], paraV[3], paraft[3], parafdt[3]; i[3] = paradx, parady, paraid */
void orc_set_parasite(orc *o, const double *d, const int32_t *i);
void orc_get_parasite(const orc *o, double *d, int32_t *i);
int64_t orc_n_ligands(const orc *o);
void orc_set_ligands(orc *o, const int32_t *ligi);
void orc_get_ligands(const orc *o, int32_t *ligi);
void orc_get_ligand_xy(const orc *o, double *x, double *y);
/* per-cell results of the last collision */
void orc_get_cells(const orc *o, double *vcm, int32_t *nc, double *ncc, double *rand1,
                   double *rand2, double *dvv, double *the);
/* g[0..9] = Sarea, Volume, Sarea0, Volume0, theta0, wallspeed, center[3], nsolb */
void orc_get_globals(const orc *o, double *g);
/* per-edge/triangle derived topology: l0, kappap (ne), edgest (2*ne), iedge (ne), tarea (nt) */
void orc_get_topology(const orc *o, double *l0, double *kappap, int32_t *edgest, int32_t *iedge,
                      double *tarea);
/* cell -> triangle lists as stored (the reference never shrinks them) */
int64_t orc_llist_total(const orc *o);
void orc_get_llist(const orc *o, int64_t *off, int32_t *ent);
/* candidate ligand pairs of the last bond preparation: cnt[nv], lig[sum cnt] */
int64_t orc_get_pairs(const orc *o, int32_t *cnt, int32_t *lig, int64_t cap);

/* ---- stages (one reference function or inline loop each) ---- */
void orc_collision(orc *o);     /* CollisionStep + CollisionDyn/mem/para */
void orc_llistupdate(orc *o);
void orc_tricm(orc *o);
void orc_stream(orc *o, double wallspeed); /* streaming + bounce-back + walls */
double orc_reflection(orc *o, int64_t i);
void orc_kick1(orc *o);
void orc_drift(orc *o);         /* membrane + parasite drift */
void orc_totalva(orc *o);
void orc_totalva0(orc *o);
void orc_pfdt(orc *o);
void orc_forces(orc *o);        /* force_calculate_dt */
void orc_external(orc *o);      /* gravity, bond springs, wall repulsion */
void orc_bond_prep(orc *o);
void orc_bond_dynamics(orc *o);
void orc_kick2(orc *o);         /* membrane + parasite second half-kick */
void orc_centerofmass(orc *o);
void orc_makeiedge(orc *o);
void orc_sortsolvent(orc *o);   /* inside/outside flags + std::sort by flag */
void orc_classify(orc *o);      /* inside/outside flags only, particle order kept */
void orc_set_wallspeed(orc *o, double w);
int orc_solve_p3(double *x, double a, double b, double c);
double orc_solve(double a, double b, double c, double dtmd);

/* the driver: step counters and one full MD step in the reference's order */
void orc_set_counters(orc *o, int64_t step, int32_t colsteps, int32_t upd);
void orc_get_counters(const orc *o, int64_t *step, int32_t *colsteps, int32_t *upd);
void orc_md_step(orc *o);

/* init of the README run (maincode.cpp:58-446) from raw mesh data; edges and
 * triangles 0-based.  stats[8] = mean edge, SD, Sarea0, Volume0, shear,
 * Young's, Poisson, knobs selected. */
/* bounce-back diagnostics since the last call: out[4] = hits, hits at a
 * negative contact time, smallest such time, plane crossings examined */
void orc_hit_stats(double *out);
int orc_init_readme(orc *o, int32_t seed, int64_t ns, int nv, const double *xyz_raw, int ne,
                    const int32_t *edges, int nt, const int32_t *tris, int nknobs, double *stats);

#ifdef __cplusplus
}
#endif

#endif
