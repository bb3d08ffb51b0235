/*
 * irbc.h — C ABI of the MI355X-native iRBC adhesive-dynamics step.
 *
 * One context owns a batch of independent replicas on one GPU; a single
 * replica (the reference's whole program) is a batch of one.  All state
 * lives in device memory; the host side of this ABI moves plain host arrays
 * in and out (SoA, the reference's units, FP64 on the host side) and
 * launches hand-written HIP kernels for gfx950.
 *
 * Every call returns an int status: IRBC_OK (0) or a negative IRBC_ERR_*.
 * No exceptions, no exit().  A context is used by one host thread at a time;
 * calls are ordered on the context's HIP streams.  Different contexts may be
 * driven from different threads at once: a graph capture inside
 * irbc_md_steps excludes the other threads' calls for its duration.  Context
 * memory is stream-ordered from the device's default pool, which keeps freed
 * memory for the next context.
 *
 * Reference surface replaced (citations are to the reference sources):
 *   define.h:35-73 macros + maincode.cpp literals   -> irbc_params (irbc_params.h)
 *   define.h:182-429 globals (solvent, Cells, vts,
 *     edges, tries, ligands, parasite)               -> irbc_ctx + set/get calls
 *   CollisionStep()        Include/solfuncs.h:21     -> irbc_collision_step
 *   llistupdate()          Include/solfuncs.h:22     -> irbc_llist_update
 *   reflectionFed()/Solve()/bbound(), the streaming
 *     loop maincode.cpp:510-562                      -> irbc_stream
 *   TotalVA()              Include/memfuncs.h:25     -> irbc_membrane_va
 *   force_calculate_dt(), pfdt(),
 *     maincode.cpp:606-665                           -> irbc_membrane_forces
 *   bond_dynamics() maincode.cpp:16 + :668-723       -> irbc_bond_step
 *   the time loop maincode.cpp:461-764               -> irbc_md_steps
 *   sortsolvent_initial1() Include/solfuncs.h:20     -> irbc_classify_solvent
 *   TotalVA0/makeiedge/centerofmass + maincode.cpp
 *     :228-375 mesh setup                            -> irbc_set_mesh
 *   CRandomMersenne / StochasticLib1 streams         -> irbc_set_*_tape (parity
 *                                                       harness) or the built-in
 *                                                       counter-based streams
 */
#ifndef IRBC_H
#define IRBC_H

#include <stdint.h>

#include "irbc_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IRBC_OK 0
#define IRBC_ERR_ARG -1      /* bad argument (null, out of range, size mismatch) */
#define IRBC_ERR_OOM -2      /* device or host allocation failed */
#define IRBC_ERR_HIP -3      /* a HIP runtime call failed */
#define IRBC_ERR_STATE -4    /* call out of order (e.g. no mesh loaded) */
#define IRBC_ERR_DEVICE -5   /* device code raised a status flag (NaN, overflow) */
#define IRBC_ERR_BUDGET -6   /* request does not fit the memory budget */
#define IRBC_ERR_NODEVICE -7 /* no HIP device */

typedef struct irbc_ctx irbc_ctx;

/* One replica of a batch: its parameters and its sizes. */
typedef struct irbc_replica_spec {
    irbc_params params;
    int64_t n_solvent;
    int32_t n_vertices, n_edges, n_triangles;
    int32_t pad;
    uint64_t seed; /* key of the counter-based streams */
} irbc_replica_spec;

/* ---- context ---- */
int irbc_create(const irbc_replica_spec *specs, int n_replicas, int device, irbc_ctx **out);
int irbc_destroy(irbc_ctx *ctx);
/* replica r's spec as created: sizes, parameters, seed (the box and mesh
 * sizes the reference keeps in define.h:42-45 and its mesh files) */
int irbc_get_spec(const irbc_ctx *ctx, int r, irbc_replica_spec *out);
const char *irbc_last_error(void);
int irbc_device_bytes(const irbc_ctx *ctx, int64_t *bytes);
/* bytes a batch of these replicas would take on the device (admission) */
int64_t irbc_estimate_bytes(const irbc_replica_spec *specs, int n_replicas);
int irbc_sync(irbc_ctx *ctx);  /* waits, then reports device status flags */

/* ---- scenario setup ---- */
/* mesh in box coordinates; edges (ne x 2) and triangles (nt x 3), 0-based.
 * Builds hinge partners, rest lengths, kappap, rest areas, A0, V0, theta0
 * and hinge orientation as maincode.cpp:228-387 does. */
int irbc_set_mesh(irbc_ctx *ctx, int r, const double *xyz, const int32_t *edges,
                  const int32_t *tris);
/* any pointer may be NULL to leave that field as it is */
int irbc_set_vertices(irbc_ctx *ctx, int r, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *fx,
                      const double *fy, const double *fz, const int32_t *ix, const int32_t *iy,
                      const int32_t *pid, const int32_t *bondi, const int32_t *bondl,
                      const int32_t *knob);
int irbc_get_vertices(irbc_ctx *ctx, int r, double *x, double *y, double *z, double *vx,
                      double *vy, double *vz, double *fx, double *fy, double *fz, double *poff,
                      double *pon, int32_t *ix, int32_t *iy, int32_t *pid, int32_t *bondi,
                      int32_t *bondl, int32_t *knob);
/* solvent state in particle order; indio may be NULL (all outside) */
int irbc_set_solvent(irbc_ctx *ctx, int r, int64_t n, const double *x, const double *y,
                     const double *z, const double *vx, const double *vy, const double *vz,
                     const int32_t *indio);
/* a fresh solvent drawn on the device from the replica's counter-based
 * streams (keyed by its seed): n_solvent positions uniform in the box,
 * velocities Maxwellian at kT = Theta (maxwell = 1) or uniform in
 * +-sqrt(3 kT / m) as maincode.cpp:96-107 draws them (maxwell = 0), stored
 * in cell order; inside flags all outside until irbc_classify_solvent.  No
 * counterpart in the reference, whose MT19937 draws irbc_set_solvent takes
 * from the host; for set-up without a host round trip (request loops). */
int irbc_init_solvent(irbc_ctx *ctx, int r, int maxwell);
/* positions/velocities as stored (FP32, widened), last shifted-cell key and
 * inside flag, all in particle order; any pointer may be NULL */
int irbc_get_solvent(irbc_ctx *ctx, int r, double *x, double *y, double *z, double *vx,
                     double *vy, double *vz, int32_t *key, int32_t *indio);
/* parasite: d[12] = pos, vel, force(t), force(t+dt); i[3] = image x, image y, cell */
int irbc_set_parasite(irbc_ctx *ctx, int r, const double *d, const int32_t *i);
int irbc_get_parasite(irbc_ctx *ctx, int r, double *d, int32_t *i);
int64_t irbc_n_ligands(const irbc_ctx *ctx, int r);
int irbc_set_ligands(irbc_ctx *ctx, int r, const int32_t *ligi);
int irbc_get_ligands(irbc_ctx *ctx, int r, int32_t *ligi);
/* g[13] = Sarea, Volume, Sarea0, Volume0, theta0, wallspeed, status,
 * bond draws of the last step, step, bonded knobs, particles queued for
 * bounce-back in the last step, the XCD the replica's last membrane chain
 * ran on (-1: none, or device-scope hand-offs), how the last collision ran
 * (1: sorted, or none yet; 2: in place, see IRBC_OPT_SORT_EVERY) */
int irbc_get_globals(irbc_ctx *ctx, int r, double *g);
/* the time-loop counters before the top-of-step increment (maincode.cpp:450-475) */
int irbc_set_counters(irbc_ctx *ctx, int r, int64_t step, int32_t colsteps, int32_t upd);
int irbc_get_counters(irbc_ctx *ctx, int r, int64_t *step, int32_t *colsteps, int32_t *upd);
int irbc_set_wallspeed(irbc_ctx *ctx, int r, double wallspeed);

/* ---- random streams ---- */
int irbc_set_seed(irbc_ctx *ctx, int r, uint64_t seed);
/* parity harness: the next collision / bond step reads these tapes.  Layout:
 * shift_u[3]; cell_u[2*ncell]; wall_z[3*ncell]; mem_z[3*ncell]; bond u[n]
 * (see oracle/oracle.h for the draw order they stand for) */
int irbc_use_tapes(irbc_ctx *ctx, int r, int on);
int irbc_set_collision_tape(irbc_ctx *ctx, int r, const double *shift_u, const double *cell_u,
                            const double *wall_z, const double *mem_z);
int irbc_set_bond_tape(irbc_ctx *ctx, int r, const double *u, int64_t n);

/* ---- stages (every replica of the batch) ---- */
int irbc_collision_step(irbc_ctx *ctx);
int irbc_llist_update(irbc_ctx *ctx);
int irbc_stream(irbc_ctx *ctx);
int irbc_classify_solvent(irbc_ctx *ctx);
#define IRBC_PH_KICK1 1
#define IRBC_PH_DRIFT 2
#define IRBC_PH_VA 4
#define IRBC_PH_FORCES 8 /* IRBC_PH_PARASITE then IRBC_PH_MEMBRANE, one pass */
#define IRBC_PH_EXTERNAL 16
#define IRBC_PH_BONDS 32
#define IRBC_PH_KICK2 64
#define IRBC_PH_COUNTERS 128
/* pfdt() alone (Src/memfuncs.cpp:443-484, Include/memfuncs.h:28): the
 * bead-membrane repulsion, subtracted from the vertex forces and added to
 * the bead's */
#define IRBC_PH_PARASITE 512
/* force_calculate_dt() alone (Include/memfuncs.h:27): springs, bending,
 * area and volume terms added to the vertex forces, with the area and
 * volume of the last IRBC_PH_VA */
#define IRBC_PH_MEMBRANE 1024
int irbc_membrane_phases(irbc_ctx *ctx, int phases);
int irbc_membrane_va(irbc_ctx *ctx, int r, double *area, double *volume);
int irbc_membrane_forces(irbc_ctx *ctx);
int irbc_bond_step(irbc_ctx *ctx);
/* per-cell results of the last collision: vcm[3*ncell], nc, ncc, rand1,
 * rand2, dvv, ThE; any pointer may be NULL */
int irbc_get_cells(irbc_ctx *ctx, int r, double *vcm, int32_t *nc, double *ncc, double *rand1,
                   double *rand2, double *dvv, double *the);
/* cell -> triangle lists: off[ncell+1], ent[cap]; returns entries or < 0 */
int64_t irbc_get_llist(irbc_ctx *ctx, int r, int64_t *off, int32_t *ent, int64_t cap);
/* ligand candidates of the last bond step: cnt[nv], lig[cap]; returns total */
int64_t irbc_get_pairs(irbc_ctx *ctx, int r, int32_t *cnt, int32_t *lig, int64_t cap);

/* ---- the time loop ---- */
int irbc_md_steps(irbc_ctx *ctx, int64_t nsteps);
/* launch-shape options (no counterpart in the reference):
 * IRBC_OPT_MEMBRANE_FUSED  1: the membrane part of a step is one workgroup
 *   per replica; 0 (default): five grid-wide kernels.  Both give the same
 *   forces up to the summation order of the global area and volume. */
#define IRBC_OPT_MEMBRANE_FUSED 1
/* IRBC_OPT_GRAPHS  1 (default): irbc_md_steps captures chunks of 10 steps
 *   as HIP graphs (one per chunk shape) and replays them; 0: launches
 *   kernel by kernel.  Ignored while event profiling is on. */
#define IRBC_OPT_GRAPHS 2
/* IRBC_OPT_NEAR_MERGED  1 (default for fewer than 8 replicas): bounce-back
 *   off the membrane in one kernel, 16 lanes per particle; 0 (default for 8
 *   or more): a bounding-sphere pass, four lanes per particle, then a hit
 *   pass over the particles that can reach a triangle.  Same results. */
#define IRBC_OPT_NEAR_MERGED 3
/* IRBC_OPT_XCD_CHUNK  1 (default): the binning, scatter and cell kernels
 *   hand each of the 8 XCDs a contiguous eighth of the tiles; 0: plain
 *   round-robin dealing.  Same results. */
#define IRBC_OPT_XCD_CHUNK 4
/* IRBC_OPT_STREAM4  1: bulk streaming with four consecutive particles per
 *   lane (16-byte loads and stores); 0 (default): one per lane.  Same
 *   results. */
#define IRBC_OPT_STREAM4 5
/* IRBC_OPT_STREAM_GRID  1: bulk streaming on a grid of 4 workgroups per CU,
 *   each looping over tiles, which leaves room on every CU for the membrane
 *   kernels running beside it; 0 (default): one workgroup per tile.  Same
 *   results. */
#define IRBC_OPT_STREAM_GRID 6
/* IRBC_OPT_MEM_AFTER_STREAM  1: the membrane part of a step starts after
 *   that step's bulk streaming and overlaps the bounce-back; 0 (default): it
 *   starts after the previous step's bulk streaming (on collision and list
 *   steps after this step's collision and lists) and overlaps this step's
 *   streaming.  Same results. */
#define IRBC_OPT_MEM_AFTER_STREAM 7
/* IRBC_OPT_MEMBRANE_CHAIN  1 (default): the membrane part of a step is one
 *   launch whose workgroups take their phase (begin, element terms, triangle
 *   terms, gather, finish) by ticket and hand over through device-scope
 *   release/acquire; it takes precedence over IRBC_OPT_MEMBRANE_FUSED.
 *   0: IRBC_OPT_MEMBRANE_FUSED decides.  Same results as the five grid
 *   kernels, bit for bit. */
#define IRBC_OPT_MEMBRANE_CHAIN 8
/* IRBC_OPT_CHAIN_DEVICE_SCOPE  1: the chain's phase hand-offs release and
 *   acquire at device scope (L2 write-back and invalidation at each);
 *   0 (default): XCD-local hand-offs (every workgroup of a replica runs on
 *   one XCD, whose L2 is their point of coherence).  Same results. */
#define IRBC_OPT_CHAIN_DEVICE_SCOPE 9
/* IRBC_OPT_NEAR_MASKS  1 (default): a particle goes to the bounce-back
 *   kernels only when its 1/4-cell sub-cell lies near a listed triangle or
 *   it travels far this step; 0: every particle of a cell that lists a
 *   triangle goes (a check of the masks).  Same results. */
#define IRBC_OPT_NEAR_MASKS 10
/* IRBC_OPT_ONE_QUEUE  1 (default): with the membrane chain, a step is two
 *   launches on one queue -- the chain's force phases, the triangle
 *   kinematics and the bulk streaming in one grid, the bounce-back and the
 *   membrane finish in the next; 0: the membrane on a second HIP stream with
 *   events between the two.  Same results. */
#define IRBC_OPT_ONE_QUEUE 11
/* IRBC_OPT_CLASSIFY_ALL  1: irbc_classify_solvent tests every triangle for
 *   every particle, as sortsolvent_initial1 does; 0 (default): only the
 *   triangles whose (y, z) box holds the particle's (y, z), the only ones
 *   its +x ray can count.  Same flags. */
#define IRBC_OPT_CLASSIFY_ALL 12
/* IRBC_OPT_ALIGN_LISTS  1: a replica updates its cell->triangle lists at
 *   the last collision step before its uplist deadline (or at the
 *   deadline), so that list work falls on collision steps; 0 (default): the
 *   reference's cadence, every uplist steps.  The same hits while the
 *   membrane moves less than the lists' margin between updates, the
 *   reference's own assumption. */
#define IRBC_OPT_ALIGN_LISTS 13
/* IRBC_OPT_SORT_EVERY  value k >= 1 (default 1): the collision of a step
 *   sorts the particles into cell order when (step / csteps) is a multiple
 *   of k, i.e. every k-th collision; the others leave each particle in its
 *   slot and reduce the cell sums with LDS-aggregated 64-bit atomics.  The
 *   cell sums are fixed point, so the results are the same for every k.
 *   On an MI355X the in-place collision is the slower one (about 160 us
 *   against 111 us on a 64x64x48 box: its four atomics per cell and tile
 *   outweigh the sort's data movement), hence the default.
 *   Stage-by-stage collisions (irbc_collision_step) always sort. */
#define IRBC_OPT_SORT_EVERY 14
/* IRBC_OPT_BIN_SPAN  value 1 (default), 2 or 4: the bin-sort tile shape --
 *   particle tiles (256 particles each) per binning workgroup of a sorting
 *   collision, counted in one LDS hash so a cell costs one global atomic per
 *   span; irbc_autotune picks it.  Collisions that run in place
 *   (IRBC_OPT_SORT_EVERY > 1) bin one tile per workgroup.  Same results. */
#define IRBC_OPT_BIN_SPAN 15
int irbc_set_option(irbc_ctx *ctx, int option, int value);
int irbc_get_option(const irbc_ctx *ctx, int option, int *value);
/* Times `steps` MD steps under 48 combinations of the options above on
 * this context's own replicas, times the four fastest again twice each
 * over at least four collisions and keeps the fastest of those, then on
 * that shape times the bin spans 1, 2 and 4 (IRBC_OPT_BIN_SPAN, over at
 * least four collisions) and keeps the fastest: the membrane as the
 * chain on one queue, the chain on its own stream or the five grid kernels
 * on its own stream, times bounce-back shape, graphs, XCD chunking and
 * stream shape.  The state is snapshot on the device first and restored
 * afterwards, so the trajectory is not advanced.  report (may be NULL)
 * receives 512 ms-per-step figures, index = fused + 2 near_merged +
 * 4 graphs + 8 xcd_chunk + 16 stream4 + 32 stream_grid + 64 mem_after +
 * 128 mem_chain + 256 one_queue, NaN where not timed.  Returns the chosen
 * index or < 0.  Needs free device memory for one copy of the context. */
int irbc_autotune(irbc_ctx *ctx, int32_t steps, double *report);
/* the autotuner's second stage, on the launch shape it chose: ms per step
 * with bin spans 1, 2, 4 (NaN when not timed) */
int irbc_autotune_spans(const irbc_ctx *ctx, double *ms);
/* keep per-cell results (irbc_get_cells) also for collisions run inside
 * irbc_md_steps; costs ~56 B per cell per collision */
int irbc_set_diagnostics(irbc_ctx *ctx, int on);

/* ---- checkpoint / resume (no counterpart in the reference) ---- */
/* A replica's resumable state as one flat byte buffer: solvent in its
 * current order, membrane, bonds, ligand occupancy, parasite, step
 * counters, stream key and the cell->triangle lists of the last update.
 * Loading it into a replica of the same sizes (any context, any slot) and
 * continuing gives the same bits as never stopping.  Save between calls of
 * irbc_md_steps. */
int64_t irbc_checkpoint_bytes(irbc_ctx *ctx, int r);
/* returns the bytes written or < 0 */
int64_t irbc_checkpoint_save(irbc_ctx *ctx, int r, void *buf, int64_t cap);
/* the replica's mesh must be loaded (irbc_set_mesh) first */
int irbc_checkpoint_load(irbc_ctx *ctx, int r, const void *buf, int64_t n);
/* A replica moved device to device: replica rs of src into slot rd of dst
 * (same device, same sizes and parameters): the state a checkpoint holds,
 * the mesh (dst needs no irbc_set_mesh) and the time series (cadence,
 * capacity, records).  src is left as it was.  dst then continues bit for
 * bit as src would have.  A request loop re-packs its running replicas
 * with it into a new context when requests arrive or finish (continuous
 * batching in one launch sequence). */
int irbc_copy_replica(irbc_ctx *dst, int rd, irbc_ctx *src, int rs);

/* ---- time series and asynchronous use (request loops) ---- */
/* every > 0: after every `every`-th step of irbc_md_steps record
 * IRBC_SERIES_REC = 8 doubles (step, membrane centroid x y z with periodic
 * images unwrapped, centroid velocity vx vy vz, bonded knobs) into a device
 * buffer of `cap` records; every = 0: off.  Resets the count.  No
 * counterpart in the reference, which writes frames instead
 * (maincode.cpp:746-754). */
int irbc_set_series(irbc_ctx *ctx, int r, int32_t every, int32_t cap);
/* copies min(count, cap) records; returns the count or < 0 */
int64_t irbc_get_series(irbc_ctx *ctx, int r, double *out, int64_t cap);
/* 1: all launched work of the context has finished; 0: still running */
int irbc_query(irbc_ctx *ctx);

/* ---- measurement: HIP-event timing on the context's stream ---- */
/* every > 0: bracket the launches of every every-th step of irbc_md_steps
 * with events; 0: off.  Resets the accumulators. */
int irbc_profile(irbc_ctx *ctx, int every);
/* t[0] = total ms of timed bulk-stream launches, t[1] = their count,
 * t[2] = total ms of collision sequences, t[3] = their count,
 * t[4] = total ms of membrane launches, t[5] = their count,
 * t[6] = total ms of near-membrane stream launches, t[7] = their count */
int irbc_profile_read(irbc_ctx *ctx, double *t);

/* ---- TotalVA with no allocation (no counterpart in the reference) ---- */
/* For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * global area and volume of TotalVA() (Src/memfuncs.cpp:161-190) of a
 * membrane held in the caller's device memory, with the same arithmetic and
 * summation order as irbc_membrane_va, so the same bits.  x, y, z (FP64) and
 * ix, iy (periodic images) have nv entries, tris is nt x 3 (0-based), all
 * on the calling thread's current device; area and volume go to out[0],
 * out[1] (device memory) in `stream` order (a hipStream_t; NULL: the null
 * stream).  Calls no allocating routine: the per-block partials live in a
 * fixed device array of 1024 blocks of 256 triangles, so nt <= 262144
 * (IRBC_ERR_ARG beyond).  A triangle index outside [0, nv) makes the result
 * NaN.  Every call on a device shares that device's array: calls that
 * overlap on one device race. */
int irbc_membrane_va_static(const double *x, const double *y, const double *z, const int32_t *ix,
                            const int32_t *iy, int32_t nv, const int32_t *tris, int32_t nt, double Lx, double Ly,
                            double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* IRBC_H */
