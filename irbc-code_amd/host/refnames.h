/*
 * refnames.h — the reference's function names over the MI355X C ABI.
 *
 * A caller written against the reference's call surface
 * (Include/solfuncs.h:19-30, Include/memfuncs.h:17-32, maincode.cpp:16)
 * binds a context once with irbc_ref_bind() and then calls the same
 * zero-argument functions.  Each forwards to the C ABI on the bound replica.
 * As in the reference (FatalError, Src/userintf.cpp:57-81) a failure prints
 * a message and ends the process; callers that need status codes use the C
 * ABI directly.
 */
#ifndef IRBC_HOST_REFNAMES_H
#define IRBC_HOST_REFNAMES_H

#include "irbc.h"

/* the context and replica the reference-named functions act on (C
 * linkage, so an FFI can bind a handle it created) */
extern "C" void irbc_ref_bind(irbc_ctx *ctx, int replica);

/* Src/solfuncs.cpp */
void CollisionStep();        /* Include/solfuncs.h:21 */
void llistupdate();          /* Include/solfuncs.h:22 */
void sortsolvent_initial1(); /* Include/solfuncs.h:20 */
/* the streaming loop of maincode.cpp:510-562 with reflectionFed and bbound */
void streamstep();

/* Src/memfuncs.cpp: results land in the globals below, as in the reference */
void TotalVA();              /* Include/memfuncs.h:25 */
void force_calculate_dt();   /* Include/memfuncs.h:27 */
void pfdt();                 /* Include/memfuncs.h:28 */
/* gravity, bond springs and wall repulsion, maincode.cpp:606-665 */
void externalforces();

/* maincode.cpp:16 and the bond loop :668-723 */
void bond_dynamics();

/* the time loop of maincode.cpp:461-764 for n steps */
void mdsteps(long n);

/* Src/memfuncs.cpp: TotalVA0 (Include/memfuncs.h:26), centerofmass (:20) */
void TotalVA0();
void centerofmass();
/* Src/solfuncs.cpp:603-622 (Include/solfuncs.h:25) */
double bbound(double x, double lx);

extern double Sarea, Volume;   /* set by TotalVA() (define.h:391) */
extern double Sarea0, Volume0; /* set by TotalVA0() (define.h:338) */
extern double center[3];       /* set by centerofmass() (define.h:281) */

#endif
