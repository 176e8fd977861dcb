/* conserve.h -- mass/momentum/energy projection, MI355X implementation.
 *
 * Drop-in for the reference's src/conserve.h.  C C^T and its LU factors are
 * built once on the host in FP64 (as the reference does); conserveAllMoments
 * runs the moment reduction, the (Ns+4)x(Ns+4) solve and the correction on
 * the GPU in FP64.
 */
#ifndef _CONSERVE_H
#define _CONSERVE_H

#include "species.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/conserve.h:6 (src/conserve.c:17-43).  vel and mix are borrowed.
 * mix[0].mass also sets the mass (and the KB=1 rule of
 * src/momentRoutines.c:26-29) used by ComputeQ_maxPreserve's Maxwellian. */
void initialize_conservation(int nodes, double h_v, double *vel, species *mix, int num_spec);

/* Replaces src/conserve.h:8 (src/conserve.c:45-73): one species of mass 1. */
void initialize_conservation_fast(int nodes, double h_v, double *vel);

/* Replaces src/conserve.h:10 (src/conserve.c:76-84). */
void dealloc_conservation(void);

/* Replaces src/conserve.h:12 (src/conserve.c:89-168): in-place scaled-pivot LU
 * of the module's C C^T; det is unused, as in the reference.  A zero pivot
 * column prints "Exiting factorMatrixintoLU...." and exits, as the reference. */
void factorMatrixIntoLU(int nElem, double det);

/* Replaces src/conserve.h:14 (src/conserve.c:173-202): solves (C C^T) x = b in
 * place with the module's LU factors. */
void solveWithCCt(int nElem, double *b);

/* Replaces src/conserve.h:16 (src/conserve.c:207-264): Q[n*Ns+m] (Ns*Ns host
 * arrays of N^3) <- Q - C_m^T (C C^T)^-1 sum C Q / Ns.  Runs on the GPU. */
void conserveAllMoments(double **Q);

/* Replaces src/conserve.h:18 (src/conserve.c:268-317): builds C C^T for the
 * module's grid and species and factors it. */
void createCCtAndPivot(void);

#ifdef __cplusplus
}
#endif

#endif
