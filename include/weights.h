/* weights.h -- convolution-weight table, MI355X implementation.
 *
 * Drop-in for the reference's src/weights.h.  The table keeps the reference's
 * host layout (N^3 malloc'd rows of N^3 doubles, row = output mode) and file
 * format (Weights/N%d_isotropic_L_v%g_lambda%g.wts, raw native-endian FP64,
 * no header).  Generation runs on the GPU (closed forms for lambda = 0 and 1,
 * Gauss-Legendre quadrature otherwise) instead of per-entry GSL QAG.
 */
#ifndef _WEIGHTS_H
#define _WEIGHTS_H

#include "species.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/weights.h:12 (src/weights.c:26-31). */
void alloc_weights(int N, double ****conv_weights, int total_species);

/* Replaces src/weights.h:37 (src/weights.c:33-125): allocates the N^3 rows,
 * loads Weights/<name> if present and weightFlag == 0, else generates and
 * stores it.  isoFlag = 1 loads N%d_AnIso_L_v%g_lambda%g_Landau.wts and exits
 * with the reference's message if it is missing.  A short file prints
 * "Error reading weight file" and exits(1), as the reference. */
void initialize_weights(int nodes, double *eta, double Lv, double lam, int weightFlag, int isoFlag, double **conv_weights,
                        species species_i, species species_j);

/* Replaces src/weights.h:43 (src/weights.c:265-281): fills conv_weights for
 * the parameters of the last initialize_weights call, on the GPU. */
void generate_conv_weights_iso(double **conv_weights);

/* Replaces src/weights.h:45 (src/weights.c:127-133).  Also drops the GPU copy. */
void dealloc_weights(int N, double **conv_weights);

/* Replaces src/weights.h:47 (src/weights.c:156-160): the radial integrand,
 * args = {mu|zeta|, |xi|, |xi - mu zeta|}. */
double ghat(double r, void *args);

/* Replaces src/weights.h:49 (src/weights.c:181-262): prefactor times the
 * radial integral for one (xi, zeta) pair, evaluated in closed form
 * (lambda 0, 1) or by composite Gauss-Legendre quadrature. */
double gHat3(double ki1, double ki2, double ki3, double zeta1, double zeta2, double zeta3);

#ifdef __cplusplus
}
#endif

#endif
