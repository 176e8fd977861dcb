/* aniso_weights.h -- the anisotropic weight table, MI355X implementation.
 *
 * Drop-in for the reference's src/aniso_weights.h for the entry points boltz_
 * uses.  glance = 0 is the Landau table (gHat3L), generated on the GPU with the
 * reference's nested 64-point Gauss-Legendre rules.  The glancing-collision
 * table (glance > 0) is written by the weight generator
 * (spbte_op_generate_weights_glance, spectralbte_amd/weightgen_amd) and loaded
 * here.  The
 * GSL-typed integrand callbacks of the reference file (ghat_theta_AnIso,
 * ghat_theta2, ghat_phi_AnIso, ghat_r_AnIso, ghatL2, ghatL, gHat3_AnIso) are
 * internal to it and not exported.
 */
#ifndef _ANISO_WEIGHTS_H
#define _ANISO_WEIGHTS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/aniso_weights.h:1 (src/aniso_weights.c:37-146).  The third
 * argument is L_v and the fourth lambda, as the definition and its caller
 * exec/boltz.c:154 have them (the reference header names them the other way
 * round).  Allocates the N^3 rows of conv_weights, loads
 * Weights/N%d_Aniso_L_v%g_lambda%g_Landau.wts (glance 0) or
 * Weights/N%d_AnIso_L_v%g_lambda%g_glance%g.wts (glance > 0) if present and
 * weightFlag == 0, else generates the Landau table on the GPU and stores it
 * under that name.  That is the Landau table for glance > 0 too: the
 * reference's generator reads a file-scope glance that is never assigned
 * (src/aniso_weights.c:31, :65-68, :417). */
void initialize_weights_AnIso(int nodes, double *zeta, double Lv, double lam, int weightFlag, double **conv_weights,
                              double glance);

/* Replaces src/aniso_weights.h:19 (src/aniso_weights.c:371-398): one Landau
 * entry for row mode zeta and column mode xi, on the host, with the
 * parameters of the last initialize_weights_AnIso call. */
double gHat3L(double zeta1, double zeta2, double zeta3, double xi1, double xi2, double xi3);

/* Replaces src/aniso_weights.h:21 (src/aniso_weights.c:400-428): fills the
 * rows for the last initialize_weights_AnIso call's grid with the Landau
 * table, on the GPU. */
void generate_conv_weights_AnIso(double **conv_weights);

#ifdef __cplusplus
}
#endif

#endif
