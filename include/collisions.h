/* collisions.h -- spectral collision operator, MI355X implementation.
 *
 * Drop-in for the reference's src/collisions.h: same names, signatures and
 * semantics; the reference's collisions.o is replaced by libspectralbte_amd.
 * All arrays are host arrays owned by the caller; every call is synchronous
 * (results are in host memory on return).  The operator itself runs on the GPU.
 *
 * fftw_complex: the reference takes it from <fftw3.h> (src/collisions.h:15).
 * This header provides a binary-identical double[2] when FFTW's header has
 * not been included first.
 */
#ifndef _COLLISIONS_H
#define _COLLISIONS_H

#if !defined(FFTW3_H)
typedef double fftw_complex[2];
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/collisions.h:27 (defined at src/collisions.c:33-74).
 * modes = N velocity nodes per axis; length = L_v; vel (N) and zeta (N) are
 * borrowed, as in the reference, and must outlive the module.  Sets up the
 * GPU-side transform tables for this grid. */
void initialize_coll(int modes, double length, double *vel, double *zeta);

/* Replaces src/collisions.h:33 (src/collisions.c:81-89).  Releases the GPU
 * state of this module (including cached weight tables). */
void dealloc_coll(void);

/* Replaces src/collisions.h:38 (src/collisions.c:212-221):
 * Q = Re F^-1[ sum_c W[r][c] Ghat[c] Fhat[r - c + N/2] ], Fhat = F(f), Ghat = F(g).
 * Q is overwritten.  conv_weights are N^3 row pointers (src/weights.c:26-31,
 * :61-63); the first call with a given pointer uploads the table to HBM. */
void ComputeQ(double *f, double *g, double *Q, double **conv_weights);

/* Replaces src/collisions.h:40 (src/collisions.c:178-210): the Maxwellian-
 * preserving split f = M_f + g_f, g = M_g + g_g,
 * Q = Q(M_f, g - M_f) + Q(g_f, M_g) + Q(g_f, g - M_f)  (the reference subtracts
 * the first argument's Maxwellian in both splits, src/collisions.c:104), done
 * as ONE fused weighted-sum pass.  Q is overwritten. */
void ComputeQ_maxPreserve(double *f, double *g, double *Q, double **conv_weights);

/* Replaces src/collisions.h:42 (src/collisions.c:232-283): the phase-shifted,
 * trapezoid-weighted 3D transform; invert = 0 forward (v -> eta), 1 inverse.
 * in may alias out. */
void fft3D(fftw_complex *in, fftw_complex *out, int invert);

/* src/collisions.h:44 declares ifft3D but the reference never defines it.
 * Here it is fft3D(in, out, 1). */
void ifft3D(fftw_complex *in, fftw_complex *out);

#ifdef __cplusplus
}
#endif

#endif
