// k_weights.hip -- convolution weight tables generated on the GPU: the
// isotropic table (reference: src/weights.c:265-281 generate_conv_weights_iso
// with gHat3, :181-262, and the integrand ghat, :156-160) and the Landau
// table of the anisotropic generator with glance = 0, and its glancing-
// collision table with glance > 0 (further below).
//
//   W[r][c] = w_c1 w_c2 w_c3 * dfac * P * I_lam(mu|eta_r|, |eta_c|, |eta_c - mu eta_r|)
//   I_lam(a0,a1,a2) = int_0^L r^(lam+2) [sinc(a0 r) sinc(a2 r) - sinc(a1 r)] dr
//   P = 16 pi^2 h_eta^3 (2 pi)^-3/2 / (4 pi)
//
// The reference integrates every entry with GSL QAG (GAUSS21, 1e-8/1e-8): CPU-
// days at N=32 (1.07e9 entries).  Here lam = 0 and lam = 1 use the closed forms
// of the integral (gain - loss), one thread per entry; any other lam uses a
// composite 64-panel x 20-point Gauss-Legendre rule per entry.
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "spbte_internal.h"

namespace {

__device__ __forceinline__ double sinc(double x) { return (x != 0.0) ? sin(x) / x : 1.0; }
__device__ __forceinline__ double fcos(double x) { return (fabs(x) > 1e-12) ? (cos(x) - 1.0) / (x * x) : -0.5 + x * x / 24.0; }

// int_0^L r^2 sinc(a r) sinc(b r) dr  and  int_0^L r^2 sinc(c r) dr
__device__ double gain0(double a, double b, double L) {
  const double tiny = 1e-12;
  if (a > tiny && b > tiny) return L / (2 * a * b) * (sinc((a - b) * L) - sinc((a + b) * L));
  if (a > tiny || b > tiny) {
    const double c = a > tiny ? a : b;
    return (sin(c * L) - c * L * cos(c * L)) / (c * c * c);
  }
  return L * L * L / 3.0;
}
__device__ double loss0(double c, double L) {
  return (c > 1e-12) ? (sin(c * L) - c * L * cos(c * L)) / (c * c * c) : L * L * L / 3.0;
}
// int_0^L r^3 sinc(a r) sinc(b r) dr  and  int_0^L r^3 sinc(c r) dr
__device__ double gain1(double a, double b, double L) {
  const double tiny = 1e-12;
  if (a > tiny && b > tiny) {
    double am = (a - b) * L;
    const double ap = (a + b) * L;
    if (fabs(a - b) < 1e-10) am = 0.0;
    return (L * L * 0.5 / (a * b)) * (sinc(am) - sinc(ap) + fcos(am) - fcos(ap));
  }
  if (a > tiny || b > tiny) {
    const double c = a > tiny ? a : b, x = c * L;
    return (2 * x * sin(x) - x * x * cos(x) + 2 * cos(x) - 2) / (c * c * c * c);
  }
  return 0.25 * L * L * L * L;
}
__device__ double loss1(double c, double L) {
  if (c > 1e-12) {
    const double x = c * L;
    return (cos(x) * (2.0 - x * x) + 2.0 * x * sin(x) - 2.0) / (c * c * c * c);
  }
  return 0.25 * L * L * L * L;
}

struct GL20 {
  double x[20], w[20];
};

__device__ double integral_gl(double a0, double a1, double a2, double L, double lam, const GL20 &q) {
  const int panels = 64;
  const double h = L / panels;
  double sum = 0.0;
  for (int p = 0; p < panels; ++p) {
    const double lo = p * h;
    double ps = 0.0;
    for (int i = 0; i < 20; ++i) {
      const double r = lo + 0.5 * h * (q.x[i] + 1.0);
      ps += q.w[i] * pow(r, lam + 2.0) * (sinc(r * a0) * sinc(r * a2) - sinc(r * a1));
    }
    sum += 0.5 * h * ps;
  }
  return sum;
}

__global__ void __launch_bounds__(256) k_weights_iso(int N, const double *__restrict__ eta, double L, double lam, double P,
                                                     double dfac, double mu, GL20 q, unsigned chunks,
                                                     double *__restrict__ W) {
  const size_t n3 = (size_t)N * N * N;
  const size_t r = blockIdx.x / chunks;                 // one row per `chunks` workgroups
  const size_t c0 = (size_t)(blockIdx.x % chunks) * blockDim.x + threadIdx.x;
  const int i = (int)(r / ((size_t)N * N)), j = (int)((r / N) % N), k = (int)(r % N);
  const double z1 = eta[i], z2 = eta[j], z3 = eta[k];
  const double a0 = mu * sqrt(z1 * z1 + z2 * z2 + z3 * z3);
  for (size_t c = c0; c < n3; c += (size_t)chunks * blockDim.x) {
    const int l = (int)(c / ((size_t)N * N)), m = (int)((c / N) % N), n = (int)(c % N);
    const double k1 = eta[l], k2 = eta[m], k3 = eta[n];
    const double a1 = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
    const double d1 = k1 - mu * z1, d2 = k2 - mu * z2, d3 = k3 - mu * z3;
    const double a2 = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
    double I;
    if (lam == 0.0) I = gain0(a0, a2, L) - loss0(a1, L);
    else if (lam == 1.0) I = gain1(a0, a2, L) - loss1(a1, L);
    else I = integral_gl(a0, a1, a2, L, lam, q);
    const double wl = (l == 0 || l == N - 1) ? 0.5 : 1.0, wm = (m == 0 || m == N - 1) ? 0.5 : 1.0,
                 wn = (n == 0 || n == N - 1) ? 0.5 : 1.0;
    W[r * n3 + c] = wl * wm * wn * dfac * (P * I);
  }
}

void gauss_legendre20(GL20 &q) {
  const int n = 20;
  for (int i = 0; i < n / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    q.x[i] = -z;
    q.x[n - 1 - i] = z;
    q.w[i] = q.w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

// ---- Landau table (glance = 0): gHat3L of exec/MPIcollisionRoutines.c:205-248
// (src/aniso_weights.c:352-398), one thread per entry:
//   W[r][c] = 2 pi int_0^L rho^(lam+2) int_0^pi g(theta) dtheta drho,
//   g = sin(t) J0(rho a0 sin t) (-rho^2 a1 sin^2 t cos(rho a2 cos t) + 4 rho a3 sin(rho a2 cos t) cos t),
//   a0 = |xi_perp|, a1 = |zeta|^2, a2 = xi.zeta/|zeta|, a3 = |zeta| (zeta = row mode, xi = column mode),
// both integrals by the 64-point Gauss-Legendre rule summed in symmetric node
// pairs, as gsl_integration_glfixed and gauss_legendre do.  The theta nodes of
// a pair, pi/2 + y and pi/2 - y with y = (pi/2) x_k, share sin(theta) = cos(y)
// and have opposite cos(theta) = -+sin(y); g is even in cos(theta) (cos(rho a2
// cos t) is even, sin(rho a2 cos t) cos t is even), so the two terms of a pair
// are equal and each pair costs one J0 and one sincos.  The node abscissae,
// cos(y), sin(y) and rho^(lam+2) are the same for every entry and are
// evaluated once on the host with the C library.
struct LandauRule {
  double w[32];                    // weights of the positive nodes, ascending nodes
  double st[32], ct[32];           // sin(theta), |cos(theta)| of theta = pi/2 -+ (pi/2) x_k
  double r_p[32], r_m[32];         // rho = B + A x_i, B - A x_i on [0, L]
  double rl_p[32], rl_m[32];       // rho^(lam+2)
  double At, Ar;                   // half-lengths of the two intervals
};

__device__ __forceinline__ double landau_g(double r, double a0, double a1, double a2, double a3, double st, double ct) {
  double sv, cv;
  sincos(r * a2 * ct, &sv, &cv);
  return st * j0(r * a0 * st) * (-r * r * a1 * st * st * cv + 4 * r * a3 * sv * ct);
}

__global__ void __launch_bounds__(256) k_weights_landau(int N, const double *__restrict__ eta, long r0, long nrows,
                                                        LandauRule q, double *__restrict__ W) {
  const long n3 = (long)N * N * N;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nrows * n3) return;
  const long r = r0 + e / n3, c = e % n3;
  const double z1 = eta[r / ((long)N * N)], z2 = eta[(r / N) % N], z3 = eta[r % N];
  const double x1 = eta[c / ((long)N * N)], x2 = eta[(c / N) % N], x3 = eta[c % N];
  const double zetalen2 = z3 * z3 + z2 * z2 + z1 * z1;
  const double xilen2 = x1 * x1 + x2 * x2 + x3 * x3;
  const double xizeta = x1 * z1 + x2 * z2 + x3 * z3;
  const double zetalen = sqrt(zetalen2);
  const double perp2 = xilen2 - xizeta * xizeta / zetalen2;
  const double a0 = (perp2 < 0 || zetalen2 == 0) ? 0.0 : sqrt(perp2);
  const double a2 = zetalen != 0 ? xizeta / zetalen : 0.0;
  double sr = 0.0;
  for (int i = 0; i < 32; ++i) {
    double f[2];
    for (int h = 0; h < 2; ++h) {
      const double rho = h ? q.r_m[i] : q.r_p[i];
      double st = 0.0;
      for (int k = 0; k < 32; ++k)
        st += q.w[k] * (2.0 * landau_g(rho, a0, zetalen2, a2, zetalen, q.st[k], q.ct[k]));
      f[h] = (h ? q.rl_m[i] : q.rl_p[i]) * (q.At * st);
    }
    sr += q.w[i] * (f[0] + f[1]);
  }
  W[(r - r0) * n3 + c] = 2.0 * M_PI * (q.Ar * sr);
}

// n-point Gauss-Legendre rule: the n/2 positive nodes ascending and their
// weights, Newton's iteration on P_n in long double.
void gauss_legendre_half(int n, double *x, double *w) {
  for (int i = 0; i < n / 2; i++) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), pp = 0.0L;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= n; j++) {
        const long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0L);
      const long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    x[n / 2 - 1 - i] = (double)z;
    w[n / 2 - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
  }
}

}  // namespace

extern "C" int spbte_launch_weights_landau(const spbte_grid_dev *g, const double *eta, double lam, double *W,
                                           hipStream_t s) {
  const int N = g->N;
  const long n3 = (long)N * N * N;
  LandauRule q;
  double x[32];
  gauss_legendre_half(64, x, q.w);
  q.At = 0.5 * (M_PI - 0.0);
  q.Ar = 0.5 * (g->Lv - 0.0);
  const double Br = 0.5 * (g->Lv + 0.0);
  for (int k = 0; k < 32; k++) {
    const double At_x = q.At * x[k], Ar_x = q.Ar * x[k];
    q.st[k] = cos(At_x);
    q.ct[k] = sin(At_x);
    q.r_p[k] = Br + Ar_x;
    q.r_m[k] = Br - Ar_x;
    q.rl_p[k] = pow(q.r_p[k], lam + 2);
    q.rl_m[k] = pow(q.r_m[k], lam + 2);
  }
  // rows per launch: about 8M entries (a fraction of a second each at N = 32)
  const long rows = n3 >= (8L << 20) ? 1 : (8L << 20) / n3;
  for (long r0 = 0; r0 < n3; r0 += rows) {
    const long nr = r0 + rows <= n3 ? rows : n3 - r0;
    const unsigned blocks = (unsigned)((nr * n3 + 255) / 256);
    hipLaunchKernelGGL(k_weights_landau, dim3(blocks), dim3(256), 0, s, N, eta, r0, nr, q, W + r0 * n3);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
  }
  return 0;
}

// ---- Glancing-collision table (glance g > 0): gHat3 of
// exec/MPIcollisionRoutines.c:149-203, the table exec/MPIWeightGenerator.c
// writes.  The reference nests three adaptive quadratures per entry,
//   W = 4 pi^2 int_0^L r^(lam+2) int_0^pi sin(phi) J0(r sin(phi) xi_perp) (R1 + R2) dphi dr,
//   R1 = int_sqrt(g)^pi b(theta) [cos(a7 (1 - cos theta) - a6) J0(a8 sin theta) - cos a6] dtheta,
//   R2 = C(r, phi) (2/pi) Lam,  Lam = log g / log sin(g/2),  b(theta) = (8/pi) g / theta^4,
// (a6 = r cos(phi) a1, a7 = r cos(phi) |zeta|/2, a8 = r sin(phi) |zeta|/2,
// a1 = xi.zeta/|zeta|; C as in ghat_phi :111-114).  The same integral is
// evaluated here in a reduced form.  With J0(x s) J0(y s) = (1/pi) int_0^pi
// J0(s sqrt(x^2 + y^2 - 2 x y cos psi)) dpsi (Neumann's addition theorem) and
// int_0^pi sin(phi) cos(X cos phi) J0(Z sin phi) dphi = 2 sinc(sqrt(X^2 + Z^2))
// (the sphere average of a plane wave), the phi integral of R1 is a psi
// average of sincs, and the r integral of a sinc is one function D of one
// variable:
//   R1 part = 8 pi^2 int_sqrt(g)^pi b(theta) [ (1/pi) int_0^pi D(kappa) dpsi - D(|xi|) ] dtheta,
//   kappa^2 = |xi|^2 + (|zeta|^2 - 2 xi.zeta) sin^2(theta/2) - xi_perp |zeta| sin(theta) cos(psi),
//   D(k) = int_0^L r^(lam+2) (sinc(k r) - 1) dr.
// The R2 part is Lam/4 times the exact Landau integral, whose phi integral is
// a combination of derivatives of the same sinc (S = sinc):
//   R2 part = (pi Lam / 2) [2 |zeta|^2 (c_perp^2 E1(q) + (2 - c_perp^2) E2(q)) - 8 |zeta| c_par E3(q)],
//   q = |xi|, c_perp = xi_perp / q, c_par = a1 / q,
//   E1 = int r^(lam+4) S''(q r) dr,  E2 = int r^(lam+4) S'(q r)/(q r) dr,  E3 = int r^(lam+3) S'(q r) dr.
// D, E1, E2, E3 are tabulated once per table on the host (values from
// composite Gauss-Legendre sums in r) on panels of width 3/L: D, evaluated at
// every node, as degree-15 polynomials in the local variable t in [-1, 1]
// (Chebyshev interpolants converted to monomials, evaluated by Horner's rule;
// the coefficients stay of the size of the function, and the interpolation
// error is about 1e-15 of it), held in LDS; E1..E3, evaluated once per entry,
// as degree-15 Chebyshev series, read from global memory.  Narrower panels of
// lower degree would read fewer bytes per node, but the lanes of a wave then
// spread over more panels than the LDS banks separate.  An entry then costs a 2-D sum of polynomial evaluations instead of
// a 3-D sum of Bessel functions.  theta: 16-point panels in log(theta) on
// [sqrt(g), 1] and in theta on [max(1, sqrt(g)), pi]; psi: the trapezoid rule,
// exact to rounding for the smooth even periodic integrand.
namespace {

struct GlanceDims {
  int nl;                  // theta nodes of the log(theta) part (16 per panel; 0 if glance >= 1)
  int nvt, nvp;            // theta rule variants (2 + v uniform panels), psi variants (8 k intervals)
  double tL, pL;           // per-row picks: v = ceil(|zeta| tL), k = ceil((12 + |zeta| pL) / 8)
  int dofs;                // offset of the D table in LDS
  int nD, nE;              // panels of D and of E1..E3
  double invhD, invhE;     // 1 / panel widths
  double lam4;             // pi Lam / 2
  int nlds;                // doubles of the rule buffer copied to LDS (nodes and D)
};
constexpr int kChebN = 16, kChebStride = 17;  // Chebyshev coefficients per panel, padded stride
// monomial coefficients per panel; an odd stride in doubles puts 32 consecutive
// panels on distinct LDS banks (the lanes of a wave read panels close together)
constexpr int kPolyN = 16, kPolyStride = 17;

__device__ __forceinline__ double poly_eval(const double *tab, int P, double invh, double x) {
  const double u = x * invh;
  int p = (int)u;
  p = p < 0 ? 0 : (p >= P ? P - 1 : p);
  const double t = 2.0 * (u - p) - 1.0;
  const double *c = tab + p * kPolyStride;
  double r = c[kPolyN - 1];
#pragma unroll
  for (int k = kPolyN - 2; k >= 0; --k) r = fma(r, t, c[k]);
  return r;
}

__device__ __forceinline__ double cheb_eval(const double *tab, int P, double invh, double x) {
  const double u = x * invh;
  int p = (int)u;
  p = p < 0 ? 0 : (p >= P ? P - 1 : p);
  const double t = 2.0 * (u - p) - 1.0, t2 = 2.0 * t;
  const double *c = tab + p * kChebStride;
  double b1 = 0.0, b2 = 0.0;
#pragma unroll
  for (int k = kChebN - 1; k >= 1; --k) {
    const double b0 = fma(t2, b1, c[k] - b2);
    b2 = b1;
    b1 = b0;
  }
  return fma(t, b1, c[0] - b2);
}

// Node counts per row.  Over the theta and psi sums kappa stays within
// |xi| -+ |zeta| (|k - xi| <= |zeta| sin(theta/2)), so the number of
// oscillations of D(kappa) there is bounded by L |zeta| / pi: a row takes the
// theta rule with 2 + ceil(L |zeta| / 60) uniform panels and the psi rule with
// 8 ceil((12 + L |zeta| / 5) / 8) intervals.  Every variant is built on the
// host; variant v of the theta rule holds nl + 16 (2 + v) nodes, variant k of
// the psi rule 8 k + 1.
__host__ __device__ inline int glance_tofs(int nl, int v) { return 3 * (v * nl + 16 * (2 * v + v * (v - 1) / 2)); }
__host__ __device__ inline int glance_pofs(int k) { return 8 * k * (k - 1) + 2 * (k - 1); }

// rule buffer: theta variants v = 0 .. nvt-1, each s2[nt] = sin^2(theta/2),
// sth[nt] = sin(theta), wt[nt] = theta weight x b(theta) x 8 pi^2; psi variants
// k = 1 .. nvp, each cps[8k+1] = cos(psi), wps[8k+1] = psi weight / pi; the D
// table (nD panels) at dofs: the first nlds doubles, copied to LDS; then the
// E1, E2, E3 tables (nE panels each), read in place.
constexpr int kGlanceThreads = 256;
__global__ void __launch_bounds__(kGlanceThreads) k_weights_glance(int N, const double *__restrict__ eta, long r0, long nrows,
                                                        const double *__restrict__ rule, GlanceDims d,
                                                        double *__restrict__ W) {
  extern __shared__ double sm[];
  for (int i = threadIdx.x; i < d.nlds; i += kGlanceThreads) sm[i] = rule[i];
  __syncthreads();
  const double *tD = sm + d.dofs;
  const double *tE1 = rule + d.nlds, *tE2 = tE1 + d.nE * kChebStride, *tE3 = tE2 + d.nE * kChebStride;
  const long n3 = (long)N * N * N;
  const long e = (long)blockIdx.x * kGlanceThreads + threadIdx.x;
  if (e >= nrows * n3) return;
  const long r = r0 + e / n3, c = e % n3;
  const double z1 = eta[r / ((long)N * N)], z2 = eta[(r / N) % N], z3 = eta[r % N];
  const double x1 = eta[c / ((long)N * N)], x2 = eta[(c / N) % N], x3 = eta[c % N];
  // geometry as gHat3 (:165-183)
  const double zetalen2 = z1 * z1 + z2 * z2 + z3 * z3;
  const double xilen2 = x1 * x1 + x2 * x2 + x3 * x3;
  const double xizeta = x1 * z1 + x2 * z2 + x3 * z3;
  const double zetalen = sqrt(zetalen2);
  const double perp2 = xilen2 - xizeta * xizeta / zetalen2;
  const double xp = (perp2 < 0 || zetalen2 == 0) ? 0.0 : sqrt(perp2);
  const double a1 = zetalen != 0 ? xizeta / zetalen : 0.0;
  const double q2 = a1 * a1 + xp * xp, q = sqrt(q2);
  const double P = zetalen2 - 2.0 * zetalen * a1, Q = xp * zetalen;
  int v = (int)ceil(zetalen * d.tL), k = (int)ceil((12.0 + zetalen * d.pL) * 0.125);
  v = v < 0 ? 0 : (v >= d.nvt ? d.nvt - 1 : v);
  k = k < 1 ? 1 : (k > d.nvp ? d.nvp : k);
  const int nt = d.nl + 16 * (2 + v), np = 8 * k + 1;
  const double *s2 = sm + glance_tofs(d.nl, v), *sth = s2 + nt, *wt = sth + nt;
  const double *cps = sm + glance_tofs(d.nl, d.nvt) + glance_pofs(k), *wps = cps + np;
  const double D0 = poly_eval(tD, d.nD, d.invhD, q);
  double t1 = 0.0;
  for (int i = 0; i < nt; ++i) {
    const double A = fma(P, s2[i], q2), B = Q * sth[i];
    double acc = 0.0;
    for (int j = 0; j < np; ++j) {
      const double k2 = fma(-B, cps[j], A);
      acc = fma(wps[j], poly_eval(tD, d.nD, d.invhD, sqrt(k2 > 0.0 ? k2 : 0.0)), acc);
    }
    t1 = fma(wt[i], acc - D0, t1);
  }
  double cperp2 = 0.0, cpar = 0.0;
  if (q > 0.0) {
    cperp2 = xp * xp / q2;
    cpar = a1 / q;
  }
  const double E1 = cheb_eval(tE1, d.nE, d.invhE, q), E2 = cheb_eval(tE2, d.nE, d.invhE, q),
               E3 = cheb_eval(tE3, d.nE, d.invhE, q);
  const double t2 = d.lam4 * (2.0 * zetalen2 * (cperp2 * E1 + (2.0 - cperp2) * E2) - 8.0 * zetalen * cpar * E3);
  W[(r - r0) * n3 + c] = t1 + t2;
}

// sinc and its derivatives; power series below |x| = 1 (no cancellation)
enum SincKind { kSm1, kS1, kS1x, kS2 };  // S - 1, S', S'/x, S''
double sinc_fn(SincKind k, double x) {
  if (fabs(x) < 1.0) {
    // S = sum_n (-1)^n x^(2n) / (2n+1)!
    double s = 0.0, fact = 1.0, xp = 1.0;  // fact = (2n+1)!, xp = x^(2n-2)
    for (int n = 1; n <= 12; ++n) {
      fact *= (2.0 * n) * (2.0 * n + 1.0);
      const double sg = (n & 1) ? -1.0 : 1.0;
      double term;
      switch (k) {
        case kSm1: term = xp * x * x; break;
        case kS1: term = 2.0 * n * xp * x; break;
        case kS1x: term = 2.0 * n * xp; break;
        default: term = 2.0 * n * (2.0 * n - 1.0) * xp; break;
      }
      s += sg * term / fact;
      xp *= x * x;
    }
    return s;
  }
  const double sx = sin(x), cx = cos(x);
  switch (k) {
    case kSm1: return sx / x - 1.0;
    case kS1: return (x * cx - sx) / (x * x);
    case kS1x: return (x * cx - sx) / (x * x * x);
    default: return ((2.0 - x * x) * sx - 2.0 * x * cx) / (x * x * x);
  }
}

// int_0^L r^pw f(k r) dr by 16-point panels of width <= 2/kmax, the first
// panel split geometrically towards r = 0 (r^pw need not be smooth there)
struct RadialRule {
  std::vector<double> r, w;
  RadialRule(double L, double kmax, double pw) {
    double gx[8], gw[8];
    gauss_legendre_half(16, gx, gw);
    const int P = (int)ceil(kmax * L / 2.0) + 2;
    const double h = L / P;
    auto panel = [&](double a, double b) {
      const double c = 0.5 * (a + b), hh = 0.5 * (b - a);
      for (int i = 0; i < 8; ++i)
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          const double x = c + sgn * hh * gx[i];
          r.push_back(x);
          w.push_back(hh * gw[i] * pow(x, pw));
        }
    };
    double a = h / 1024.0;
    panel(0.0, a);
    for (; a < h; a *= 4.0) panel(a, fmin(4.0 * a, h));
    for (int p = 1; p < P; ++p) panel(p * h, (p + 1) * h);
  }
  double sum(SincKind k, double kappa) const {
    double s = 0.0;
    for (size_t i = 0; i < r.size(); ++i) s += w[i] * sinc_fn(k, kappa * r[i]);
    return s;
  }
};

// piecewise Chebyshev coefficients of f on [0, P h], appended to out
template <class F>
void cheb_table(int P, double h, F f, std::vector<double> &out) {
  double fv[kChebN];
  for (int p = 0; p < P; ++p) {
    for (int j = 0; j < kChebN; ++j) {
      const double t = cos(M_PI * (j + 0.5) / kChebN);
      fv[j] = f(h * (p + 0.5 * (t + 1.0)));
    }
    for (int k = 0; k < kChebN; ++k) {
      double c = 0.0;
      for (int j = 0; j < kChebN; ++j) c += fv[j] * cos(M_PI * k * (j + 0.5) / kChebN);
      out.push_back((k == 0 ? 1.0 : 2.0) * c / kChebN);
    }
    out.push_back(0.0);  // pad to kChebStride
  }
}

// piecewise degree-(kPolyN-1) polynomials of f in t in [-1, 1] on [0, P h]:
// Chebyshev interpolant at kPolyN nodes, converted to monomials in long double
template <class F>
void poly_table(int P, double h, F f, std::vector<double> &out) {
  long double Tk[kPolyN][kPolyN] = {};  // monomial coefficients of T_k
  Tk[0][0] = 1.0L;
  Tk[1][1] = 1.0L;
  for (int k = 2; k < kPolyN; ++k)
    for (int i = 0; i < kPolyN; ++i) Tk[k][i] = (i > 0 ? 2.0L * Tk[k - 1][i - 1] : 0.0L) - Tk[k - 2][i];
  double fv[kPolyN];
  for (int p = 0; p < P; ++p) {
    for (int j = 0; j < kPolyN; ++j) fv[j] = f(h * (p + 0.5 * (cos(M_PI * (j + 0.5) / kPolyN) + 1.0)));
    long double a[kPolyN] = {};
    for (int k = 0; k < kPolyN; ++k) {
      long double c = 0.0L;
      for (int j = 0; j < kPolyN; ++j) c += (long double)fv[j] * cosl(3.14159265358979323846264338327950288L * k * (j + 0.5L) / kPolyN);
      c *= (k == 0 ? 1.0L : 2.0L) / kPolyN;
      for (int i = 0; i < kPolyN; ++i) a[i] += c * Tk[k][i];
    }
    for (int i = 0; i < kPolyN; ++i) out.push_back((double)a[i]);
    for (int i = kPolyN; i < kPolyStride; ++i) out.push_back(0.0);
  }
}

}  // namespace

// The glance table: 0 < g < pi (log sin(g/2) < 0), checked by the caller.
extern "C" int spbte_launch_weights_glance(const spbte_grid_dev *g, const double *eta_host, const double *eta,
                                           double lam, double glance, double *W, hipStream_t s) {
  const int N = g->N;
  const long n3 = (long)N * N * N;
  const double L = g->Lv;
  double emax = 0.0;
  for (int i = 0; i < N; ++i) emax = fmax(emax, fabs(eta_host[i]));
  const double zmax = sqrt(3.0) * emax, kmax = 2.0 * zmax;  // |xi|, |zeta| <= zmax; kappa <= |xi| + |zeta|
  std::vector<double> buf;
  GlanceDims d;
  // theta and psi rules, every per-row variant
  double gx[8], gw[8];
  gauss_legendre_half(16, gx, gw);
  auto panels = [&](std::vector<double> &th, std::vector<double> &tw, double a, double b, int P, bool logv) {
    const double h = (b - a) / P;
    for (int p = 0; p < P; ++p)
      for (int i = 0; i < 8; ++i)
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          const double u = a + h * (p + 0.5) + sgn * 0.5 * h * gx[i], w = 0.5 * h * gw[i];
          th.push_back(logv ? exp(u) : u);
          tw.push_back(logv ? w * exp(u) : w);
        }
  };
  const double t0 = sqrt(glance);
  const int P1 = t0 < 1.0 ? 1 + (int)ceil(-log(t0) / 1.2) : 0;
  d.nl = 16 * P1;
  d.tL = L / 60.0;
  d.pL = L / 5.0;
  d.nvt = (int)ceil(zmax * d.tL) + 1;
  d.nvp = (int)ceil((12.0 + zmax * d.pL) * 0.125);
  for (int v = 0; v < d.nvt; ++v) {
    std::vector<double> th, tw;
    if (P1) panels(th, tw, log(t0), 0.0, P1, true);
    panels(th, tw, fmax(t0, 1.0), M_PI, 2 + v, false);
    for (size_t i = 0; i < th.size(); ++i) buf.push_back(sin(0.5 * th[i]) * sin(0.5 * th[i]));
    for (size_t i = 0; i < th.size(); ++i) buf.push_back(sin(th[i]));
    for (size_t i = 0; i < th.size(); ++i)
      buf.push_back(8.0 * M_PI * M_PI * tw[i] * (8.0 / M_PI) * glance / (th[i] * th[i] * th[i] * th[i]));
  }
  for (int k = 1; k <= d.nvp; ++k) {
    const int M = 8 * k;
    for (int j = 0; j <= M; ++j) buf.push_back(cos(M_PI * j / M));
    for (int j = 0; j <= M; ++j) buf.push_back((j == 0 || j == M ? 0.5 : 1.0) / M);
  }
  if (buf.size() != (size_t)(glance_tofs(d.nl, d.nvt) + glance_pofs(d.nvp + 1))) return (int)hipErrorInvalidValue;
  d.dofs = (int)buf.size();
  // radial tables: D in panels of width 3/L, widened if the LDS part would
  // pass 64 KiB (very large N only; the error grows as the 16th power of the width)
  double hp = 3.0 / L;
  while ((buf.size() + (size_t)kPolyStride * ((int)ceil(kmax / hp) + 1)) * sizeof(double) > 64 * 1024) hp *= 1.1;
  d.nD = (int)ceil(kmax / hp) + 1;
  d.invhD = 1.0 / hp;
  const double hc = 3.0 / L;
  d.nE = (int)ceil(zmax / hc) + 1;
  d.invhE = 1.0 / hc;
  {
    const RadialRule r2(L, d.nD * hp, lam + 2.0), r3(L, d.nE * hc, lam + 3.0), r4(L, d.nE * hc, lam + 4.0);
    poly_table(d.nD, hp, [&](double k) { return r2.sum(kSm1, k); }, buf);
    d.nlds = (int)buf.size();
    cheb_table(d.nE, hc, [&](double k) { return r4.sum(kS2, k); }, buf);
    cheb_table(d.nE, hc, [&](double k) { return r4.sum(kS1x, k); }, buf);
    cheb_table(d.nE, hc, [&](double k) { return r3.sum(kS1, k); }, buf);
  }
  d.lam4 = 0.5 * M_PI * log(glance) / log(sin(0.5 * glance));
  const size_t lds = sizeof(double) * d.nlds, bytes = sizeof(double) * buf.size();
  double *drule = NULL;
  hipError_t e = hipMalloc((void **)&drule, bytes);
  if (e != hipSuccess) return (int)e;
  e = hipMemcpy(drule, buf.data(), bytes, hipMemcpyHostToDevice);
  // rows per launch: about 4M entries
  const long rows = n3 >= (4L << 20) ? 1 : (4L << 20) / n3;
  for (long r0 = 0; e == hipSuccess && r0 < n3; r0 += rows) {
    const long nr = r0 + rows <= n3 ? rows : n3 - r0;
    const unsigned blocks = (unsigned)((nr * n3 + kGlanceThreads - 1) / kGlanceThreads);
    hipLaunchKernelGGL(k_weights_glance, dim3(blocks), dim3(kGlanceThreads), lds, s, N, eta, r0, nr, drule, d,
                       W + r0 * n3);
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(s), e3 = hipFree(drule);
  return e != hipSuccess ? (int)e : (e2 != hipSuccess ? (int)e2 : (int)e3);
}

extern "C" int spbte_launch_weights_iso(const spbte_grid_dev *g, const double *eta, double lam, double dfac, double mu,
                                        double *W, hipStream_t s) {
  const int N = g->N;
  const size_t n3 = (size_t)N * N * N;
  const double P = 16.0 * M_PI * M_PI * g->heta * g->heta * g->heta / pow(2.0 * M_PI, 1.5) / (4.0 * M_PI);
  GL20 q;
  gauss_legendre20(q);
  const unsigned chunks = (unsigned)((n3 + 255) / 256 < 64 ? (n3 + 255) / 256 : 64);
  hipLaunchKernelGGL(k_weights_iso, dim3((unsigned)(n3 * chunks)), dim3(256), 0, s, N, eta, g->Lv, lam, P, dfac, mu, q, chunks,
                     W);
  return (int)hipGetLastError();
}
