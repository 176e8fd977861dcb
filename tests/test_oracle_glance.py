"""Pins of the oracle's glancing-collision entry (gHat3 of
exec/MPIcollisionRoutines.c:33-203, glance > 0), which restates the
reference's triple integral over (r, phi, theta) with fixed composite rules.
No reference fixture holds such a table, so the oracle is pinned by what the
mathematics fixes, each check computed from a different formula than the
oracle's:

* a zero row mode zeta gives exactly 0 (every integrand vanishes);
* xi = 0: the phi integral is the sphere average of a plane wave,
  int_0^pi sin(phi) cos(X cos phi) J0(Z sin phi) dphi = 2 sinc(sqrt(X^2 + Z^2)),
  which leaves a 2-D integral over (theta, r), done here by scipy's adaptive
  quadrature, plus the Taylor term in closed form;
* xi parallel to zeta: the same identity with the xi.zeta phase, and the
  Taylor term by 2-D scipy quadrature over (r, cos phi);
* general xi: Neumann's addition theorem for J0 turns the phi integral into
  a psi average of sincs (an independent 2-D form over (theta, psi), numpy
  Gauss-Legendre and trapezoid sums here), and the Taylor term equals
  Lam/4 times the Landau entry gHat3L, Lam = log g / log sin(g/2), tying it to
  the oracle's Landau entry that test_oracle_math pins.
Tolerance 1e-9 relative: the checks are accurate to about 1e-11, the oracle's
panels to about 1e-12 on these grids."""
import math

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss
from scipy import integrate, special


def _b(th, g):
    return (8.0 / math.pi) * g / th ** 4  # ghat_theta, exec/MPIcollisionRoutines.c:51


def _lam_fac(g):
    return math.log(g) / math.log(math.sin(0.5 * g))  # :119


def _D(kappa, Lv, lam):
    """int_0^L r^(lam+2) (sinc(kappa r) - 1) dr by adaptive quadrature."""
    def f(r):
        x = kappa * r
        s1 = -x * x / 6 + x ** 4 / 120 - x ** 6 / 5040 + x ** 8 / 362880 if x < 1e-2 else math.sin(x) / x - 1
        return r ** (lam + 2) * s1
    return integrate.quad(f, 0, Lv, epsabs=0, epsrel=1e-13, limit=400)[0]


def _grid(N, Lv):
    hv = 2 * Lv / (N - 1)
    he = 2 * math.pi / (N * hv)
    return -0.5 * N * he + he * np.arange(N)


def test_zero_row_mode_gives_zero(orc):
    eta = _grid(6, 3.0)
    assert orc.glance_entry((0, 0, 0), (eta[1], eta[4], eta[0]), 3.0, 1.0, 0.01) == 0.0


@pytest.mark.parametrize("lam,g", [(1.0, 0.01), (-3.0, 0.2), (0.0, 1e-3)])
def test_xi_zero_reduces_to_sinc_average(orc, lam, g):
    """xi = 0: W = 8 pi^2 int_sqrt(g)^pi b(theta) D(|zeta| sin(theta/2)) dtheta
    - (2 pi / 3) Lam |zeta|^2 L^(lam+5) / (lam+5)."""
    Lv = 3.0
    eta = _grid(6, Lv)
    zeta = np.array([eta[0], eta[4], eta[2]])
    zl = float(np.linalg.norm(zeta))
    t1 = 8 * math.pi ** 2 * integrate.quad(lambda th: _b(th, g) * _D(zl * math.sin(0.5 * th), Lv, lam),
                                           math.sqrt(g), math.pi, epsabs=0, epsrel=1e-12, limit=400)[0]
    t2 = -(2 * math.pi / 3) * _lam_fac(g) * zl * zl * Lv ** (lam + 5) / (lam + 5)
    got = orc.glance_entry(zeta, (0, 0, 0), Lv, lam, g, (6, 6, 5, 5))
    assert abs(got - (t1 + t2)) < 1e-9 * abs(t1 + t2), (got, t1, t2)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_xi_parallel_to_zeta_reduces_to_sinc_average(orc, sign):
    """xi = s zeta/|zeta|: the R1 part is 8 pi^2 int b(theta) [D(kappa) - D(|s|)]
    dtheta with kappa = |(|zeta| sin^2(theta/2) - s, |zeta| sin(theta)/2)|; the
    Taylor part 4 pi^2 (2/pi) Lam int r^(lam+2) int_-1^1 C du dr with
    C = -(1/16) r^2 |zeta|^2 (1-u^2) cos(r s u) + (1/4) r |zeta| u sin(r s u)."""
    Lv, lam, g = 3.0, 1.0, 0.01
    eta = _grid(6, Lv)
    zeta = np.array([eta[4], eta[4], eta[4]])
    xi = sign * np.array([eta[5], eta[5], eta[5]])
    zl = float(np.linalg.norm(zeta))
    s = float(xi @ zeta) / zl

    def kappa(th):
        s2 = math.sin(0.5 * th) ** 2
        return math.hypot(zl * s2 - s, 0.5 * zl * math.sin(th))
    Ds = _D(abs(s), Lv, lam)
    t1 = 8 * math.pi ** 2 * integrate.quad(lambda th: _b(th, g) * (_D(kappa(th), Lv, lam) - Ds),
                                           math.sqrt(g), math.pi, epsabs=0, epsrel=1e-12, limit=400)[0]

    def c(u, r):
        return r ** (lam + 2) * (-(1 / 16) * r * r * zl * zl * (1 - u * u) * math.cos(r * s * u)
                                 + 0.25 * r * zl * u * math.sin(r * s * u))
    t2 = 4 * math.pi ** 2 * (2 / math.pi) * _lam_fac(g) * integrate.dblquad(
        c, 0, Lv, -1, 1, epsabs=1e-13, epsrel=1e-12)[0]
    got = orc.glance_entry(zeta, xi, Lv, lam, g, (6, 6, 5, 5))
    assert abs(got - (t1 + t2)) < 1e-9 * abs(t1 + t2), (got, t1, t2)


def _composite(a, b, P, n=16):
    x, w = leggauss(n)
    h = (b - a) / P
    c = a + h * (np.arange(P)[:, None] + 0.5)
    return (c + 0.5 * h * x).ravel(), np.tile(0.5 * h * w, P)


def _neumann_r1(zeta, xi, Lv, lam, g):
    """8 pi^2 int b(theta) [(1/pi) int_0^pi D(kappa) dpsi - D(|xi|)] dtheta,
    kappa^2 = |xi|^2 + (|zeta|^2 - 2 xi.zeta) sin^2(theta/2) - xi_perp |zeta| sin(theta) cos(psi)."""
    zl = float(np.linalg.norm(zeta))
    a1 = float(xi @ zeta) / zl
    xp = math.sqrt(max(float(xi @ xi) - a1 * a1, 0.0))
    r, wr = _composite(0, Lv, 24)

    def D(k):
        x = np.multiply.outer(k, r)
        return ((np.sinc(x / np.pi) - 1) * wr * r ** (lam + 2)).sum(-1)
    u, wu = _composite(0.5 * math.log(g), 0.0, 6)
    t2, w2 = _composite(1.0, math.pi, 6)
    th = np.concatenate([np.exp(u), t2])
    wt = np.concatenate([wu * np.exp(u), w2]) * _b(th, g)
    M = 96
    psi = math.pi * np.arange(M + 1) / M
    wp = np.full(M + 1, 1.0 / M)
    wp[0] = wp[-1] = 0.5 / M
    k2 = (a1 * a1 + xp * xp + (zl * zl - 2 * zl * a1) * np.sin(0.5 * th)[:, None] ** 2
          - xp * zl * np.sin(th)[:, None] * np.cos(psi)[None, :])
    avg = (D(np.sqrt(np.maximum(k2, 0))) * wp).sum(-1)
    return 8 * math.pi ** 2 * (wt * (avg - D(np.array([math.hypot(a1, xp)]))[0])).sum()


@pytest.mark.parametrize("lam,g", [(1.0, 0.01), (0.0, 0.1), (-3.0, 1e-3)])
def test_general_entries_against_neumann_form_and_landau(orc, lam, g):
    """Generic (zeta, xi) with xi_perp != 0: the oracle's triple integral equals
    the psi-average form of the R1 part plus Lam/4 times gHat3L."""
    Lv = 3.0
    eta = _grid(6, Lv)
    rng = np.random.default_rng(11)
    for _ in range(4):
        zeta = eta[rng.integers(0, 6, 3)]
        xi = eta[rng.integers(0, 6, 3)]
        if np.linalg.norm(zeta) == 0 or np.linalg.norm(np.cross(zeta, xi)) < 0.1:
            continue
        want = _neumann_r1(zeta, xi, Lv, lam, g) + 0.25 * _lam_fac(g) * orc.landau_entry(zeta, xi, Lv, lam)
        got = orc.glance_entry(zeta, xi, Lv, lam, g, (6, 6, 5, 5))
        assert abs(got - want) < 1e-9 * abs(want), (zeta, xi, got, want)


def test_entries_binding_matches_single_entries(orc):
    """orc.weights_glance_entries indexes rows and columns as k + N(j + N i)."""
    N, Lv, lam, g = 4, 2.0, 1.0, 0.05
    eta = _grid(N, Lv)
    rows, cols = np.array([0, 21, 63, 42]), np.array([63, 5, 0, 42])
    got = orc.weights_glance_entries(N, eta, Lv, lam, g, rows, cols, (3, 3, 3, 3))
    for e, (r, c) in enumerate(zip(rows, cols)):
        zeta = (eta[r // 16], eta[(r // 4) % 4], eta[r % 4])
        xi = (eta[c // 16], eta[(c // 4) % 4], eta[c % 4])
        assert got[e] == orc.glance_entry(zeta, xi, Lv, lam, g, (3, 3, 3, 3))
