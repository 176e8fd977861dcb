"""Independent checks of the CPU oracle: each pins one function by a route other
than its own formula (a library routine, a closed form, an invariant)."""
import numpy as np
import pytest
from scipy import integrate

from conftest import rel_l2


@pytest.mark.parametrize("N,Lv,grid", [(4, 3.0, "0d"), (8, 5.0, "0d"), (6, 4.0, "1d"), (8, 9.0, "1d")])
@pytest.mark.parametrize("invert", [0, 1])
def test_fft3d_matches_library_fft(orc, N, Lv, grid, invert):
    """src/collisions.c:232-283 is pre-twiddle * FFTW * post-twiddle; rebuild it with
    numpy's FFT (pocketfft) and compare with the oracle's direct separable DFT."""
    v, eta = (orc.grid_0d if grid == "0d" else orc.grid_1d)(N, Lv)
    rng = np.random.default_rng(N * 10 + invert)
    x = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    out = orc.fft3d(N, v, eta, x, invert)
    dv, deta = v[1] - v[0], eta[1] - eta[0]
    if not invert:
        delta, Ls, Le, a, s = dv, -eta[0], Lv, eta, 1.0
    else:
        delta, Ls, Le, a, s = deta, Lv, -eta[0], v, -1.0
    w = np.ones(N); w[0] = w[-1] = 0.5
    idx = np.arange(N)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    pre = (2 * np.pi) ** -1.5 * delta ** 3 * w[I] * w[J] * w[K] * np.exp(1j * s * (I + J + K) * Ls * delta)
    y = pre * x
    Y = np.fft.ifftn(y) * N ** 3 if invert else np.fft.fftn(y)
    post = np.exp(1j * s * Le * (a[I] + a[J] + a[K]))
    assert rel_l2(out, post * Y) < 1e-14


def test_fft3d_gaussian_closed_form(orc):
    """Forward transform of a centred Gaussian approaches its continuous Fourier
    transform (2 pi)^-3/2 int f e^{-i eta.v} dv = (2 pi)^-3/2 (2 pi)^3/2 e^{-|eta|^2/2}."""
    N, Lv = 32, 10.0
    v, eta = orc.grid_0d(N, Lv)
    V2 = v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2
    f = (2 * np.pi) ** -1.5 * np.exp(-V2 / 2)
    out = orc.fft3d(N, v, eta, f.astype(complex), 0)
    E2 = eta[:, None, None] ** 2 + eta[None, :, None] ** 2 + eta[None, None, :] ** 2
    ref = (2 * np.pi) ** -1.5 * np.exp(-E2 / 2)
    # trapezoid aliasing ~ exp(-(2 pi/h_v - |eta|)^2/2): below 1e-13 for |eta_d| <= 2
    m = (np.abs(eta)[:, None, None] <= 2) & (np.abs(eta)[None, :, None] <= 2) & (np.abs(eta)[None, None, :] <= 2)
    assert m.sum() > 100
    assert np.abs(out - ref)[m].max() < 1e-12


@pytest.mark.parametrize("N", [4, 6, 8])
def test_qhat_convolution_theorem(orc, N):
    """With W == 1, sum_c G[c] F[(r + N/2 - c) mod N] is a circular convolution
    shifted by N/2 (src/collisions.c:141-160): check it against FFT convolution."""
    rng = np.random.default_rng(N)
    F = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    G = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    W = np.ones((N ** 3, N ** 3))
    Q = orc.qhat(N, W, F, G)
    conv = np.fft.ifftn(np.fft.fftn(G) * np.fft.fftn(F))
    ref = np.roll(conv, shift=(-(N // 2),) * 3, axis=(0, 1, 2))
    assert rel_l2(Q, ref) < 1e-13


def test_qhat_single_entry_index_map(orc):
    """W with one non-zero entry (r0, c0): only Qhat[r0] = W G[c0] F[s] is non-zero,
    s_d = (r0_d + N/2 - c0_d) mod N, for an asymmetric (r0, c0)."""
    N = 6
    rng = np.random.default_rng(7)
    F = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    G = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    r0, c0 = (1, 4, 5), (5, 0, 2)
    ri = (r0[0] * N + r0[1]) * N + r0[2]
    ci = (c0[0] * N + c0[1]) * N + c0[2]
    W = np.zeros((N ** 3, N ** 3)); W[ri, ci] = 2.5
    Q = orc.qhat(N, W, F, G)
    s = tuple((r0[d] + N // 2 - c0[d]) % N for d in range(3))
    expect = np.zeros((N, N, N), complex); expect[r0] = 2.5 * G[c0] * F[s]
    mask = np.ones((N, N, N), bool); mask[r0] = False
    assert np.abs(Q[mask]).max() == 0.0
    assert abs(Q[r0] - expect[r0]) < 4e-16 * abs(expect[r0])


def _cmat(N, v):
    w = np.ones(N); w[0] = w[-1] = 0.5
    dv = v[1] - v[0]
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    p = (w[I] * w[J] * w[K] * dv ** 3).ravel()
    return np.stack([p, p * v[I].ravel(), p * v[J].ravel(), p * v[K].ravel(),
                     p * 0.5 * (v[I] ** 2 + v[J] ** 2 + v[K] ** 2).ravel()])


@pytest.mark.parametrize("N,Lv", [(8, 5.0), (12, 7.0)])
def test_conserve_is_the_orthogonal_projection(orc, N, Lv):
    """src/conserve.c:207-264 with one species is Q - C^T (C C^T)^-1 C Q."""
    v, _ = orc.grid_0d(N, Lv)
    rng = np.random.default_rng(N)
    Q = rng.standard_normal(N ** 3)
    cons = orc.Conserver(N, v)
    P = cons(Q)
    C = _cmat(N, v)
    ref = Q - C.T @ np.linalg.solve(C @ C.T, C @ Q)
    assert rel_l2(P, ref) < 1e-13
    assert np.abs(C @ P).max() < 1e-13 * np.abs(C @ np.abs(Q)).max()
    assert rel_l2(cons(P), P) < 1e-14  # idempotent


def test_conserve_two_species_moments(orc):
    """Ns=2 (src/conserve.c:213-263): after projection the total mass of each
    species and the total momentum/energy over species vanish."""
    N, Lv = 6, 4.0
    v, _ = orc.grid_0d(N, Lv)
    masses = np.array([1.0, 2.0])
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((4, N ** 3))
    P = orc.Conserver(N, v, masses)(Q)
    C = _cmat(N, v)
    S = [P[0] + P[2], P[1] + P[3]]  # species m: arrays n*Ns+m
    mass = [masses[m] * C[0] @ S[m] for m in range(2)]
    mom_en = sum(masses[m] * C[1:] @ S[m] for m in range(2))
    scale = np.abs(C @ np.abs(Q).sum(0)).max()
    assert max(abs(mass[0]), abs(mass[1])) < 1e-13 * scale
    assert np.abs(mom_en).max() < 1e-13 * scale


def test_lu_pivot_rule(orc):
    """src/conserve.c:115-126 takes the FIRST row whose scaled entry beats the
    diagonal, not the largest; the factors still solve the system exactly."""
    import ctypes
    A = np.array([[1.0, 2.0, 0.0], [3.0, 1.0, 1.0], [9.0, 1.0, 2.0]])
    LU = A.copy().ravel()
    piv = np.zeros(3, dtype=np.int32)
    assert orc.lib().orc_lu_factor(3, LU.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    assert piv[0] == 1  # row 1 (3/3=1 > 1/2) comes before the larger row 2 (9/9=1, ties do not count)
    b = np.array([1.0, -2.0, 0.5])
    x = b.copy()
    orc.lib().orc_lu_solve(3, LU.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.abs(A @ x - b).max() < 1e-14


@pytest.mark.parametrize("lam", [0.0, 1.0, 0.5, -0.5])
def test_radial_integral_against_quadpack(orc, lam):
    """src/weights.c:156-160,203: I = int_0^L r^(lam+2)[sinc(a0 r)sinc(a2 r) - sinc(a1 r)] dr,
    closed forms (lam 0, 1) or Gauss-Legendre (else) against scipy's QUADPACK."""
    import warnings
    warnings.simplefilter("ignore", integrate.IntegrationWarning)
    L = 7.0
    rng = np.random.default_rng(int(lam * 10) + 5)
    cases = [(0.0, 0.0, 0.0), (0.0, 1.3, 0.9), (0.8, 0.0, 0.0), (0.0, 0.4, 0.0), (1.1, 1.1, 1.1)]
    cases += [tuple(rng.uniform(0, 3, 3)) for _ in range(8)]

    def sinc(x):
        return np.sinc(x / np.pi)

    for a0, a1, a2 in cases:
        ref, _ = integrate.quad(lambda r: r ** (lam + 2) * (sinc(a0 * r) * sinc(a2 * r) - sinc(a1 * r)), 0, L,
                                epsabs=1e-14, epsrel=1e-13, limit=400)
        got = orc.ghat_integral(a0, a1, a2, L, lam)
        scale = L ** (lam + 3) / (lam + 3)
        assert abs(got - ref) < 1e-12 * scale, (a0, a1, a2, got, ref)


def test_moments_of_discrete_maxwellian(orc):
    """src/momentRoutines.c:58-183 on a wide grid reproduce (rho, u, T) of a
    Maxwellian to quadrature accuracy, and the KB rule (mass != 1) scales T."""
    N, Lv = 32, 10.0
    v, _ = orc.grid_0d(N, Lv)
    f = orc.maxwellian(v, 1.3, np.array([0.4, -0.2, 0.1]), 0.9)
    rho = orc.density(N, v, f)
    u = orc.bulk_velocity(N, v, f, rho)
    T = orc.temperature(N, v, f, u, rho)
    assert abs(rho - 1.3) < 1e-8 and np.abs(u - [0.4, -0.2, 0.1]).max() < 1e-8 and abs(T - 0.9) < 1e-8


def test_maxpreserve_decomposition(orc):
    """Q(M,g)+Q(g,M)+Q(g,g) = Q(f,f) - Q(M,M) by bilinearity (src/collisions.c:178-210)."""
    N, Lv = 8, 5.0
    v, eta = orc.grid_0d(N, Lv)
    W = orc.weights_iso(N, eta, Lv, 0.0)
    f = orc.bkw_initial(N, v)
    Qmp = orc.compute_q_maxpreserve(N, v, eta, W, f, f)
    rho = orc.density(N, v, f); u = orc.bulk_velocity(N, v, f, rho); T = orc.temperature(N, v, f, u, rho)
    M = orc.maxwellian(v, rho, u, T)
    ref = orc.compute_q(N, v, eta, W, f, f) - orc.compute_q(N, v, eta, W, M, M)
    assert np.abs(Qmp - ref).max() < 1e-12 * np.abs(ref).max()


def _advection_errors(orc, order, sizes):
    """Transport alone of a smooth bump f(x, v) = (1 + 0.3 exp(-((x - 0.5)/0.06)^2)) M(v)
    on [0, 1] with no-flux / copy walls, CFL 0.4; L1 error against the exact
    translation f(x - v_x t, v) at the cell centres.  upwindOne advances by dt;
    advectTwo by dt/2 (two half-dt stages averaged: a Heun step of dt/2, the
    transport half of the Strang splitting of exec/boltz.c:285-354)."""
    N, Lv, T = 4, 2.0, 0.05
    v, _ = orc.grid_1d(N, Lv)
    M = orc.maxwellian_1d(v, 1.0, 0.0, 1.0).ravel()
    vx = np.repeat(v, N * N)
    bump = lambda x: 1.0 + 0.3 * np.exp(-((x - 0.5) / 0.06) ** 2)  # noqa: E731
    errs = []
    for nX in sizes:
        x, dx = orc.uniform_mesh(nX, 1.0, order)
        dt = 0.4 * dx[0] / Lv
        steps = int(round(T / dt))
        f = bump(x)[:, None] * M[None, :]
        for _ in range(steps):
            if order == 1:
                f, fc = orc.upwind_one(N, nX, v, dx, dt, 0, f)
            else:
                f, fc = orc.advect_two(N, nX, v, x, dx, dt, 0, f)
            f = fc.copy()
        xi = x[order:order + nX]
        tfin = steps * dt * (1.0 if order == 1 else 0.5)
        exact = bump(xi[:, None] - vx[None, :] * tfin) * M[None, :]
        errs.append(np.abs(f[order:order + nX] - exact).sum() * dx[0])
    return np.array(errs)


def test_transport_convergence_orders(orc):
    """upwindOne converges at first order and the minmod MUSCL advectTwo at second
    order (src/transportroutines.c:94-238, :241-492) on a smooth profile: pins the
    order-2 restatement, for which the reference holds no fixture."""
    e1 = _advection_errors(orc, 1, [40, 80, 160])
    e2 = _advection_errors(orc, 2, [40, 80, 160])
    r1 = np.log2(e1[:-1] / e1[1:])
    r2 = np.log2(e2[:-1] / e2[1:])
    assert np.all((r1 > 0.8) & (r1 < 1.2)), r1
    assert np.all(r2 > 1.5), r2
    assert e2[-1] < 0.2 * e1[-1]


def _nonuniform_mesh(nX, order, seed):
    """Cell widths drawn in [0.05, 0.15], centres with faces that match:
    x[l+1] - x[l] = (dx[l] + dx[l+1]) / 2 (ghosts included)."""
    rng = np.random.default_rng(seed)
    dx = rng.uniform(0.05, 0.15, nX + 2 * order)
    x = np.concatenate([[0.0], np.cumsum(0.5 * (dx[:-1] + dx[1:]))])
    return x, dx


@pytest.mark.parametrize("stage", ["upwind_two", "advect_two"])
def test_order2_transport_exact_on_linear_data(orc, stage):
    """On data linear in x every minmod slope is the exact gradient (all three
    divided differences of src/transportroutines.c:406-420,443-449 agree), so
    the faces x_l +- dx_l/2 carry the exact values and one upwindTwo stage
    translates the data by 0.5 dt v_i (CFL_NUM = 0.5 dt v/dx, :422,451) on any
    mesh; advectTwo's two stages and average (:477-492) translate it by the
    same amount.  A wrong width, face index, upwind side or time factor breaks
    this on a non-uniform mesh.  Cells next to the walls and the extrapolated
    ghosts (:271-274, :293-296) are excluded."""
    N, Lv, nX = 4, 2.0, 16
    v, _ = orc.grid_1d(N, Lv)
    x, dx = _nonuniform_mesh(nX, 2, seed=3)
    M = orc.maxwellian_1d(v, 1.0, 0.0, 1.0).ravel()
    vx = np.repeat(v, N * N)
    a, b = 1.3, -0.7
    f = (a + b * x)[:, None] * M[None, :]
    dt = 0.8 * dx.min() / Lv
    if stage == "upwind_two":
        _, fc = orc.upwind_two(N, nX, v, x, dx, dt, 0, f)
        cells = range(4, nX)
    else:
        _, fc = orc.advect_two(N, nX, v, x, dx, dt, 0, f)
        cells = range(6, nX - 2)
    for l in cells:
        exact = (a + b * (x[l] - 0.5 * dt * vx)) * M
        assert np.abs(fc[l] - exact).max() <= 1e-13 * np.abs(f).max(), l


def test_order2_transport_is_tvd_on_steps(orc):
    """A minmod MUSCL stage with 0.5 dt |v|/dx <= 0.4 neither raises the total
    variation of a velocity node's profile nor leaves the range of its data
    (src/transportroutines.c:80-90,406-466): two steps on a non-uniform mesh,
    away from the walls."""
    N, Lv, nX = 4, 2.0, 24
    v, _ = orc.grid_1d(N, Lv)
    x, dx = _nonuniform_mesh(nX, 2, seed=5)
    M = orc.maxwellian_1d(v, 1.0, 0.0, 1.0).ravel()
    prof = np.where((np.arange(nX + 4) >= 9) & (np.arange(nX + 4) < 17), 1.0, 0.2)
    prof[13:15] = 0.6
    f = prof[:, None] * M[None, :]
    dt = 0.8 * dx.min() / Lv
    inner = slice(2, nX + 2)
    for fn in (orc.upwind_two, orc.advect_two):
        _, fc = fn(N, nX, v, x, dx, dt, 0, f)
        tv0 = np.abs(np.diff(f[inner], axis=0)).sum(axis=0)
        tv1 = np.abs(np.diff(fc[inner], axis=0)).sum(axis=0)
        assert np.all(tv1 <= tv0 * (1 + 1e-12)), fn.__name__
        assert np.all(fc[inner] <= f.max(axis=0) * (1 + 1e-12))
        assert np.all(fc[inner] >= f.min(axis=0) * (1 - 1e-12))
        assert np.abs(tv1 - tv0).max() > 0  # the data did move


def test_left_diffuse_wall_composes_diffuse_and_copy(orc):
    """bc 3 (ICChoice 1: diffuse left wall, copied right ghost) equals the
    diffuse-wall result away from the right wall and the copy result away from
    the left wall, at order 1 and order 2."""
    N, Lv, nX, dt = 8, 5.0, 6, 0.002
    v, _ = orc.grid_1d(N, Lv)
    rng = np.random.default_rng(5)
    for order in (1, 2):
        x, dx = orc.uniform_mesh(nX, 1.0, order)
        f = np.zeros((nX + 2 * order, N ** 3))
        for c in range(nX):
            f[order + c] = orc.maxwellian_1d(v, 1 + 0.2 * rng.random(), 0.3 * rng.random(), 1 + rng.random()).ravel()
        if order == 1:
            run = lambda bc: orc.upwind_one(N, nX, v, dx, dt, bc, f, T0=2.0)[1]  # noqa: E731
        else:
            run = lambda bc: orc.upwind_two(N, nX, v, x, dx, dt, bc, f, T0=2.0)[1]  # noqa: E731
        mixed, diffuse, copy = run(3), run(1), run(0)
        lo, hi = order, order + nX
        assert np.array_equal(mixed[lo:hi - 1], diffuse[lo:hi - 1])
        assert np.array_equal(mixed[lo + 1:hi], copy[lo + 1:hi])
        assert not np.array_equal(mixed[lo], copy[lo]) and not np.array_equal(mixed[hi - 1], diffuse[hi - 1])


def test_poiseuille_force_adds_y_momentum(orc):
    """The order-2 body force of ICChoice 5 (src/transportroutines.c:429-437,
    :458-466) is -Ma dt/2 d f/d v_y per stage: on a smooth Maxwellian it adds
    Ma dt/2 rho to the y momentum (integration by parts) and leaves density,
    x momentum and the z direction alone, to the quadrature error."""
    N, Lv, nX, dt, Ma = 24, 6.0, 4, 0.01, 1.0
    v, _ = orc.grid_1d(N, Lv)
    x, dx = orc.uniform_mesh(nX, 1.0, 2)
    f = np.tile(orc.maxwellian_1d(v, 1.0, 0.0, 1.0).ravel(), (nX + 4, 1))
    base = orc.upwind_two(N, nX, v, x, dx, dt, 0, f)[1]
    forced = orc.upwind_two(N, nX, v, x, dx, dt, 0, f, Ma=Ma)[1]
    w = np.ones(N); w[0] = w[-1] = 0.5
    h3 = (v[1] - v[0]) ** 3
    W3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() * h3
    VX = np.repeat(v, N * N)
    VY = np.tile(np.repeat(v, N), N)
    VZ = np.tile(v, N * N)
    rho = W3 @ f[3]
    for l in range(2, nX + 2):
        d = forced[l] - base[l]
        assert abs(W3 @ (VY * d) - Ma * 0.5 * dt * rho) < 1e-3 * Ma * 0.5 * dt * rho
        assert abs(W3 @ d) < 1e-9 and abs(W3 @ (VX * d)) < 1e-9 and abs(W3 @ (VZ * d)) < 1e-9


def test_gauss_legendre64_rule(orc):
    """The 64-point rule of gsl_integration_glfixed / gauss_legendre: published
    nodes and weights (Abramowitz & Stegun 25.4 to double precision), weights
    summing to 2, and exactness on even powers up to degree 126."""
    x, w = orc.gauss_legendre64()
    assert x[0] == 0.0243502926634244325089558 and w[0] == 0.0486909570091397203833654
    assert x[31] == 0.9993050417357721394569056 and w[31] == 0.0017832807216964329472961
    assert abs(2 * w.sum() - 2) < 1e-15
    for k in (2, 10, 64, 126):
        assert abs(2 * (w * x ** k).sum() - 2 / (k + 1)) < 5e-15 * 2 / (k + 1), k


@pytest.mark.parametrize("lam", [0.0, 1.0, 2.0])
def test_landau_entry_closed_form_and_quadrature(orc, lam):
    """gHat3L (exec/MPIcollisionRoutines.c:205-248): with the column mode xi = 0
    the integrand is -r^(lam+4) |zeta|^2 sin^3(theta), so the entry is
    -(8 pi / 3) |zeta|^2 L^(lam+5) / (lam+5); a zero row mode gives 0; generic
    entries agree with scipy's adaptive 2D quadrature of the same integral."""
    import math
    from scipy import integrate, special
    Lv = 5.0
    z = (0.3, -0.7, 1.1)
    z2 = sum(t * t for t in z)
    want = -(8 * math.pi / 3) * z2 * Lv ** (lam + 5) / (lam + 5)
    assert abs(orc.landau_entry(z, (0, 0, 0), Lv, lam) / want - 1) < 1e-13
    assert orc.landau_entry((0, 0, 0), (0.55, 1.1, -0.55), Lv, lam) == 0.0
    for zeta, xi in [((0.55, -1.1, 0.55), (1.65, 0.55, -0.55)), ((-1.65, 0.0, 0.55), (0.55, 2.2, 1.1))]:
        zeta, xi = np.array(zeta), np.array(xi)
        zl = np.linalg.norm(zeta)
        a2 = xi @ zeta / zl
        xp = math.sqrt(max(xi @ xi - a2 * a2, 0.0))

        def f(th, r):
            s, c = math.sin(th), math.cos(th)
            return r ** (lam + 2) * s * special.j0(r * xp * s) * (
                -r * r * zl * zl * s * s * math.cos(r * a2 * c) + 4 * r * zl * math.sin(r * a2 * c) * c)
        ref = 2 * math.pi * integrate.dblquad(f, 0, Lv, 0, math.pi, epsabs=1e-12, epsrel=1e-12)[0]
        got = orc.landau_entry(zeta, xi, Lv, lam)
        assert abs(got - ref) < 1e-10 * abs(ref), (got, ref)


def test_cross_species_maxwellians_are_an_equilibrium(orc):
    """Two species at a common temperature, Maxwellians of their own masses,
    are a collision equilibrium: with the table of the pair built as
    exec/boltz.c:157-160 and src/weights.c:190 orient it (species_i the first
    argument's, mu = m_j / (m_i + m_j)), Q(M_0, M_1) vanishes to the spectral
    error; with the masses swapped it does not.  Pins the mass orientation of
    orc_weights_iso and the argument roles of orc_compute_q for unequal masses."""
    N, Lv = 16, 6.0
    v, eta = orc.grid_0d(N, Lv)
    V2 = v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2
    M = lambda m: (m / (2 * np.pi)) ** 1.5 * np.exp(-m * V2 / 2)  # noqa: E731
    right = orc.weights_iso(N, eta, Lv, 1.0, 2.0, 2.0, 1.0, 2.0)
    swapped = orc.weights_iso(N, eta, Lv, 1.0, 2.0, 2.0, 2.0, 1.0)
    q_right = np.abs(orc.compute_q(N, v, eta, right, M(1.0), M(2.0))).max()
    q_swapped = np.abs(orc.compute_q(N, v, eta, swapped, M(1.0), M(2.0))).max()
    assert q_right < 2e-4 and q_swapped > 1e-2, (q_right, q_swapped)


def test_two_identical_species_equal_one_species_at_half_knudsen(orc):
    """With two copies of one species (same mass, same table, same state) every
    pair's Q is Q(f, f), the joint projection splits evenly, and each copy
    advances as one species with Kn / 2: pins the pair sums and the update
    order of the multi-species drivers (exec/boltz.c:193-241, :285-354)."""
    N, Lv, dt, Kn = 8, 5.0, 0.01, 1.0
    v, eta = orc.grid_0d(N, Lv)
    W = orc.weights_iso(N, eta, Lv, 1.0)
    f0 = orc.bkw_initial(N, v).ravel()
    two = orc.run_0d_species(N, v, eta, [W] * 4, [f0, f0], [1.0, 1.0], dt, 3, order=2, Kn=Kn)
    one = orc.run_0d(N, v, eta, W, f0, dt, 3, order=2, Kn=Kn / 2)
    states = []
    orc_f = f0.copy()
    cons = orc.Conserver(N, v)
    for _ in range(3):  # the same single-species Heun step, returning states
        Q = cons(orc.compute_q_maxpreserve(N, v, eta, W, orc_f, orc_f).ravel())
        f1 = orc_f + dt * Q / (Kn / 2)
        Q = cons(orc.compute_q_maxpreserve(N, v, eta, W, f1, f1).ravel())
        orc_f = 0.5 * (orc_f + f1) + 0.5 * dt * Q / (Kn / 2)
        states.append(orc_f.copy())
    assert len(one) == 4
    for t in range(3):
        for l in range(2):
            assert np.abs(two[t][l] - states[t]).max() < 1e-13 * np.abs(states[t]).max()
    x, dx = orc.uniform_mesh(6, 1.0, 1)
    v1, eta1 = orc.grid_1d(N, Lv)
    W1 = orc.weights_iso(N, eta1, Lv, 1.0)
    g0 = np.stack([orc.maxwellian_1d(v1, 1.0 + 0.1 * c, 0.2, 1.0 + 0.05 * c).ravel() for c in range(6)])
    two = orc.run_1d_species(N, v1, eta1, [W1] * 4, [g0, g0], [1.0, 1.0], x, dx, 0.002, 2, 1, Kn, 1)
    ref = []
    orc.run_1d(N, v1, eta1, W1, g0, x, dx, 0.002, 2, 1, Kn / 2, 1, callback=lambda t, f: ref.append(f[1:7].copy()))
    for t in range(2):
        for l in range(2):
            assert np.abs(two[t][l] - ref[t]).max() < 1e-13 * np.abs(ref[t]).max()
