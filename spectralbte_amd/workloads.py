"""Initial states of the benchmark workloads (synthetic, seeded).

All Maxwellians use the moment convention of the operator itself,
M = rho (2 pi T)^-3/2 exp(-|v-u|^2/(2T)) (src/collisions.c:98-103), so T is the
kinetic temperature reported by src/momentRoutines.c:168-183 (KB = 1, m = 1).
The reference's 1D initial conditions use exp(-|v-u|^2/T)/(pi T)^3/2 instead
(src/initializer.c:366,410), i.e. a temperature parameter twice the kinetic one;
these workloads state their temperatures in kinetic units.
"""
import numpy as np

GAMMA = 5.0 / 3.0  # monatomic gas

# The Mach-1.2 shock's domain length: its 256 cells are then 0.039 wide, so
# the order-2 stages' Courant number 0.5 dt |v| / dx stays at 0.21 or less for
# |v| <= L_v = 8 at dt = 0.002 (on a unit domain it is 2.0 and the run
# diverges within 50 steps).
SHOCK_LENGTH = 10.0
# The heat-transfer run's time step with the glancing-collision table as the
# reference writes it (no trapezoid or h_eta^3 factors, so Q is about 1e4
# times the isotropic table's): at Kn = 0.05 the explicit step diverges
# within 200 steps at 1e-6 and 3e-7.
HEAT_GLANCE_DT = 1e-7


def grid_0d(N, Lv):
    """src/initializer.c:66-82."""
    dv = 2 * Lv / (N - 1)
    v = -Lv + np.arange(N) * dv
    deta = (2 * np.pi / N) / dv
    Leta = 0.5 * N * deta if N % 2 == 0 else 0.5 * (N - 1) * deta
    return v, -Leta + np.arange(N) * deta


def grid_1d(N, Lv):
    """src/initializer.c:256-266."""
    dv = 2 * Lv / (N - 1)
    v = -Lv + np.arange(N) * dv
    Leta = 0.5 * (N - 1) * np.pi / Lv
    deta = np.pi * (N - 1) / (N * Lv)
    return v, -Leta + np.arange(N) * deta


def maxwellian(v, rho, u, T):
    V2 = (v[:, None, None] - u[0]) ** 2 + (v[None, :, None] - u[1]) ** 2 + (v[None, None, :] - u[2]) ** 2
    return rho * (2 * np.pi * T) ** -1.5 * np.exp(-V2 / (2 * T))


def rankine_hugoniot(mach, rho1=1.0, T1=1.0):
    """Upstream (rho1, u1, T1) and downstream states of a normal shock in a
    monatomic gas: u1 = M sqrt(gamma T1), rho2/rho1 = 4M^2/(M^2+3),
    T2/T1 = (5M^2-1)(M^2+3)/(16 M^2), u2 = u1 rho1/rho2."""
    m2 = mach * mach
    u1 = mach * np.sqrt(GAMMA * T1)
    rho2 = rho1 * (GAMMA + 1) * m2 / ((GAMMA - 1) * m2 + 2)
    T2 = T1 * (2 * GAMMA * m2 - (GAMMA - 1)) * ((GAMMA - 1) * m2 + 2) / ((GAMMA + 1) ** 2 * m2)
    return (rho1, u1, T1), (rho2, u1 * rho1 / rho2, T2)


def shock_profile(N, Lv, nX, mach=1.2, blend_cells=20.0):
    """Mach-`mach` normal shock on nX cells in the Mott-Smith form: every cell
    holds (1 - s_l) M_up + s_l M_down, the upstream and Rankine-Hugoniot
    downstream Maxwellians weighted by a tanh blend s_l whose 1%-99% width is
    `blend_cells` cells (tanh argument (l + 1/2 - nX/2) / (blend_cells / (2 atanh 0.98))).
    Inside the layer the cells are bimodal, away from equilibrium, so Q is
    far from zero there.  Returns (nX, N^3)."""
    v, _ = grid_1d(N, Lv)
    (r1, u1, T1), (r2, u2, T2) = rankine_hugoniot(mach)
    width = blend_cells / (2 * np.arctanh(0.98))
    m_up = maxwellian(v, r1, (u1, 0.0, 0.0), T1).ravel()
    m_dn = maxwellian(v, r2, (u2, 0.0, 0.0), T2).ravel()
    f = np.empty((nX, N ** 3))
    for l in range(nX):
        s = 0.5 * (1 + np.tanh((l + 0.5 - nX / 2) / width))
        f[l] = (1 - s) * m_up + s * m_dn
    return f


def two_maxwellian_ensemble(N, Lv, n, seed):
    """n 0D problems: seeded sums of two Maxwellians, weights U(0.2,0.8), means
    U(-1.5,1.5)^3, temperatures U(0.5,2), renormalised to unit density on the
    grid's own quadrature (src/momentRoutines.c:58-72).  Returns (n, N^3)."""
    v, _ = grid_0d(N, Lv)
    w = np.ones(N); w[0] = w[-1] = 0.5
    dv3 = (v[1] - v[0]) ** 3
    W3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() * dv3
    rng = np.random.default_rng(seed)
    out = np.empty((n, N ** 3))
    for i in range(n):
        a = rng.uniform(0.2, 0.8, 2)
        f = sum(a[k] * maxwellian(v, 1.0, rng.uniform(-1.5, 1.5, 3), rng.uniform(0.5, 2.0)).ravel() for k in range(2))
        out[i] = f / (W3 @ f)
    return out


def bkw_state(N, Lv, K):
    """The BKW exact solution at K on the 0D grid (src/initializer.c:123-135,
    unit density, zero bulk velocity, unit temperature):
    (2 pi K)^-3/2 exp(-|v|^2/(2K)) [(5K-3)/(2K) + (1-K)/(2K^2) |v|^2].  (N^3,)"""
    v, _ = grid_0d(N, Lv)
    V2 = (v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2).ravel()
    return (2 * np.pi * K) ** -1.5 * np.exp(-V2 / (2 * K)) * ((5 * K - 3) / (2 * K) + (1 - K) / (2 * K * K) * V2)
