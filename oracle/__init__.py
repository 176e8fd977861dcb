"""CPU oracle for SpectralBTE's collision-operator path (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this package; the product library never does.  The numerical work is in
``spbte_oracle.c`` (built into ``liboracle.so`` by the Makefile beside it); this
module wraps it with ctypes/numpy and restates the reference's time loops:

* ``run_0d``   -- exec/boltz.c:186-250 (0D: ComputeQ_maxPreserve + conserve,
                  Euler or Heun), output rows of src/output.c:159-213;
* ``run_1d``   -- exec/boltz.c:254-345 (1D: advection, per-cell ComputeQ +
                  conserve, Euler or Heun, second half advection), output rows of
                  src/output.c:295-344;
* ``bkw_initial``       -- src/initializer.c:123-135;
* ``maxwellian_1d``     -- src/initializer.c:391-396 (1D convention exp(-|v-u|^2/T)).

Parity status of each function is recorded in the C file's header.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def build_native(outdir):
    """Compile spbte_oracle.c for the host it runs on (-O3 -march=native
    -fopenmp) into outdir; returns the library's path.  This is the
    CPU-baseline build: the tests use the portable liboracle.so."""
    path = os.path.join(outdir, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-std=gnu99", "-fPIC", "-fopenmp", "-shared", "-o", path,
                    os.path.join(_HERE, "spbte_oracle.c"), "-lm"], check=True)
    return path


def use_library(path):
    """Load the oracle functions from `path` (a build of spbte_oracle.c) from now on."""
    global _LIB_PATH, _lib
    _LIB_PATH = path
    _lib = None
    return lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        I, D = ctypes.c_int, ctypes.c_double
        sig = {
            "orc_grid_0d": (None, [I, D, _dp, _dp]),
            "orc_grid_1d": (None, [I, D, _dp, _dp]),
            "orc_fft3d": (None, [I, _dp, _dp, _dp, _dp, I]),
            "orc_qhat": (None, [I, _dp, _dp, _dp, _dp, I]),
            "orc_compute_q": (None, [I, _dp, _dp, _dp, _dp, _dp, _dp, I]),
            "orc_compute_q_maxpreserve": (None, [I, _dp, _dp, D, _dp, _dp, _dp, _dp, I]),
            "orc_density": (D, [I, _dp, D, _dp]),
            "orc_bulk_velocity": (None, [I, _dp, D, _dp, D, _dp]),
            "orc_temperature": (D, [I, _dp, D, _dp, _dp, D]),
            "orc_energy_parts": (None, [I, _dp, _dp, _dp]),
            "orc_lu_factor": (I, [I, _dp, _ip]),
            "orc_lu_solve": (None, [I, _dp, _ip, _dp]),
            "orc_conserve_setup": (I, [I, _dp, I, _dp, _dp, _ip]),
            "orc_conserve": (None, [I, _dp, I, _dp, _dp, _ip, _dp]),
            "orc_ghat_integral": (D, [D, D, D, D, D]),
            "orc_weights_iso": (None, [I, _dp, D, D, D, D, D, D, _dp, I]),
            "orc_gauss_legendre64": (None, [_dp, _dp]),
            "orc_set_wall_kb": (None, [D]),
            "orc_landau_entry": (D, [D, D, D, D, D, D, D, D]),
            "orc_weights_landau_rows": (None, [I, _dp, D, D, ctypes.c_long, ctypes.c_long, _dp]),
            "orc_glance_entry": (D, [D, D, D, D, D, D, D, D, D, I, I, I, I]),
            "orc_weights_glance_entries": (None, [I, _dp, D, D, D, ctypes.c_long, _lp, _lp, I, I, I, I, _dp, I]),
            "orc_weights_load": (I, [ctypes.c_char_p, I, _dp]),
            "orc_diffuse_bc": (None, [I, _dp, D, _dp, _dp, D, I]),
            "orc_upwind_one": (None, [I, I, _dp, D, _dp, D, I, D, D, _dp, _dp]),
            "orc_upwind_two": (None, [I, I, _dp, D, _dp, _dp, D, I, D, D, D, _dp, _dp]),
            "orc_advect_two": (None, [I, I, _dp, D, _dp, _dp, D, I, D, D, D, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _lpp(a):
    return a.ctypes.data_as(_lp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def default_threads():
    return int(os.environ.get("SPBTE_ORACLE_THREADS", os.cpu_count() or 1))


# ----------------------------------------------------------------- grids
def grid_0d(N, Lv):
    v = np.zeros(N); eta = np.zeros(N)
    lib().orc_grid_0d(N, Lv, _p(v), _p(eta))
    return v, eta


def grid_1d(N, Lv):
    v = np.zeros(N); eta = np.zeros(N)
    lib().orc_grid_1d(N, Lv, _p(v), _p(eta))
    return v, eta


# ---------------------------------------------------------- operator parts
def fft3d(N, v, eta, x, invert):
    """x: complex array of N^3 (any shape).  Returns complex (N,N,N)."""
    a = _c(np.asarray(x, dtype=np.complex128).reshape(-1).view(np.float64))
    out = np.zeros_like(a)
    lib().orc_fft3d(N, _p(_c(v)), _p(_c(eta)), _p(a), _p(out), int(invert))
    return out.view(np.complex128).reshape(N, N, N)


def qhat(N, W, Fhat, Ghat, nthreads=None):
    W = _c(W)
    F = _c(np.asarray(Fhat, dtype=np.complex128).reshape(-1).view(np.float64))
    G = _c(np.asarray(Ghat, dtype=np.complex128).reshape(-1).view(np.float64))
    out = np.zeros_like(F)
    lib().orc_qhat(N, _p(W), _p(F), _p(G), _p(out), nthreads or default_threads())
    return out.view(np.complex128).reshape(N, N, N)


def compute_q(N, v, eta, W, f, g, nthreads=None):
    Q = np.zeros(N ** 3)
    lib().orc_compute_q(N, _p(_c(v)), _p(_c(eta)), _p(_c(W)), _p(_c(f).reshape(-1)), _p(_c(g).reshape(-1)), _p(Q),
                        nthreads or default_threads())
    return Q.reshape(N, N, N)


def compute_q_maxpreserve(N, v, eta, W, f, g, mass=1.0, nthreads=None):
    Q = np.zeros(N ** 3)
    lib().orc_compute_q_maxpreserve(N, _p(_c(v)), _p(_c(eta)), mass, _p(_c(W)), _p(_c(f).reshape(-1)),
                                    _p(_c(g).reshape(-1)), _p(Q), nthreads or default_threads())
    return Q.reshape(N, N, N)


# ------------------------------------------------------------------ moments
def density(N, v, f, mass=1.0):
    return lib().orc_density(N, _p(_c(v)), mass, _p(_c(f).reshape(-1)))


def bulk_velocity(N, v, f, rho, mass=1.0):
    u = np.zeros(3)
    lib().orc_bulk_velocity(N, _p(_c(v)), mass, _p(_c(f).reshape(-1)), rho, _p(u))
    return u


def temperature(N, v, f, u, rho, mass=1.0):
    return lib().orc_temperature(N, _p(_c(v)), mass, _p(_c(f).reshape(-1)), _p(_c(u)), rho)


def energy_parts(N, v, f):
    out = np.zeros(2)
    lib().orc_energy_parts(N, _p(_c(v)), _p(_c(f).reshape(-1)), _p(out))
    return out


# ------------------------------------------------------------- conservation
class Conserver:
    """src/conserve.c:17-43 + :207-264: set up C C^T once, project Q arrays."""

    def __init__(self, N, v, masses=(1.0,)):
        self.N = N
        self.v = _c(v)
        self.Ns = len(masses)
        self.masses = _c(masses)
        n = self.Ns + 4
        self.LU = np.zeros(n * n)
        self.pivot = np.zeros(n, dtype=np.int32)
        st = lib().orc_conserve_setup(N, _p(self.v), self.Ns, _p(self.masses), _p(self.LU),
                                      self.pivot.ctypes.data_as(_ip))
        if st != 0:
            raise ValueError("singular C C^T")

    def __call__(self, Q):
        """Q: Ns*Ns arrays of N^3 (array n*Ns+m).  Returns the projected copy."""
        q = _c(Q).copy().reshape(-1)
        lib().orc_conserve(self.N, _p(self.v), self.Ns, _p(self.masses), _p(self.LU), self.pivot.ctypes.data_as(_ip),
                           _p(q))
        return q.reshape(np.shape(Q))


# ------------------------------------------------------------------ weights
def ghat_integral(a0, a1, a2, Lv, lam):
    return lib().orc_ghat_integral(a0, a1, a2, Lv, lam)


def weights_iso(N, eta, Lv, lam, d_i=2.0, d_j=2.0, m_i=1.0, m_j=1.0, nthreads=None):
    W = np.zeros((N ** 3, N ** 3))
    lib().orc_weights_iso(N, _p(_c(eta)), Lv, lam, d_i, d_j, m_i, m_j, _p(W), nthreads or default_threads())
    return W


def gauss_legendre64():
    """(x, w): the 32 positive nodes (ascending) and weights of the 64-point rule."""
    x, w = np.zeros(32), np.zeros(32)
    lib().orc_gauss_legendre64(_p(x), _p(w))
    return x, w


def landau_entry(zeta, xi, Lv, lam):
    """gHat3L (exec/MPIcollisionRoutines.c:219-248) for row mode zeta, column mode xi."""
    return lib().orc_landau_entry(*map(float, zeta), *map(float, xi), float(Lv), float(lam))


def weights_landau(N, eta, Lv, lam, rows=None):
    """Rows of the glance = 0 (Landau) table; rows = (r0, nrows) or all."""
    r0, nr = rows if rows is not None else (0, N ** 3)
    W = np.zeros((nr, N ** 3))
    lib().orc_weights_landau_rows(N, _p(_c(eta)), Lv, lam, r0, nr, _p(W))
    return W


def glance_entry(zeta, xi, Lv, lam, glance, panels=(4, 4, 3, 3)):
    """gHat3 (exec/MPIcollisionRoutines.c:149-203) for row mode zeta, column mode
    xi: fixed composite 16-point rules with (r, phi, log-theta, theta) panel
    counts `panels` in place of the reference's adaptive GSL rules."""
    return lib().orc_glance_entry(*map(float, zeta), *map(float, xi), float(Lv), float(lam), float(glance),
                                  *map(int, panels))


def weights_glance_entries(N, eta, Lv, lam, glance, rows, cols, panels=(4, 4, 3, 3), nthreads=None):
    """Entries (rows[e], cols[e]) of the glance > 0 table."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(len(rows))
    lib().orc_weights_glance_entries(N, _p(_c(eta)), Lv, lam, glance, len(rows), _lpp(rows), _lpp(cols),
                                     *map(int, panels), _p(out), nthreads or default_threads())
    return out


def weights_load(path, N):
    W = np.zeros((N ** 3, N ** 3))
    st = lib().orc_weights_load(str(path).encode(), N, _p(W))
    if st != 0:
        raise IOError(f"weight file {path}: status {st}")
    return W


# ---------------------------------------------------------------- transport
def diffuse_bc(N, v, fin, fout, TW, right, mass=1.0):
    """Returns a copy of fout with the incoming half set."""
    o = _c(fout).copy().reshape(-1)
    lib().orc_diffuse_bc(N, _p(_c(v)), mass, _p(_c(fin).reshape(-1)), _p(o), TW, int(right))
    return o.reshape(np.shape(fout))


def upwind_one(N, nX, v, dx, dt, bc, f, T0=1.0, T1=2.0, mass=1.0):
    """f: (nX+2, N^3) with ghosts; returns (f_with_ghosts_set, fc)."""
    f = _c(f).copy()
    fc = np.zeros_like(f)
    lib().orc_upwind_one(N, nX, _p(_c(v)), mass, _p(_c(dx)), dt, bc, T0, T1, _p(f), _p(fc))
    return f, fc


def upwind_two(N, nX, v, x, dx, dt, bc, f, T0=1.0, T1=2.0, mass=1.0, Ma=0.0):
    f = _c(f).copy()
    fc = np.zeros_like(f)
    lib().orc_upwind_two(N, nX, _p(_c(v)), mass, _p(_c(x)), _p(_c(dx)), dt, bc, T0, T1, Ma, _p(f), _p(fc))
    return f, fc


def advect_two(N, nX, v, x, dx, dt, bc, f, T0=1.0, T1=2.0, mass=1.0, Ma=0.0):
    f = _c(f).copy()
    fc = np.zeros_like(f)
    tmp = np.zeros_like(f)
    lib().orc_advect_two(N, nX, _p(_c(v)), mass, _p(_c(x)), _p(_c(dx)), dt, bc, T0, T1, Ma, _p(f), _p(fc), _p(tmp))
    return f, fc


# --------------------------------------------------------- initial states
def bkw_initial(N, v, K=None, T=1.0):
    """src/initializer.c:123-135 (K = 1 - exp(-5.5/6) there)."""
    if K is None:
        K = 1.0 - np.exp(-5.5 / 6.0)
    V2 = (v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2)
    return (np.exp(-V2 / (2 * K * T * T)) / (2.0 * (2 * np.pi * K * T * T) ** 1.5)
            * ((5 * K - 3) / K + (1 - K) * V2 / (K * K * T * T)))


def maxwellian_1d(v, rho, ux, T):
    """src/initializer.c:366,388,395,410: rho exp(-|v-u|^2/T)/(pi T)^(3/2)."""
    V2 = ((v[:, None, None] - ux) ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2)
    return rho * np.exp(-V2 / T) / ((T * np.pi) * np.sqrt(T * np.pi))


def maxwellian(v, rho, u, T):
    """rho (2 pi T)^-3/2 exp(-|v-u|^2/(2T)) (src/collisions.c:98-103)."""
    V2 = ((v[:, None, None] - u[0]) ** 2 + (v[None, :, None] - u[1]) ** 2 + (v[None, None, :] - u[2]) ** 2)
    return rho * (2 * np.pi * T) ** -1.5 * np.exp(-V2 / (2 * T))


# ------------------------------------------------------------------ drivers
def moments_row_0d(N, v, f, t):
    """src/output.c:159-213: time, rho, u_x, T, p, neg/pos energy, slice."""
    rho = density(N, v, f)
    u = bulk_velocity(N, v, f, rho)
    T = temperature(N, v, f, u, rho)
    e = energy_parts(N, v, f)
    ff = np.asarray(f).reshape(-1)
    sl = [ff[N // 2 + N * (N // 2 + N * l)] for l in range(N)]
    return [t, rho, u[0], T, rho * T, e[1] / e[0]] + sl


def run_0d(N, v, eta, W, f0, dt, nsteps, order=2, Kn=1.0, out_every=1, nthreads=None):
    """exec/boltz.c:186-250 for one species.  Returns list of moment rows."""
    cons = Conserver(N, v)
    f = _c(f0).reshape(N, N, N).copy()
    rows = [moments_row_0d(N, v, f, 0.0)]
    for t in range(nsteps):
        Q = cons(compute_q_maxpreserve(N, v, eta, W, f, f, nthreads=nthreads))
        if order == 1:
            f = f + dt * Q / Kn
        else:
            f1 = f + dt * Q / Kn
            Q = cons(compute_q_maxpreserve(N, v, eta, W, f1, f1, nthreads=nthreads))
            f = 0.5 * (f + f1) + 0.5 * dt * Q / Kn
        if (t + 1) % out_every == 0:
            rows.append(moments_row_0d(N, v, f, dt * (t + 1)))
    return rows


def run_0d_species(N, v, eta, Ws, f0, masses, dt, nsteps, order=2, Kn=1.0, nthreads=None):
    """exec/boltz.c:193-241 with Ns species: Ws[m*Ns + l] is the table of pair
    (l, m), Q[m*Ns + l] = ComputeQ_maxPreserve(f_l, f_m) (whose Maxwellian uses
    species 0's mass, src/collisions.c:95-97), the joint projection, and f_l
    updated with Q[m*Ns + l] in m order.  f0: (Ns, N^3).  Returns the states
    (Ns, N^3) after each step."""
    Ns = len(masses)
    cons = Conserver(N, v, masses)
    f = [np.array(x, dtype=np.float64).reshape(-1).copy() for x in f0]

    def all_q(ff):
        Q = np.zeros((Ns * Ns, N ** 3))
        for l in range(Ns):
            for m in range(Ns):
                Q[m * Ns + l] = compute_q_maxpreserve(N, v, eta, Ws[m * Ns + l], ff[l], ff[m], mass=masses[0],
                                                      nthreads=nthreads).ravel()
        return cons(Q)

    out = []
    for _ in range(nsteps):
        Q = all_q(f)
        if order == 1:
            for l in range(Ns):
                for m in range(Ns):
                    f[l] = f[l] + dt * Q[m * Ns + l] / Kn
        else:
            f1 = [x.copy() for x in f]
            for l in range(Ns):
                for m in range(Ns):
                    f1[l] = f1[l] + dt * Q[m * Ns + l] / Kn
            Q = all_q(f1)
            for l in range(Ns):
                f[l] = 0.5 * (f[l] + f1[l])
                for m in range(Ns):
                    f[l] = f[l] + 0.5 * dt * Q[m * Ns + l] / Kn
        out.append(np.array(f))
    return out


def run_1d_species(N, v, eta, Ws, f0, masses, x, dx, dt, nsteps, order, Kn, bc, T0=1.0, T1=2.0, nthreads=None):
    """exec/boltz.c:254-354 with Ns species on one rank: each species
    transported with its own mass at the walls (src/boundaryConditions.c:56-62),
    per cell Q[n*Ns + m] = ComputeQ(f_m, f_n) with table Ws[n*Ns + m], the joint
    projection, Euler or Heun.  f0: (Ns, nX, N^3).  Returns the interior states
    (Ns, nX, N^3) after each step."""
    Ns = len(masses)
    f0 = _c(f0)
    nX = f0.shape[1]
    n3 = N ** 3
    cons = Conserver(N, v, masses)
    f = []
    for m in range(Ns):
        a = np.zeros((nX + 2 * order, n3))
        a[order:order + nX] = f0[m]
        f.append(a)

    def cell_q(cells):
        Q = np.zeros((Ns * Ns, n3))
        for m in range(Ns):
            for n in range(Ns):
                Q[n * Ns + m] = compute_q(N, v, eta, Ws[n * Ns + m], cells[m], cells[n], nthreads=nthreads).ravel()
        return cons(Q)

    lib().orc_set_wall_kb(1.0 if masses[0] == 1.0 else 1.380658e-23)
    out = []
    for _ in range(nsteps):
        fc = []
        for m in range(Ns):
            if order == 1:
                f[m], c = upwind_one(N, nX, v, dx, dt, bc, f[m], T0, T1, mass=masses[m])
            else:
                f[m], c = advect_two(N, nX, v, x, dx, dt, bc, f[m], T0, T1, mass=masses[m])
            fc.append(c)
        for l in range(order, nX + order):
            Q = cell_q([fc[m][l] for m in range(Ns)])
            if order == 1:
                for m in range(Ns):
                    f[m][l] = fc[m][l]
                    for n in range(Ns):
                        f[m][l] = f[m][l] + dt * Q[n * Ns + m] / Kn
            else:
                f1 = []
                for m in range(Ns):
                    a = fc[m][l].copy()
                    for n in range(Ns):
                        a = a + dt * Q[n * Ns + m] / Kn
                    f1.append(a)
                Q = cell_q(f1)
                for m in range(Ns):
                    fc[m][l] = 0.5 * fc[m][l] + 0.5 * f1[m]
                    for n in range(Ns):
                        fc[m][l] = fc[m][l] + 0.5 * dt * Q[n * Ns + m] / Kn
        if order == 2:
            for m in range(Ns):
                _, fnew = advect_two(N, nX, v, x, dx, dt, bc, fc[m].copy(), T0, T1, mass=masses[m])
                f[m][order:order + nX] = fnew[order:order + nX]
        out.append(np.array([f[m][order:order + nX] for m in range(Ns)]))
    lib().orc_set_wall_kb(0.0)
    return out


def uniform_mesh(nX, length, order):
    """src/mesh_setup.c:121-146 for one rank (x = cell centres, ghosts by
    extension), with the right ghosts of :165-175 filled in."""
    dxz = length / nX
    n = nX + 2 * order
    x = np.zeros(n); dx = np.full(n, dxz)
    # the reference accumulates x_edge += dx; reproduce that rounding
    xe = 0.0
    for j in range(nX):
        x[order + j] = xe + 0.5 * dxz
        xe += dxz
    for g in range(order - 1, -1, -1):
        x[g] = x[g + 1] - dx[g + 1]
    for g in range(order + nX, n):
        x[g] = x[g - 1] + dx[g - 1]
    return x, dx


def run_1d(N, v, eta, W, f0, x, dx, dt, nsteps, order, Kn, bc, T0=1.0, T1=2.0, out_every=1, nthreads=None,
           callback=None, Ma=0.0):
    """exec/boltz.c:254-345 for one species on one rank.  f0: (nX, N^3).
    Returns list of snapshots, each (nX, 4) = rho, u_x, T, p per cell."""
    nX = f0.shape[0]
    n3 = N ** 3
    cons = Conserver(N, v)
    f = np.zeros((nX + 2 * order, n3))
    f[order:order + nX] = _c(f0).reshape(nX, n3)

    def snap(ff):
        out = np.zeros((nX, 4))
        for l in range(nX):
            c = ff[order + l]
            rho = density(N, v, c)
            u = bulk_velocity(N, v, c, rho)
            T = temperature(N, v, c, u, rho)
            out[l] = (rho, u[0], T, rho * T)
        return out

    snaps = [snap(f)]
    for t in range(nsteps):
        if order == 1:
            f, fc = upwind_one(N, nX, v, dx, dt, bc, f, T0, T1)
        else:
            f, fc = advect_two(N, nX, v, x, dx, dt, bc, f, T0, T1, Ma=Ma)
        for l in range(order, nX + order):
            Q = cons(compute_q(N, v, eta, W, fc[l], fc[l], nthreads=nthreads).reshape(-1))
            if order == 1:
                f[l] = fc[l] + dt * Q / Kn
            else:
                f1 = fc[l] + dt * Q / Kn
                Q = cons(compute_q(N, v, eta, W, f1, f1, nthreads=nthreads).reshape(-1))
                fc[l] = 0.5 * fc[l] + 0.5 * f1 + 0.5 * dt * Q / Kn
        if order == 2:
            fc2 = fc.copy()
            fc2, fnew = advect_two(N, nX, v, x, dx, dt, bc, fc2, T0, T1, Ma=Ma)
            f[order:order + nX] = fnew[order:order + nX]
        if callback is not None:
            callback(t, f)
        if (t + 1) % out_every == 0:
            snaps.append(snap(f))
    return snaps
