"""GPU parity of the glancing-collision weight table (glance > 0; gHat3 of
exec/MPIcollisionRoutines.c:33-203, written by exec/MPIWeightGenerator.c) and
of the aniso_weights.h entry point with glance > 0 (src/aniso_weights.c:37-146).

The GPU evaluates a reduced 2-D form of the integral from tabulated radial
functions; the oracle sums the reference's 3-D form with composite
Gauss-Legendre rules.  Tolerance: 1e-10 of the largest |W| of the compared
entries (both sides are accurate to about 1e-12 of it; the reference's own
adaptive rules stop at 1e-6)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sample(N, rng, n=40):
    n3 = N ** 3
    h = N // 2
    zero = h + N * (h + N * h)  # the zero mode (eta[N/2] = 0)
    rows = np.concatenate([[0, n3 - 1, zero, zero, 0, n3 - 1], rng.integers(0, n3, n)])
    cols = np.concatenate([[n3 - 1, 0, 0, n3 - 1, zero, zero], rng.integers(0, n3, n)])
    return rows, cols


@pytest.mark.parametrize("N,Lv,lam,g,panels", [(6, 3.0, -3.0, 1e-3, (6, 6, 5, 4)), (8, 4.0, 1.0, 0.01, (6, 6, 4, 4)),
                                               (8, 5.0, 0.5, 0.3, (6, 6, 3, 4)), (8, 4.0, 0.0, 2.0, (6, 6, 3, 4))])
def test_glance_table_matches_oracle(spb, orc, N, Lv, lam, g, panels):
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(lam, g)
    W = op.get_weights()
    op.close()
    assert np.isfinite(W).all()
    rows, cols = _sample(N, np.random.default_rng(N))
    ref = orc.weights_glance_entries(N, eta, Lv, lam, g, rows, cols, panels)
    scale = np.abs(ref).max()
    err = np.abs(W[rows, cols] - ref)
    print(N, g, "max err / max|W|", err.max() / scale)
    assert err.max() <= 1e-10 * scale
    h = N // 2
    zero = h + N * (h + N * h)
    assert np.abs(W[zero]).max() <= 1e-13 * scale  # a zero row mode gives 0


def test_glance_table_n16_symmetry_and_oracle(spb, orc):
    """N = 16: the whole table is invariant under the same cyclic permutation of
    the axes of both modes (the entry depends on |zeta|, xi.zeta and |xi| only),
    and sampled entries match the oracle."""
    N, Lv, lam, g = 16, 5.0, 1.0, 1e-4
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(lam, g)
    W = op.get_weights()
    op.close()
    n = N
    idx = np.arange(n ** 3)
    i, j, k = idx // (n * n), (idx // n) % n, idx % n
    perm = k * n * n + i * n + j  # (i, j, k) -> (k, i, j)
    scale = np.abs(W).max()
    for r0 in range(0, n ** 3, 257):
        assert np.abs(W[perm[r0], perm] - W[r0]).max() <= 1e-12 * scale, r0
    rows, cols = _sample(N, np.random.default_rng(16), 10)
    ref = orc.weights_glance_entries(N, eta, Lv, lam, g, rows, cols, (10, 10, 7, 6))
    assert np.abs(W[rows, cols] - ref).max() <= 1e-10 * np.abs(ref).max()


def test_glance_refuses_out_of_range(spb, orc):
    v, eta = orc.grid_0d(4, 3.0)
    op = spb.Operator(4, 3.0, v, eta)
    for g in (0.0, -0.1, 3.2, float("nan")):
        with pytest.raises(spb.SpbteError):
            op.generate_weights_glance(1.0, g)
    op.close()


def test_weightgen_writes_glance_file(spb, orc, tmp_path):
    """weightgen_amd N glance beta lambda L_v with glance > 0 writes
    Weights/N6_AnIso_L_v3_lambda1_glance0.01.wts (exec/MPIWeightGenerator.c:80-83,
    :101-118) on the generator's grid, and keeps it on a second call."""
    N, Lv, lam, g = 6, 3.0, 1.0, 0.01
    (tmp_path / "Weights").mkdir()
    spb.weightgen(N, g, 1, lam, Lv, workdir=tmp_path)
    path = tmp_path / "Weights" / "N6_AnIso_L_v3_lambda1_glance0.01.wts"
    W = np.fromfile(path).reshape(N ** 3, N ** 3)
    hv = 2 * Lv / (N - 1)
    he = 2 * np.pi / (N * hv)
    eta = -0.5 * N * he + he * np.arange(N)
    rows, cols = _sample(N, np.random.default_rng(6), 20)
    ref = orc.weights_glance_entries(N, eta, Lv, lam, g, rows, cols, (6, 6, 5, 4))
    assert np.abs(W[rows, cols] - ref).max() <= 1e-10 * np.abs(ref).max()
    mtime = os.path.getmtime(path)
    spb.weightgen(N, g, 1, lam, Lv, workdir=tmp_path)
    assert os.path.getmtime(path) == mtime


def test_initialize_weights_aniso_glance_file_and_missing_file(spb, orc, tmp_path):
    """initialize_weights_AnIso with glance > 0 (src/aniso_weights.c:37-146):
    the glance file, when present, is loaded; when missing, the Landau table is
    generated and stored under the glance name, as the reference does (its
    file-scope glance is never assigned, :31, :417)."""
    import ctypes
    N, Lv, lam, g = 4, 5.0, 1.0, 0.01
    v, eta = orc.grid_0d(N, Lv)
    L = spb.lib()
    L.initialize_weights_AnIso.restype = None
    L.initialize_weights_AnIso.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]
    eta_c = np.ascontiguousarray(eta)

    def call(wd):
        cw = ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))()
        L.alloc_weights(N, ctypes.byref(cw), 1)
        cwd = os.getcwd()
        os.chdir(wd)
        try:
            L.initialize_weights_AnIso(N, eta_c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), Lv, lam, 0,
                                       ctypes.cast(cw[0], ctypes.c_void_p), g)
        finally:
            os.chdir(cwd)
        got = np.stack([np.ctypeslib.as_array(cw[0][r], shape=(N ** 3,)).copy() for r in range(N ** 3)])
        L.dealloc_weights(N, cw[0])
        return got

    name = "N4_AnIso_L_v5_lambda1_glance0.01.wts"
    (tmp_path / "a" / "Weights").mkdir(parents=True)
    got = call(tmp_path / "a")
    landau = orc.weights_landau(N, eta, Lv, lam)
    assert np.abs(got - landau).max() <= 1e-12 * np.abs(landau).max()
    assert np.array_equal(np.fromfile(tmp_path / "a" / "Weights" / name).reshape(64, 64), got)
    (tmp_path / "b" / "Weights").mkdir(parents=True)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(lam, g)
    op.store_weights_file(str(tmp_path / "b" / "Weights" / name))
    Wg = op.get_weights()
    op.close()
    assert np.array_equal(call(tmp_path / "b"), Wg)


@pytest.mark.parametrize("order", [1, 2])
def test_heat_transfer_run_with_glance_table(spb, orc, order):
    """The heat-transfer configuration with the anisotropic glancing-collision
    table (grazing parameter 0.1, lambda 1, L_v 7): a Maxwellian at rho 1,
    u 0, T 1.5 with 1% seeded Gaussian density noise, diffuse walls at T 1 and
    T 2, Kn 0.05; the GPU run (graph path) against the oracle's 1D driver
    (exec/boltz.c:285-354) with the same table, which the tests above pin.  As
    the reference writes the table (no trapezoid or h_eta^3 factors), Q is
    about 1e4 times the isotropic table's, so the explicit step needs dt near
    1e-6 at Kn 0.05."""
    N, Lv, nX, Kn, dt, steps = 8, 7.0, 6, 0.05, 2e-6, 3
    v, eta = orc.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(1.0, 0.1)
    W = op.get_weights()
    x, dx = orc.uniform_mesh(nX, 1.0, order)
    noise = 1.0 + 0.01 * np.random.default_rng(128).standard_normal(nX)
    f0 = np.stack([orc.maxwellian_1d(v, noise[c], 0.0, 1.5).ravel() for c in range(nX)])
    run = spb.Run.mesh(op, f0, x, dx, order=order, dt=dt, Kn=Kn, bc=spb.BC_DIFFUSE)
    run.use_graph(True)
    states = []
    orc.run_1d(N, v, eta, W, f0, x, dx, dt, steps, order, Kn, spb.BC_DIFFUSE,
               callback=lambda t, f: states.append(f[order:nX + order].copy()))
    for t in range(steps):
        run.step(1)
        g = run.state()
        assert np.abs(g - states[t]).max() < 1e-12 * np.abs(states[t]).max(), t
    assert np.abs(states[-1] - f0).max() > 1e-4 * np.abs(f0).max()  # the collisions did act
    run.close()
    op.close()
