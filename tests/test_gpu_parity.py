"""GPU parity: every kernel of the MI355X operator against the CPU oracle on the
same seeded inputs, through the C ABI.  Tolerances: FP64 paths 1e-12 relative
L2 (the reference-method bar; transforms 1e-13); reduced-precision forms 1e-6
relative L2 against the FP64 result; conservation |C.Q| <= 1e-13 relative."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu


def _grid(orc, N, Lv, kind):
    return (orc.grid_0d if kind == "0d" else orc.grid_1d)(N, Lv)


def _two_maxwellians(orc, v, rng):
    """Seeded sum of two Maxwellians with unit density (the ensemble workload's initial state)."""
    f = 0
    ws = rng.uniform(0.2, 0.8, 2)
    for w in ws:
        f = f + w * orc.maxwellian(v, 1.0, rng.uniform(-1.5, 1.5, 3), rng.uniform(0.5, 2.0))
    return f / ws.sum()


def _cmoments(N, v, Q):
    w = np.ones(N); w[0] = w[-1] = 0.5
    dv = v[1] - v[0]
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    p = (w[I] * w[J] * w[K] * dv ** 3).ravel()
    C = np.stack([p, p * v[I].ravel(), p * v[J].ravel(), p * v[K].ravel(),
                  p * 0.5 * (v[I] ** 2 + v[J] ** 2 + v[K] ** 2).ravel()])
    Q = np.asarray(Q).reshape(-1, N ** 3)
    return Q @ C.T, np.abs(Q) @ np.abs(C.T)


# ------------------------------------------------------------------ transforms
# Grids of 2^a 3^b points run as LDS FFTs (radix 4, 2, 3: N=6 is 2x3, 9 is 3x3,
# 18 is 2x3x3, 48 is 4x4x3); N=10 keeps the per-axis matrix form.
@pytest.mark.parametrize("N,Lv,kind", [(8, 5.0, "0d"), (16, 5.0, "0d"), (24, 6.0, "0d"), (32, 8.0, "1d"), (12, 7.0, "1d"),
                                       (6, 4.0, "0d"), (9, 5.0, "1d"), (18, 6.0, "0d"), (48, 8.0, "1d"), (10, 5.0, "0d")])
@pytest.mark.parametrize("invert", [0, 1])
def test_transform(spb, orc, N, Lv, kind, invert):
    v, eta = _grid(orc, N, Lv, kind)
    rng = np.random.default_rng(N + 100 * invert)
    x = rng.standard_normal((3, N, N, N)) + 1j * rng.standard_normal((3, N, N, N))
    op = spb.Operator(N, Lv, v, eta)
    y = op.transform(x, invert=invert)
    for b in range(3):
        assert rel_l2(y[b], orc.fft3d(N, v, eta, x[b], invert)) < 1e-13


def test_fft3D_abi(spb, orc):
    """fft3D (src/collisions.h:42) through the reference interface, in == out allowed."""
    N, Lv = 16, 5.0
    v, eta = orc.grid_0d(N, Lv)
    spb.initialize_coll(N, Lv, v, eta)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    assert rel_l2(spb.fft3D(x, 0), orc.fft3d(N, v, eta, x, 0)) < 1e-13
    assert rel_l2(spb.fft3D(x, 1), orc.fft3d(N, v, eta, x, 1)) < 1e-13
    spb.dealloc_coll()


# ------------------------------------------------------------- weighted sum
@pytest.mark.parametrize("N,ncell", [(8, 2), (16, 3), (24, 3), (12, 2), (9, 2), (10, 3)])
def test_qhat_fp64_random(spb, orc, N, ncell):
    """Random W, F, G (ragged cell count vs the tile's cells per workgroup)."""
    v, eta = orc.grid_0d(N, 5.0)
    rng = np.random.default_rng(N)
    W = rng.standard_normal((N ** 3, N ** 3))
    F = rng.standard_normal((ncell, N, N, N)) + 1j * rng.standard_normal((ncell, N, N, N))
    G = rng.standard_normal((ncell, N, N, N)) + 1j * rng.standard_normal((ncell, N, N, N))
    op = spb.Operator(N, 5.0, v, eta)
    op.set_weights(W)
    Q = op.qhat(F, G, "fp64")
    for b in range(ncell):
        assert rel_l2(Q[b], orc.qhat(N, W, F[b], G[b])) < 1e-12
    if N in (16, 24):
        # FP32 products through the FP32 operand layout (values, then part sums)
        Q32 = op.qhat(F, G, "fp32a")
        for b in range(ncell):
            assert rel_l2(Q32[b], Q[b]) < 1e-5, (b, rel_l2(Q32[b], Q[b]))
    if N in (9, 10, 12):
        # grids without reduced-precision tiles run FP64 only, and say so
        with pytest.raises(spb.SpbteError, match="reduced-precision"):
            op.qhat(F, G, "fp32")
    op.close()


@pytest.mark.parametrize("N", [16, 24])
def test_qhat_tiles_agree(spb, orc, N):
    """Every instantiated tile gives the same FP64 sum (index map of each lane layout)."""
    v, eta = orc.grid_0d(N, 5.0)
    rng = np.random.default_rng(5)
    W = rng.standard_normal((N ** 3, N ** 3))
    ncell = 5
    F = rng.standard_normal((ncell, N ** 3)) + 1j * rng.standard_normal((ncell, N ** 3))
    G = rng.standard_normal((ncell, N ** 3)) + 1j * rng.standard_normal((ncell, N ** 3))
    op = spb.Operator(N, 5.0, v, eta)
    op.set_weights(W)
    ref = np.stack([orc.qhat(N, W, F[b], G[b]).ravel() for b in range(ncell)])
    tiles = {16: [(4, 4, 1, 0), (4, 4, 2, 0), (4, 8, 2, 0), (4, 8, 4, 0), (4, 1, 1, 1), (4, 1, 2, 1),
                  (1, 1, 1, 16), (1, 1, 1, 64), (2, 1, 1, 32), (4, 1, 2, 64),    # x-planes cut into y-line ranges
                  (1, 2, 1, 32), (1, 4, 1, 64)],                                 # one-row waves, 2 and 4 a workgroup
             24: [(6, 8, 1, 0), (6, 8, 2, 0), (6, 8, 4, 0), (6, 3, 1, 1), (3, 3, 1, 1), (3, 3, 2, 1), (4, 2, 1, 1),
                  (4, 2, 2, 1), (8, 1, 1, 1)]}[N]
    for t in tiles:
        op.set_tile("fp64", False, *t)
        assert rel_l2(op.qhat(F, G, "fp64"), ref) < 1e-12, t


def test_n32_tiles_agree(spb, orc):
    """Every N=32 tile of both kernel kinds, one and two products, gives the FP64
    sum of the tiled kernel's (4 rows, 4 lines, 2 cells) tile, which
    test_n32_single_cell_against_oracle checks against the oracle."""
    N, Lv = 32, 6.0
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    rng = np.random.default_rng(32)
    f = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(3)])
    tiles = [(4, 4, 2, 0), (8, 2, 1, 1), (8, 2, 2, 1), (4, 4, 2, 1), (4, 2, 2, 1), (4, 4, 4, 1), (4, 4, 1, 1),
             (4, 2, 1, 1), (4, 1, 1, 1), (4, 1, 2, 1), (4, 4, 1, 2), (4, 4, 2, 4), (4, 2, 1, 8), (8, 2, 1, 4),
             (4, 4, 1, 32)]
    for mp in (False, True):
        ref = None
        for t in tiles:
            try:
                op.set_tile("fp64", mp, *t)
            except spb.SpbteError:
                continue  # not instantiated for this product count
            Q = op.compute_q(f, maxpreserve=mp)
            if ref is None:
                ref = Q
            assert rel_l2(Q, ref) < 1e-13, (mp, t)
    op.close()


def test_n32_fp32_products_streaming(spb, orc):
    """The FP32-product form on the streaming kernel (every instantiated N=32
    tile, one and two products, whole and split x-plane sums) against the FP64
    sum: within 1e-6 relative L2 per problem, like the other reduced forms.  The
    16-wave tile (four cells per workgroup, here 5 problems: one group padded)
    runs each lane's sum in the same order as the 4-wave tile: bit-identical."""
    N, Lv = 32, 6.0
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    rng = np.random.default_rng(33)
    f = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(5)])
    ran = 0
    for mp in (False, True):
        op.set_tile("fp64", mp, 4, 4, 2, 0)
        ref = op.compute_q(f, maxpreserve=mp)
        got = {}
        for t in [(4, 4, 1, 1), (4, 4, 2, 1), (4, 4, 4, 1), (4, 4, 1, 4), (4, 4, 2, 2), (4, 16, 1, 1)]:
            try:
                op.set_tile("fp32a", mp, *t)
            except spb.SpbteError:
                continue  # not instantiated for this product count
            Q = op.compute_q(f, form="fp32a", maxpreserve=mp)
            got[t] = Q
            for b in range(len(f)):
                assert rel_l2(Q[b], ref[b]) < 1e-6, (mp, t, b)
            ran += 1
        if mp:
            assert np.array_equal(got[(4, 16, 1, 1)], got[(4, 4, 1, 1)])
    assert ran >= 6
    op.close()


@pytest.mark.parametrize("N", [16, 24])
def test_maxpreserve_tiles_agree(spb, orc, N):
    """Every instantiated two-product (fused maxPreserve) tile against the oracle's
    three-pass ComputeQ_maxPreserve, in FP64 and with FP32-streamed weights."""
    Lv = 6.0
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    W = op.get_weights()
    rng = np.random.default_rng(N + 1)
    ncell = 3
    f = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(ncell)])
    ref = np.stack([orc.compute_q_maxpreserve(N, v, eta, W, f[b], f[b]).ravel() for b in range(ncell)])
    tiles = {16: [(4, 4, 1, 0), (4, 4, 2, 0), (4, 1, 1, 1), (4, 1, 2, 1)],
             24: [(6, 8, 1, 0), (6, 8, 2, 0), (6, 3, 1, 1), (3, 3, 1, 1), (4, 2, 1, 1), (8, 1, 1, 1), (4, 9, 1, 1),
                  (4, 2, 1, 4), (4, 9, 1, 2), (3, 3, 1, 8)]}[N]
    for t in tiles:
        for form, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
            op.set_tile(form, True, *t)
            Q = op.compute_q(f, form=form, maxpreserve=True)
            for b in range(ncell):
                assert rel_l2(Q[b], ref[b]) < tol, (t, form, b)


# --------------------------------------------------- reference interface
def test_computeq_abi_bkw8(spb, orc):
    """ComputeQ and ComputeQ_maxPreserve through src/collisions.h with the BKW8 table."""
    N, Lv = 8, 5.0
    v, eta = orc.grid_0d(N, Lv)
    W = orc.weights_load(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), N)
    f = orc.bkw_initial(N, v).ravel()
    spb.initialize_coll(N, Lv, v, eta)
    rows = spb.WeightRows(W)
    assert rel_l2(spb.ComputeQ(f, f, rows), orc.compute_q(N, v, eta, W, f, f)) < 1e-12
    assert rel_l2(spb.ComputeQ_maxPreserve(f, f, rows), orc.compute_q_maxpreserve(N, v, eta, W, f, f)) < 1e-12
    # two different arguments: the reference subtracts f's Maxwellian from g (src/collisions.c:104)
    g = orc.maxwellian(v, 0.9, np.array([0.3, 0, -0.2]), 1.2).ravel()
    assert rel_l2(spb.ComputeQ(f, g, rows), orc.compute_q(N, v, eta, W, f, g)) < 1e-12
    assert rel_l2(spb.ComputeQ_maxPreserve(f, g, rows), orc.compute_q_maxpreserve(N, v, eta, W, f, g)) < 1e-12
    spb.dealloc_coll()


def test_computeq_abi_sees_reinitialized_weights(spb, orc, tmp_path):
    """ComputeQ keeps a device copy of the table keyed by the conv_weights
    pointer; initialize_weights refilling the same pointer (src/weights.c:33-125)
    must replace that copy, so the next ComputeQ uses the new rows."""
    import ctypes
    import shutil
    N, Lv = 8, 5.0
    v, eta = orc.grid_0d(N, Lv)
    (tmp_path / "Weights").mkdir()
    shutil.copy(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), tmp_path / "Weights")
    L = spb.lib()
    L.initialize_weights.restype = None
    L.initialize_weights.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p, spb.Species, spb.Species]
    cw = ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))()
    L.alloc_weights(N, ctypes.byref(cw), 1)
    eta_c = np.ascontiguousarray(eta)
    dp = ctypes.POINTER(ctypes.c_double)
    sp = spb.default_species()
    f = np.ascontiguousarray(orc.bkw_initial(N, v).ravel())
    spb.initialize_coll(N, Lv, v, eta)

    def compute_q():
        Q = np.zeros_like(f)
        L.ComputeQ(f.ctypes.data_as(dp), f.ctypes.data_as(dp), Q.ctypes.data_as(dp), cw[0])
        return Q

    def table():
        return np.stack([np.ctypeslib.as_array(cw[0][r], shape=(N ** 3,)).copy() for r in range(N ** 3)])

    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        L.initialize_weights(N, eta_c.ctypes.data_as(dp), Lv, 0.0, 0, 0, ctypes.cast(cw[0], ctypes.c_void_p), sp, sp)
        W0 = table()
        Q0 = compute_q()
        assert rel_l2(Q0, orc.compute_q(N, v, eta, W0, f, f)) < 1e-12
        # regenerate lambda = 1 into the same pointer (weightFlag 1: fresh table)
        L.initialize_weights(N, eta_c.ctypes.data_as(dp), Lv, 1.0, 1, 0, ctypes.cast(cw[0], ctypes.c_void_p), sp, sp)
    finally:
        os.chdir(cwd)
    W1 = table()
    assert rel_l2(W1, W0) > 1e-3
    Q1 = compute_q()
    assert rel_l2(Q1, orc.compute_q(N, v, eta, W1, f, f)) < 1e-12
    assert rel_l2(Q1, Q0) > 1e-3
    L.dealloc_weights(N, cw[0])
    spb.dealloc_coll()


def test_conserve_abi(spb, orc):
    """conserveAllMoments (src/conserve.h:16), one and two species."""
    N, Lv = 8, 9.0
    v, eta = orc.grid_1d(N, Lv)
    rng = np.random.default_rng(2)
    Q = rng.standard_normal(N ** 3)
    spb.initialize_conservation(N, v[1] - v[0], v)
    (P,) = spb.conserveAllMoments([Q])
    assert rel_l2(P, orc.Conserver(N, v)(Q)) < 1e-12
    Cq, scale = _cmoments(N, v, P)
    assert np.abs(Cq).max() < 1e-13 * scale.max()
    spb.dealloc_conservation()
    masses = (1.0, 2.0)
    Q4 = rng.standard_normal((4, N ** 3))
    spb.initialize_conservation(N, v[1] - v[0], v, masses)
    P4 = np.array(spb.conserveAllMoments(list(Q4)))
    assert rel_l2(P4, orc.Conserver(N, v, masses)(Q4)) < 1e-12
    spb.dealloc_conservation()


def test_bkw8_run_abi(spb, orc):
    """The BKW8 test of the reference (tests/BKW8) driven through the GPU library:
    100 Heun steps of ComputeQ_maxPreserve + conserveAllMoments."""
    N, Lv, dt = 8, 5.0, 0.01
    v, eta = orc.grid_0d(N, Lv)
    W = orc.weights_load(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), N)
    spb.initialize_coll(N, Lv, v, eta)
    spb.initialize_conservation(N, v[1] - v[0], v)
    rows = spb.WeightRows(W)
    f = orc.bkw_initial(N, v).ravel()
    out = [orc.moments_row_0d(N, v, f, 0.0)]
    for t in range(100):
        (Q,) = spb.conserveAllMoments([spb.ComputeQ_maxPreserve(f, f, rows)])
        f1 = f + dt * Q
        (Q,) = spb.conserveAllMoments([spb.ComputeQ_maxPreserve(f1, f1, rows)])
        f = 0.5 * (f + f1) + 0.5 * dt * Q
        out.append(orc.moments_row_0d(N, v, f, dt * (t + 1)))
    spb.dealloc_coll()
    spb.dealloc_conservation()
    out = np.array(out)
    ref = np.loadtxt(os.path.join(GOLDEN, "BKW8", "moments_BKW8.test.in"), comments="#")
    m = np.abs(ref) > 1e-12
    assert (np.abs(out - ref)[m] / np.abs(ref)[m]).max() < 5e-7
    orc_rows = np.array(orc.run_0d(N, v, eta, W, orc.bkw_initial(N, v), dt, 100, order=2))
    m = np.abs(orc_rows) > 1e-12
    assert (np.abs(out - orc_rows)[m] / np.abs(orc_rows)[m]).max() < 1e-10


def test_heat_run_batched(spb, orc):
    """tests/heat_transport of the reference: the collision step of all 250 cells in
    one batched call (ComputeQ + projection), transport by the oracle, 10 steps."""
    N, Lv, nX, dt, Kn = 8, 9.0, 250, 1e-4, 3.2
    v, eta = orc.grid_1d(N, Lv)
    W = orc.weights_load(os.path.join(GOLDEN, "heat_transport", "N8_isotropic_L_v9_lambda1.wts"), N)
    op = spb.Operator(N, Lv, v, eta)
    op.set_weights(W)
    x, dx = orc.uniform_mesh(nX, 1.0, 1)
    f = np.zeros((nX + 2, N ** 3))
    f[1:nX + 1] = orc.maxwellian_1d(v, 1.0, 0.0, 1.5).ravel()
    snaps = []
    for t in range(10):
        f, fc = orc.upwind_one(N, nX, v, dx, dt, 1, f)
        Q = op.compute_q(fc[1:nX + 1], conserve=True)
        f[1:nX + 1] = fc[1:nX + 1] + dt * Q / Kn
        snap = np.zeros((nX, 4))
        for l in range(nX):
            c = f[1 + l]
            rho = orc.density(N, v, c); u = orc.bulk_velocity(N, v, c, rho); T = orc.temperature(N, v, c, u, rho)
            snap[l] = (rho, u[0], T, rho * T)
        snaps.append(snap)
    ref = np.loadtxt(os.path.join(GOLDEN, "heat_transport", "moments_heat_transport.test.in"),
                     comments="#").reshape(11, nX, 6)[1:]
    snaps = np.array(snaps)
    for c in (0, 2, 3):
        assert (np.abs(ref[:, :, 2 + c] - snaps[:, :, c]) / np.abs(ref[:, :, 2 + c])).max() < 5e-7


# ------------------------------------------------------------ batched path
@pytest.mark.parametrize("N,Lv,lam,ncell", [(16, 5.0, 0.0, 5), (24, 6.0, 1.0, 3), (6, 4.0, 1.0, 2), (9, 5.0, 1.0, 2),
                                          (12, 5.0, 0.0, 3)])
def test_batched_compute_q_fp64(spb, orc, N, Lv, lam, ncell):
    """Batched ComputeQ / maxPreserve (+ projection) on physical inputs vs the oracle,
    W generated on the GPU and checked against the oracle's table.  Up to N=16
    the forward transform sums the Maxwellians' moments itself and the inverse
    transform's last pass runs over whole cubes with the projection (N = 6, 9,
    12, 16: radix 2, 3 and 4 stages in registers)."""
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(lam)
    W = op.get_weights()
    Wo = orc.weights_iso(N, eta, Lv, lam)
    assert np.abs(W - Wo).max() < 1e-13 * np.abs(Wo).max()
    rng = np.random.default_rng(N)
    f = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(ncell)])
    g = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(ncell)])
    Q = op.compute_q(f)
    Qg = op.compute_q(f, g)
    Qmp = op.compute_q(f, maxpreserve=True, conserve=True)
    Qmpg = op.compute_q(f, g, maxpreserve=True)
    cons = orc.Conserver(N, v)
    for b in range(ncell):
        assert rel_l2(Q[b], orc.compute_q(N, v, eta, W, f[b], f[b])) < 1e-12
        assert rel_l2(Qg[b], orc.compute_q(N, v, eta, W, f[b], g[b])) < 1e-12
        ref = cons(orc.compute_q_maxpreserve(N, v, eta, W, f[b], f[b]).ravel())
        assert rel_l2(Qmp[b], ref) < 1e-12
        assert rel_l2(Qmpg[b], orc.compute_q_maxpreserve(N, v, eta, W, f[b], g[b])) < 1e-12
    Cq, scale = _cmoments(N, v, Qmp)
    assert np.abs(Cq).max() < 1e-13 * scale.max()


def test_weights_generation_matches_fixtures(spb, orc):
    """generate_conv_weights_iso on the GPU reproduces both reference tables."""
    for kind, Lv, lam, path in [("0d", 5.0, 0.0, "BKW8/N8_isotropic_L_v5_lambda0.wts"),
                                ("1d", 9.0, 1.0, "heat_transport/N8_isotropic_L_v9_lambda1.wts")]:
        v, eta = _grid(orc, 8, Lv, kind)
        op = spb.Operator(8, Lv, v, eta)
        op.generate_weights_iso(lam)
        Wf = orc.weights_load(os.path.join(GOLDEN, path), 8)
        assert np.abs(op.get_weights() - Wf).max() <= 2e-15 * np.abs(Wf).max()
    v, eta = orc.grid_1d(8, 6.0)
    op = spb.Operator(8, 6.0, v, eta)
    op.generate_weights_iso(0.5)
    Wo = orc.weights_iso(8, eta, 6.0, 0.5)
    assert np.abs(op.get_weights() - Wo).max() < 1e-12 * np.abs(Wo).max()


@pytest.mark.parametrize("N,Lv", [(16, 5.0), (24, 6.0)])
def test_reduced_precision_forms(spb, orc, N, Lv):
    """FP32 / FP16x2 streamed weights (FP64 or FP32 arithmetic): <= 1e-6 relative L2
    vs FP64, projection exact."""
    v, eta = orc.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    rng = np.random.default_rng(11)
    f = np.stack([_two_maxwellians(orc, v, rng).ravel() for _ in range(4)])
    ref = op.compute_q(f, conserve=True)
    for form in ("fp32", "fp16x2", "fp32a"):
        for mp in (False, True):
            q = op.compute_q(f, form=form, maxpreserve=mp, conserve=True)
            r = op.compute_q(f, maxpreserve=mp, conserve=True) if mp else ref
            err = max(rel_l2(q[b], r[b]) for b in range(4))
            assert err < 1e-6, (form, mp, err)
            Cq, scale = _cmoments(N, v, q)
            assert np.abs(Cq).max() < 1e-13 * scale.max()


def test_batch_beyond_one_pass_and_empty_batch(spb, orc):
    """A batch larger than one pass (4096 cells at N=16: the cap of the
    operator's per-pass scratch) runs in chunks: every cell, on both sides of
    the chunk boundary, equals the same cell evaluated in a small batch with the
    same tile, bit for bit; an empty batch is a no-op that touches no pointer."""
    N, Lv = 16, 5.0
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    op.set_tile("fp64", False, 4, 4, 1, 0)
    rng = np.random.default_rng(4101)
    ncell = 4096 + 5
    V = np.stack(np.meshgrid(v, v, v, indexing="ij")).reshape(3, -1)
    u = rng.uniform(-1, 1, (ncell, 3))
    T = rng.uniform(0.6, 1.6, ncell)
    d2 = ((V[None, :, :] - u[:, :, None]) ** 2).sum(1)
    f = np.exp(-d2 / (2 * T[:, None])) / (2 * np.pi * T[:, None]) ** 1.5
    Q = op.compute_q(f)
    idx = [0, 1, 4095, 4096, 4097, ncell - 1]
    assert np.array_equal(Q[idx], op.compute_q(f[idx]))
    op.compute_q_dev(None, None, None, 0)
    op.close()


def test_n32_single_cell_against_oracle(spb, orc):
    """The headline grid (N=32, L_v=8, hard spheres, 1D grid): one cell, FP64 through
    the batched path against the oracle's full O(N^6) evaluation; FP32 forms <= 1e-6."""
    N, Lv = 32, 8.0
    v, eta = orc.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    rng = np.random.default_rng(32)
    f = _two_maxwellians(orc, v, rng).ravel()
    q64 = op.compute_q(f[None])[0]
    W = op.get_weights()
    ref = orc.compute_q(N, v, eta, W, f, f)
    del W
    assert rel_l2(q64, ref) < 1e-12
    for form in ("fp32", "fp16x2", "fp32a"):
        err = rel_l2(op.compute_q(f[None], form=form)[0], q64)
        print(form, err)
        assert err < 1e-6, (form, err)


@pytest.mark.parametrize("N,Lv,lam,rows", [(8, 5.0, 1.0, list(range(0, 512, 37))), (12, 4.0, 0.0, [0, 871, 1727])])
def test_weights_landau_matches_oracle(spb, orc, N, Lv, lam, rows):
    """The glance = 0 (Landau) table generated on the GPU against the oracle's
    gHat3L rows (1e-12 of the row's max |W|: the same 64-point rules, GPU
    Bessel J0 against the C library's), and its xi = 0 column against the
    closed form -(8 pi / 3) |zeta|^2 L^(lam+5) / (lam+5) for every row."""
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_landau(lam)
    W = op.get_weights()
    for r in rows:
        ref = orc.weights_landau(N, eta, Lv, lam, rows=(r, 1))[0]
        assert np.abs(W[r] - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300), r
    h = N // 2
    c0 = h + N * (h + N * h)
    assert eta[h] == 0.0
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    z2 = (eta[I] ** 2 + eta[J] ** 2 + eta[K] ** 2).ravel()
    want = -(8 * np.pi / 3) * z2 * Lv ** (lam + 5) / (lam + 5)
    assert np.abs(W[:, c0] - want).max() <= 1e-13 * np.abs(want).max()
    op.close()


def test_initialize_weights_aniso_generates_landau_rows(spb, orc, tmp_path):
    """initialize_weights_AnIso with glance 0 and no file generates the Landau
    table on the GPU into the caller's rows (src/aniso_weights.c:82-106) and
    stores it under the reference's name."""
    import ctypes
    N, Lv, lam = 8, 5.0, 0.0
    v, eta = orc.grid_0d(N, Lv)
    (tmp_path / "Weights").mkdir()
    L = spb.lib()
    cw = ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))()
    L.alloc_weights(N, ctypes.byref(cw), 1)
    L.initialize_weights_AnIso.restype = None
    L.initialize_weights_AnIso.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]
    eta_c = np.ascontiguousarray(eta)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        L.initialize_weights_AnIso(N, eta_c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), Lv, lam, 0,
                                   ctypes.cast(cw[0], ctypes.c_void_p), 0.0)
    finally:
        os.chdir(cwd)
    rows = cw[0]
    stored = np.fromfile(tmp_path / "Weights" / "N8_Aniso_L_v5_lambda0_Landau.wts").reshape(512, 512)
    for r in (5, 300):
        ref = orc.weights_landau(N, eta, Lv, lam, rows=(r, 1))[0]
        got = np.ctypeslib.as_array(rows[r], shape=(512,))
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.array_equal(stored[r], got)
    L.dealloc_weights(N, rows)
