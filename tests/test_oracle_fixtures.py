"""The CPU oracle against the golden fixtures of the reference's own tests.

tests/BKW8 and tests/heat_transport of the reference hold two weight tables
(compared byte-for-byte by tests/run_test.sh:40-52) and two moment histories
(printed with %le, 7 significant digits, src/output.c:184-206,322-340)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_weights_lambda0_fixture(orc):
    # tests/BKW8/input/BKW8.test.in: N=8, L_v=5, lambda=0, 0D grid (src/initializer.c:66-82)
    v, eta = orc.grid_0d(8, 5.0)
    W = orc.weights_iso(8, eta, 5.0, 0.0)
    Wf = orc.weights_load(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), 8)
    assert np.abs(W - Wf).max() <= 2e-15 * np.abs(Wf).max()


def test_weights_lambda1_fixture(orc):
    # tests/heat_transport/input/heat_transport.test.in: N=8, L_v=9, lambda=1, 1D grid (:256-266)
    v, eta = orc.grid_1d(8, 9.0)
    W = orc.weights_iso(8, eta, 9.0, 1.0)
    Wf = orc.weights_load(os.path.join(GOLDEN, "heat_transport", "N8_isotropic_L_v9_lambda1.wts"), 8)
    assert np.abs(W - Wf).max() <= 2e-15 * np.abs(Wf).max()


def test_weights_file_layout(orc, tmp_path):
    # src/weights.c:78-88: raw FP64 rows; a short file is an error (:83-86)
    src = os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts")
    raw = np.fromfile(src, dtype="<f8")
    assert raw.size == 512 * 512
    short = tmp_path / "short.wts"
    raw[:-1].tofile(short)
    with pytest.raises(IOError):
        orc.weights_load(short, 8)
    with pytest.raises(IOError):
        orc.weights_load(tmp_path / "missing.wts", 8)


def test_bkw8_pipeline_fixture(orc):
    """0D BKW, maxPreserve + projection + Heun, 100 steps (exec/boltz.c:186-250)."""
    v, eta = orc.grid_0d(8, 5.0)
    W = orc.weights_load(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), 8)
    rows = np.array(orc.run_0d(8, v, eta, W, orc.bkw_initial(8, v), 0.01, 100, order=2, Kn=1.0))
    ref = np.loadtxt(os.path.join(GOLDEN, "BKW8", "moments_BKW8.test.in"), comments="#")
    assert rows.shape == ref.shape == (101, 14)
    m = np.abs(ref) > 1e-12
    rel = np.abs(rows - ref)[m] / np.abs(ref)[m]
    assert rel.max() < 5e-7, rel.max()
    # velocity is round-off noise (1e-16) in the fixture: absolute check
    assert np.abs(rows[:, 2]).max() < 1e-14


def test_heat_pipeline_fixture(orc):
    """1D heat transfer, order-1 upwind + diffuse walls + ComputeQ + projection + Euler
    (exec/boltz.c:254-345), 250 cells x 10 steps."""
    N, Lv, nX = 8, 9.0, 250
    v, eta = orc.grid_1d(N, Lv)
    W = orc.weights_load(os.path.join(GOLDEN, "heat_transport", "N8_isotropic_L_v9_lambda1.wts"), N)
    x, dx = orc.uniform_mesh(nX, 1.0, 1)
    f0 = np.tile(orc.maxwellian_1d(v, 1.0, 0.0, 1.5).reshape(1, -1), (nX, 1))
    snaps = np.array(orc.run_1d(N, v, eta, W, f0, x, dx, 1e-4, 10, 1, 3.2, bc=1))
    ref = np.loadtxt(os.path.join(GOLDEN, "heat_transport", "moments_heat_transport.test.in"),
                     comments="#").reshape(11, nX, 6)
    assert np.abs(ref[0, :, 1] - x[1:nX + 1]).max() < 1e-15
    for c in (0, 2, 3):  # rho, T, p
        r = ref[:, :, 2 + c]
        assert (np.abs(r - snaps[:, :, c]) / np.abs(r)).max() < 5e-7
    r = ref[:, :, 3]  # u_x: compare where it is above round-off
    m = np.abs(r) > 1e-8
    assert (np.abs(r - snaps[:, :, 1])[m] / np.abs(r)[m]).max() < 5e-7
    assert np.abs(r - snaps[:, :, 1]).max() < 5e-9
