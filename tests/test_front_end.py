"""The boltz_ front end's host side (include/spbte_boltz.h), no GPU: the input,
mesh and species readers against the reference's own input files and
src/input.c / src/mesh_setup.c / src/species.c semantics, the grids against the
oracle, and every Init_field state against closed-form moments."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

BKW_IN = os.path.join(GOLDEN, "BKW8", "BKW8.test.in")
HEAT_IN = os.path.join(GOLDEN, "heat_transport", "heat_transport.test.in")
HEAT_MESH = os.path.join(GOLDEN, "heat_transport", "heat_transport.test.mesh")


def test_read_input_reference_files(spb):
    a = spb.read_input(BKW_IN)
    assert (a.N, a.L_v, a.Kn, a.lambda_, a.dt, a.nT, a.order, a.data_freq) == (8, 5.0, 1.0, 0.0, 0.01, 100, 2, 1)
    assert (a.restart, a.init_flag, a.inhom, a.weight_flag, a.aniso, a.num_species) == (0, 2, 0, 0, 0, 1)
    assert a.species_names[0].value == b"default" and a.mesh_file == b"not set"
    b = spb.read_input(HEAT_IN)
    assert (b.N, b.L_v, b.Kn, b.lambda_, b.dt, b.nT, b.order, b.data_freq) == (8, 9.0, 3.2, 1.0, 1e-4, 10, 1, 1)
    assert (b.restart_time, b.init_flag, b.inhom, b.mesh_file) == (0.0, 3, 1, b"heat_transport.test.mesh")


def test_read_input_defaults_and_errors(spb, tmp_path):
    """Missing keywords keep set_default_values (src/input.c:132-183); comment
    words are skipped; a file without Stop, or a 1D run without a mesh, fails."""
    p = tmp_path / "min.in"
    p.write_text("# a comment with words\nInit_field\n5\nStop\n")
    a = spb.read_input(p)
    assert (a.N, a.L_v, a.Kn, a.lambda_, a.dt, a.nT, a.order, a.data_freq) == (16, 5.0, 1.0, 1.0, -1.0, 1000, 1, 10)
    assert (a.restart, a.restart_time, a.inhom, a.init_flag, a.num_species) == (0, 85500.0, 0, 5, 1)
    (tmp_path / "nostop.in").write_text("N\n8\n")
    with pytest.raises(spb.SpbteError, match="Stop"):
        spb.read_input(tmp_path / "nostop.in")
    (tmp_path / "nomesh.in").write_text("SpaceInhom\n1\nStop\n")
    with pytest.raises(spb.SpbteError, match="mesh"):
        spb.read_input(tmp_path / "nomesh.in")
    with pytest.raises(spb.SpbteError, match="not found"):
        spb.read_input(tmp_path / "absent.in")


def test_read_mesh_matches_oracle_and_zones(spb, orc, tmp_path):
    for order in (1, 2):
        x, dx, dmin = spb.read_mesh(HEAT_MESH, order)
        xo, dxo = orc.uniform_mesh(250, 1.0, order)
        assert np.array_equal(x, xo) and np.array_equal(dx, dxo) and dmin == 0.004
    # two zones: 4 cells over 1.0 then 2 cells over 2.0
    p = tmp_path / "two.mesh"
    p.write_text("%explanation\n6\n2\n4\n1.0\n2\n2.0\n")
    x, dx, dmin = spb.read_mesh(p, 2)
    assert np.allclose(dx, [0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 1.0, 1.0, 1.0, 1.0])
    assert np.allclose(x, [-0.375, -0.125, 0.125, 0.375, 0.625, 0.875, 1.5, 2.5, 3.5, 4.5]) and dmin == 0.25
    p.write_text("%explanation\n7\n2\n4\n1.0\n2\n2.0\n")
    with pytest.raises(spb.SpbteError, match="zones hold 6"):
        spb.read_mesh(p, 1)


def test_read_species_file(spb, tmp_path):
    """The keyword layout of the reference's input/species/Ar, with its values;
    nb_en_levels is not a keyword the reference reads (it reads num_en_levels)."""
    p = tmp_path / "Ar"
    p.write_text("#---\n# Species file for Ar\nname\nAr\nnb_en_levels\n0\nmass\n6.63e-26\nd_ref\n4.17e-10\n"
                 "T_ref\n273.0\nmu_ref\n2.117e-5\nomega\n0.81\nEn_for\n0.0\nStop\n")
    s = spb.read_species(p)
    assert s.name == b"Ar" and s.mass == 6.63e-26 and s.d_ref == 4.17e-10 and s.T_ref == 273.0
    assert s.mu_ref == 2.117e-5 and s.omega == 0.81 and s.E0 == 0.0 and s.num_levels == 0 and s.Rgas == 1.0


def test_velocity_grids_match_oracle(spb, orc):
    for N, Lv in [(8, 5.0), (9, 4.0), (24, 9.0)]:
        v, eta = spb.velocity_grid(N, Lv, 0)
        vo, eo = orc.grid_0d(N, Lv)
        assert np.array_equal(v, vo) and np.array_equal(eta, eo)
        v, eta = spb.velocity_grid(N, Lv, 1)
        vo, eo = orc.grid_1d(N, Lv)
        assert np.array_equal(v, vo) and np.array_equal(eta, eo)


def _input(spb, N, Lv, init, inhom=0, order=1):
    a = spb.Input()
    a.N, a.L_v, a.init_flag, a.inhom, a.order = N, Lv, init, inhom, order
    return a


def _moments(N, v, f, mass=1.0, kb=1.0):
    """rho, u_x, u_y, u_z, T of src/momentRoutines.c:58-183 with trapezoid weights."""
    w = np.ones(N); w[0] = w[-1] = 0.5
    W3 = ((w[:, None, None] * w[None, :, None] * w[None, None, :]) * (v[1] - v[0]) ** 3).ravel()
    V = [np.repeat(v, N * N), np.tile(np.repeat(v, N), N), np.tile(v, N * N)]
    rho1 = W3 @ f
    u = [W3 @ (Vd * f) / rho1 for Vd in V]
    T1 = W3 @ (sum((Vd - ud) ** 2 for Vd, ud in zip(V, u)) * f) / (3 * rho1)
    return mass * rho1, u[0], u[1], u[2], mass / kb * T1


def test_initial_states_0d_closed_forms(spb, orc):
    """initialize_hom (src/initializer.c:92-199): each state's moments against
    their closed forms (BKW: rho = T = 1; discontinuous Maxwellian: rho = 3/4,
    u_x = -1/(3 sqrt(pi)); perturbed Maxwellian: rho = 1 + 0.1 Im (1-i)^-3/2;
    shell: the radial integral), and BKW against the oracle's state."""
    sp = spb.default_species()
    N, Lv = 32, 5.0
    v, _ = spb.velocity_grid(N, Lv, 0)
    f2 = spb.initial_state(_input(spb, N, Lv, 2), sp, v)
    assert np.abs(f2 - orc.bkw_initial(N, v).ravel()).max() <= 1e-15 * np.abs(f2).max()
    r, ux, uy, uz, T = _moments(N, v, f2)
    # tolerance: the domain truncation of the BKW tails at L_v = 5 (rho 7.5e-9, T 5.9e-8)
    assert abs(r - 1) < 2e-7 and abs(T - 1) < 2e-7 and max(abs(ux), abs(uy), abs(uz)) < 1e-12
    f1 = spb.initial_state(_input(spb, N, Lv, 1), sp, v)
    r, ux, uy, uz, T = _moments(N, v, f1)
    assert abs(r - 0.75) < 1e-9 and abs(ux + 1 / (3 * math.sqrt(math.pi))) < 2e-3 and abs(uy) < 1e-12
    f5 = spb.initial_state(_input(spb, 48, Lv, 5), sp, spb.velocity_grid(48, Lv, 0)[0])
    r = _moments(48, spb.velocity_grid(48, Lv, 0)[0], f5)[0]
    assert abs(r - (1 + 0.1 * ((1 - 1j) ** -1.5).imag)) < 1e-6
    v8 = spb.velocity_grid(N, 8.0, 0)[0]  # exp(-|v|^2/2) needs the wider box
    f4 = spb.initial_state(_input(spb, N, 8.0, 4), sp, v8)
    r, ux, _, _, T = _moments(N, v8, f4)
    assert abs(r - 1) < 1e-9 and abs(T - 1) < 1e-9
    # shell exp(-S (|v| - sigma)^2 / sigma^2) / S^2, S = 10, sigma = 0.3 L_v
    f0 = spb.initial_state(_input(spb, 48, Lv, 0), sp, spb.velocity_grid(48, Lv, 0)[0])
    s, S = 0.3 * Lv, 10.0
    rr = np.linspace(0, 6, 200001)
    g = 4 * np.pi * rr ** 2 * np.exp(-S * (rr - s) ** 2 / s ** 2) / S ** 2
    rho_exact = np.trapezoid(g, rr)
    T_exact = np.trapezoid(g * rr ** 2, rr) / (3 * rho_exact)
    r, ux, _, _, T = _moments(48, spb.velocity_grid(48, Lv, 0)[0], f0)
    assert abs(r / rho_exact - 1) < 1e-6 and abs(T / T_exact - 1) < 1e-6 and abs(ux) < 1e-12
    # two Maxwellians at +-2 sigma along v_x, sigma = pi L_v / 10: symmetric,
    # peaked at v_x = +-2 sigma, density of the box-truncated pair
    f3 = spb.initial_state(_input(spb, 161, Lv, 3), sp, spb.velocity_grid(161, Lv, 0)[0])
    v161 = spb.velocity_grid(161, Lv, 0)[0]
    sg = math.pi * Lv / 10
    r, ux, uy, _, _ = _moments(161, v161, f3)
    e = lambda a: math.erf(a / (sg * math.sqrt(2)))  # noqa: E731
    rho_box = 0.5 * (e(Lv - 2 * sg) + e(Lv + 2 * sg)) * e(Lv) ** 2
    assert abs(r / rho_box - 1) < 2e-3 and abs(ux) < 1e-12 and abs(uy) < 1e-12
    marg = f3.reshape(161, 161, 161)[:, 80, 80]
    assert abs(abs(v161[np.argmax(marg)]) - 2 * sg) < v161[1] - v161[0]
    with pytest.raises(spb.SpbteError, match="Init_field 6"):
        spb.initial_state(_input(spb, 8, Lv, 6), sp, spb.velocity_grid(8, Lv, 0)[0])


def test_initial_states_1d(spb, orc):
    """initialize_inhom (src/initializer.c:300-423): per case the cells'
    density, velocity and temperature (the states use exp(-|v-u|^2/T)/(pi T)^3/2,
    so the moment temperature is T/2); the shock cases put the left state on
    cells whose index counting the ghost cells is below nX/2."""
    sp = spb.default_species()
    N, Lv, nX = 32, 7.0, 10
    v, _ = spb.velocity_grid(N, Lv, 1)
    cases = {0: [(1.0, 0.0, 0.5), (1.0, 0.0, 0.5)], 1: [(1.0, 0.0, 1.0)] * 2, 2: [(1.0, -1.0, 0.5)] * 2,
             3: [(1.0, 0.0, 0.75)] * 2, 5: [(1.0, 0.0, 0.5)] * 2, 6: [(1.0, 1.2972, 0.5), (1.297, 1.0, 0.5975)]}
    for order in (1, 2):
        for init, (left, right) in cases.items():
            f = spb.initial_state(_input(spb, N, Lv, init, 1, order), sp, v, nX)
            for c in range(nX):
                want = left if c + order < nX // 2 else right
                r, ux, uy, _, T = _moments(N, v, f[c])
                assert abs(r - want[0]) < 1e-9 and abs(ux - want[1]) < 1e-9 and abs(T - want[2]) < 1e-9, (init, c)
    f = spb.initial_state(_input(spb, N, Lv, 3, 1, 1), sp, v, nX)
    assert np.abs(f[0] - orc.maxwellian_1d(v, 1.0, 0.0, 1.5).ravel()).max() <= 1e-16
    with pytest.raises(spb.SpbteError, match="Init_field 4"):
        spb.initial_state(_input(spb, N, Lv, 4, 1, 1), sp, v, nX)


def test_species_mass_enters_states_and_kb(spb):
    """Init_field 4 for a named species uses its mass and the SI Boltzmann
    constant (src/initializer.c:52-55, :166): rho = 1 and T = 1 K in the
    reference's moments (mass m, KB 1.380658e-23)."""
    sp = spb.default_species()
    sp.name, sp.mass = b"Ar", 6.63e-26
    kb = 1.380658e-23
    N, Lv = 48, 120.0  # thermal speed (KB T / m)^1/2 = 14.4
    v, _ = spb.velocity_grid(N, Lv, 0)
    f = spb.initial_state(_input(spb, N, Lv, 4), sp, v)
    r, _, _, _, T = _moments(N, v, f, mass=sp.mass, kb=kb)
    assert abs(r - 1) < 1e-9 and abs(T - 1) < 1e-9


def test_boltz_main_errors_before_any_gpu_call(spb, tmp_path):
    """The front end reports a missing input file (src/input.c:24-27) and refuses
    runs it cannot do, before touching the GPU."""
    with pytest.raises(spb.SpbteError, match="input file not found"):
        spb.boltz("absent.in", "absent.out", workdir=tmp_path)
    (tmp_path / "input").mkdir()
    (tmp_path / "input" / "two.in").write_text(
        "Init_field\n2\nTime_step\n0.01\nAnisotropic\n1\nnum_species\n2\ndefault\ndefault\nStop\n")
    with pytest.raises(spb.SpbteError, match="Anisotropic 1 with 2 species"):
        spb.boltz("two.in", "x.out", workdir=tmp_path)
    (tmp_path / "input" / "nodt.in").write_text("Init_field\n2\nStop\n")
    with pytest.raises(spb.SpbteError, match="Time_step"):
        spb.boltz("nodt.in", "x.out", workdir=tmp_path)


def test_boltz_executable_usage():
    """spectralbte_amd/boltz_amd, the drop-in for boltz_, prints its usage
    without arguments (no GPU call)."""
    exe = os.path.join(ROOT, "spectralbte_amd", "boltz_amd")
    assert os.path.exists(exe), "build the library first (make -C spectralbte_amd)"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


def test_weightgen_refuses_bad_glance_and_prints_usage(spb, tmp_path):
    """glance outside (0, pi), where log(sin(glance/2)) of
    exec/MPIcollisionRoutines.c:119 is not negative, is refused before any GPU
    call; without arguments the executable prints its usage."""
    (tmp_path / "Weights").mkdir()
    for g in (4.0, -0.1):
        with pytest.raises(spb.SpbteError, match=f"glance {g:g}"):
            spb.weightgen(8, g, 1, 1.0, 5.0, workdir=tmp_path)
    exe = os.path.join(ROOT, "spectralbte_amd", "weightgen_amd")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
