"""The boltz_ front end end to end on the GPU (spbte_boltz_main through the
C ABI): the reference's own test inputs (tests/BKW8, tests/heat_transport) run
in a directory laid out as boltz_ expects, against their moment fixtures; the
weight table generated and stored when Weights/ is empty; restart dumps and a
restarted run; and the walls and body force each Init_field selects, against
the oracle's run.

Tolerances: 5e-7 relative against the fixtures (their print precision, as in
the other output tests); 1e-6 relative against the oracle's run printed
through the same %le columns; weight tables 2e-15 of max|W| (the closed forms
evaluated in a different order than GSL's)."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FLAGS = os.path.join(GOLDEN, "heat_transport", "heat_transport.test.out")


def _workdir(base, name, inputs, weights=()):
    d = base / name
    for sub in ("input", "Data", "Weights", "Restart"):
        (d / sub).mkdir(parents=True)
    for fname, content in inputs.items():
        if isinstance(content, str) and os.path.exists(content):
            shutil.copy(content, d / "input" / fname)
        else:
            (d / "input" / fname).write_text(content)
    for w in weights:
        shutil.copy(w, d / "Weights" / os.path.basename(w))
    return d


def _edit(path, **values):
    """The reference input file at path with the value after each named keyword replaced."""
    lines = open(path).read().split("\n")
    for i, ln in enumerate(lines):
        if ln.strip() in values:
            lines[i + 1] = str(values.pop(ln.strip()))
    assert not values, values
    return "\n".join(lines)


def _input_text(**kw):
    return "".join(f"{k}\n{v}\n" for k, v in kw.items()) + "num_species\n1\ndefault\nStop\n"


def _rows(path):
    return np.array([[float(t) for t in ln.split()] for ln in open(path) if ln.strip() and not ln.startswith("#")])


def _headers(path):
    return [ln for ln in open(path) if ln.startswith("#")]


@pytest.mark.parametrize("weights", ["fixture", "generated"])
def test_boltz_bkw8_matches_fixture(spb, tmp_path, weights):
    """tests/BKW8 end to end: 0D BKW, maxPreserve + projection + Heun, 100 steps.
    With an empty Weights/ the table is generated on the GPU and stored under the
    reference's file name."""
    wts = os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts")
    d = _workdir(tmp_path, weights, {"BKW8.test.in": os.path.join(GOLDEN, "BKW8", "BKW8.test.in"),
                                     "BKW8.test.out": os.path.join(GOLDEN, "BKW8", "BKW8.test.out")},
                 weights=[wts] if weights == "fixture" else [])
    spb.boltz("BKW8.test.in", "BKW8.test.out", workdir=d)
    out, ref = d / "Data" / "moments_BKW8.test.in", os.path.join(GOLDEN, "BKW8", "moments_BKW8.test.in")
    assert _headers(out) == _headers(ref)
    got, want = _rows(out), _rows(ref)
    assert got.shape == want.shape == (101, 5 + 1 + 8)
    m = np.abs(want) > 1e-12
    assert (np.abs(got - want)[m] / np.abs(want)[m]).max() < 5e-7
    stored = np.fromfile(d / "Weights" / os.path.basename(wts))
    fixture = np.fromfile(wts)
    assert np.abs(stored - fixture).max() <= 2e-15 * np.abs(fixture).max()


def test_boltz_heat_transport_matches_fixture(spb, tmp_path):
    """tests/heat_transport end to end: mesh file, 1D order 1 with diffuse walls
    at T = 1 and 2 (Init_field 3), ComputeQ + projection, 10 steps."""
    g = os.path.join(GOLDEN, "heat_transport")
    d = _workdir(tmp_path, "heat", {"heat_transport.test.in": os.path.join(g, "heat_transport.test.in"),
                                    "heat_transport.test.out": os.path.join(g, "heat_transport.test.out"),
                                    "heat_transport.test.mesh": os.path.join(g, "heat_transport.test.mesh")},
                 weights=[os.path.join(g, "N8_isotropic_L_v9_lambda1.wts")])
    spb.boltz("heat_transport.test.in", "heat_transport.test.out", workdir=d)
    out, ref = d / "Data" / "moments_heat_transport.test.in", os.path.join(g, "moments_heat_transport.test.in")
    assert _headers(out) == _headers(ref)
    got, want = _rows(out), _rows(ref)
    assert got.shape == want.shape == (11 * 250, 6)
    for c in (0, 1, 2, 4, 5):  # time, position, density, temperature, pressure
        assert (np.abs(got[:, c] - want[:, c]) / np.maximum(np.abs(want[:, c]), 1e-30)).max() < 5e-7, c
    assert np.abs(got[:, 3] - want[:, 3]).max() < 5e-7 * np.abs(want[:, 3]).max()


def test_boltz_restart_dump_and_resume(spb, tmp_path):
    """Restart_time reached at the first output: the interior pdf and the step
    index are dumped (src/restart.c:24-63) and the run ends; Restart 1 reloads
    them (:65-109), appends to the moment file with a new header and continues.
    As in exec/boltz.c:360-388 the stored index is that of the step just taken,
    so the resumed run repeats one step: its output at step k is the
    uninterrupted run's output at step k + 1."""
    g = os.path.join(GOLDEN, "heat_transport")
    src = os.path.join(g, "heat_transport.test.in")
    common = {"heat_transport.test.out": os.path.join(g, "heat_transport.test.out"),
              "heat_transport.test.mesh": os.path.join(g, "heat_transport.test.mesh")}
    wts = [os.path.join(g, "N8_isotropic_L_v9_lambda1.wts")]
    d = _workdir(tmp_path, "cut", dict(common, **{"heat_transport.test.in": _edit(
        src, Number_of_time_steps=4, Restart_time=1e-9)}), weights=wts)
    spb.boltz("heat_transport.test.in", "heat_transport.test.out", workdir=d)
    first = _rows(d / "Data" / "moments_heat_transport.test.in")
    assert first.shape == (250, 6) and np.all(first[:, 0] == 0.0)
    pdf = d / "Restart" / "heat_transport.test.in_rank0_default.plt"
    assert os.path.getsize(pdf) == 250 * 8 ** 3 * 8
    assert np.fromfile(d / "Restart" / "heat_transport.test.in_time.plt", dtype=np.int32).tolist() == [0]
    (d / "input" / "heat_transport.test.in").write_text(_edit(src, Number_of_time_steps=4, Restart=1))
    spb.boltz("heat_transport.test.in", "heat_transport.test.out", workdir=d)
    out = d / "Data" / "moments_heat_transport.test.in"
    assert len(_headers(out)) == 2
    resumed = _rows(out)[250:].reshape(4, 250, 6)
    s = _workdir(tmp_path, "straight", dict(common, **{"heat_transport.test.in": _edit(src, Number_of_time_steps=5)}),
                 weights=wts)
    spb.boltz("heat_transport.test.in", "heat_transport.test.out", workdir=s)
    straight = _rows(s / "Data" / "moments_heat_transport.test.in").reshape(6, 250, 6)
    for k in range(1, 5):
        assert resumed[k - 1][:, 0] == pytest.approx(1e-4 * k)
        assert np.array_equal(resumed[k - 1][:, 1:], straight[k + 1][:, 1:]), k


# Init_field -> (order-1 walls, order-2 walls, T0 of the left wall, body force Ma)
# as upwindOne / upwindTwo choose them (src/transportroutines.c:117-172, :359-404,
# :176-179): 0 copy / no-flux, 1 diffuse both, 2 periodic, 3 diffuse left only.
WALLS = {0: (0, 0, 1.0, 0.0), 1: (3, 3, 2.0, 0.0), 2: (0, 0, 1.0, 0.0), 3: (1, 1, 1.0, 0.0), 5: (None, 1, 1.0, 1.0),
         6: (2, 0, 1.0, 0.0)}


@pytest.mark.parametrize("init,order", [(1, 1), (1, 2), (5, 2), (6, 1), (6, 2), (0, 2), (2, 1)])
def test_boltz_init_fields_vs_oracle(spb, orc, tmp_path, init, order):
    """Each 1D Init_field through the front end (initial state, walls, body
    force, generated weights) against the oracle's run of the same state."""
    N, Lv, nX, dt, Kn, nT = 8, 5.0, 8, 0.002, 0.5, 3
    text = _input_text(N=N, L_v=Lv, Knudsen=Kn, Lambda=1.0, Time_step=dt, Number_of_time_steps=nT,
                       Space_order=order, Data_writing_frequency=1, Restart=0, Restart_time=0, Init_field=init,
                       SpaceInhom=1, Recompute_weights=0, Anisotropic=0, mesh_file="m.mesh")
    d = _workdir(tmp_path, f"i{init}o{order}", {"run.in": text, "run.out": FLAGS, "m.mesh": f"%m\n{nX}\n1\n{nX}\n1.0\n"})
    spb.boltz("run.in", "run.out", workdir=d)
    got = _rows(d / "Data" / "moments_run.in").reshape(nT + 1, nX, 6)
    v, eta = orc.grid_1d(N, Lv)
    W = orc.weights_load(str(d / "Weights" / "N8_isotropic_L_v5_lambda1.wts"), N)
    inp = spb.read_input(d / "input" / "run.in")
    f0 = spb.initial_state(inp, spb.default_species(), v, nX)
    x, dx = orc.uniform_mesh(nX, 1.0, order)
    b1, b2, T0, Ma = WALLS[init]
    snaps = orc.run_1d(N, v, eta, W, f0, x, dx, dt, nT, order, Kn, b1 if order == 1 else b2, T0=T0, T1=2.0, Ma=Ma)
    for t in range(nT + 1):
        assert np.allclose(got[t][:, 1], x[order:order + nX], rtol=1e-6, atol=0)
        ref = np.asarray(snaps[t])  # rho, u_x, T, p
        for c in range(4):
            assert np.all(np.abs(got[t][:, 2 + c] - ref[:, c]) <= 1e-6 * np.abs(ref[:, c]) + 1e-12), (t, c)


def test_boltz_refuses_what_the_reference_cannot_do(spb, tmp_path):
    """Poiseuille at order 1 (the reference's force is written outside the
    cell) and Init_field 4 in 1D (no state) end with an error, not a run."""
    for init, order, msg in [(5, 1, "Space_order 2"), (4, 1, "Init_field 4")]:
        text = _input_text(N=8, L_v=5.0, Time_step=0.001, Number_of_time_steps=1, Space_order=order,
                           Data_writing_frequency=1, Init_field=init, SpaceInhom=1, mesh_file="m.mesh")
        d = _workdir(tmp_path, f"bad{init}", {"run.in": text, "run.out": FLAGS, "m.mesh": "%m\n4\n1\n4\n1.0\n"})
        with pytest.raises(spb.SpbteError, match=msg):
            spb.boltz("run.in", "run.out", workdir=d)


def test_boltz_landau_table_generated_and_used(spb, orc, tmp_path):
    """Anisotropic 1 with an empty Weights/: the Landau table (glance 0, as
    exec/boltz.c:151-153 asks for) is generated on the GPU, stored under the
    reference's name, and the 0D run with it matches the oracle's run with the
    stored table; weightgen_amd (exec/MPIWeightGenerator.c) writes the same
    file bit for bit."""
    N, Lv, dt, nT = 8, 5.0, 1e-6, 3
    text = _input_text(N=N, L_v=Lv, Knudsen=1.0, Lambda=1.0, Time_step=dt, Number_of_time_steps=nT, Space_order=2,
                       Data_writing_frequency=1, Init_field=2, SpaceInhom=0, Anisotropic=1)
    d = _workdir(tmp_path, "landau", {"run.in": text, "run.out": os.path.join(GOLDEN, "BKW8", "BKW8.test.out")})
    spb.boltz("run.in", "run.out", workdir=d)
    name = "N8_Aniso_L_v5_lambda1_Landau.wts"
    W = orc.weights_load(str(d / "Weights" / name), N)
    v, eta = orc.grid_0d(N, Lv)
    for r in (3, 200, 511):
        ref = orc.weights_landau(N, eta, Lv, 1.0, rows=(r, 1))[0]
        assert np.abs(W[r] - ref).max() <= 1e-12 * np.abs(ref).max()
    ref_rows = np.array(orc.run_0d(N, v, eta, W, orc.bkw_initial(N, v).ravel(), dt, nT, order=2))
    got = _rows(d / "Data" / "moments_run.in")
    assert got.shape == ref_rows.shape
    m = np.abs(ref_rows) > 1e-12
    assert (np.abs(got - ref_rows)[m] / np.abs(ref_rows)[m]).max() < 1e-6
    g = tmp_path / "gen"
    (g / "Weights").mkdir(parents=True)
    spb.weightgen(N, 0, 1, 1.0, Lv, workdir=g)
    assert np.array_equal(np.fromfile(g / "Weights" / name), np.fromfile(d / "Weights" / name))
    spb.weightgen(N, 0, 1, 1.0, Lv, workdir=g)  # an existing table is kept


@pytest.mark.parametrize("N", [9, 12])
def test_boltz_other_grids_vs_oracle(spb, orc, tmp_path, N):
    """The front end on grids other than the fixtures' (odd N uses the 0D grid's
    0.5 (N-1) h_eta half-width, src/initializer.c:74-78): a BKW run with the
    isotropic Maxwell-molecule table generated on the GPU against the oracle's
    run with the stored table."""
    Lv, dt, nT = 5.0, 0.01, 3
    text = _input_text(N=N, L_v=Lv, Knudsen=1.0, Lambda=0.0, Time_step=dt, Number_of_time_steps=nT, Space_order=2,
                       Data_writing_frequency=1, Init_field=2, SpaceInhom=0)
    d = _workdir(tmp_path, f"n{N}", {"run.in": text, "run.out": os.path.join(GOLDEN, "BKW8", "BKW8.test.out")})
    spb.boltz("run.in", "run.out", workdir=d)
    W = orc.weights_load(str(d / "Weights" / f"N{N}_isotropic_L_v5_lambda0.wts"), N)
    v, eta = orc.grid_0d(N, Lv)
    assert np.abs(W - orc.weights_iso(N, eta, Lv, 0.0)).max() <= 1e-13 * np.abs(W).max()
    ref = np.array(orc.run_0d(N, v, eta, W, orc.bkw_initial(N, v).ravel(), dt, nT, order=2))
    got = _rows(d / "Data" / "moments_run.in")
    assert got.shape == ref.shape
    m = np.abs(ref) > 1e-12
    assert (np.abs(got - ref)[m] / np.abs(ref)[m]).max() < 1e-6
