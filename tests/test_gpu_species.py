"""GPU parity of multi-species runs (exec/boltz.c with Ns > 1) against the
oracle's drivers: pair tables in operator slots m*Ns + l, Q(f_l, f_m) for every
pair, the joint Ns + 4 projection, species-summed updates, species-mass walls;
and the front end with two species files.

Tolerances: states 1e-12 of max|f| against the oracle over a few steps (FP64,
same operation order up to the transform's summation); conservation of each
species' mass and of the total momentum and energy to 1e-12 relative; the
front end's printed moments 1e-6 (print precision)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _pair_tables(spb, op, N, Lv, masses, lam=1.0):
    """Generate the table of every pair (l, m) in slot m*Ns + l on the GPU and
    return them (index m*Ns + l) for the oracle."""
    Ns = len(masses)
    op.set_species(masses)
    Ws = [None] * (Ns * Ns)
    for l in range(Ns):
        for m in range(Ns):
            op.select_table(m * Ns + l)
            op.generate_weights_iso(lam, 2.0, 2.0, masses[l], masses[m])
            Ws[m * Ns + l] = op.get_weights()
    op.select_table(0)
    return Ws


def _moments(N, v, f, masses):
    w = np.ones(N); w[0] = w[-1] = 0.5
    W3 = ((w[:, None, None] * w[None, :, None] * w[None, None, :]) * (v[1] - v[0]) ** 3).ravel()
    V = [np.repeat(v, N * N), np.tile(np.repeat(v, N), N), np.tile(v, N * N)]
    E = 0.5 * sum(Vd ** 2 for Vd in V)
    mass = [m * (W3 @ f[l].T) for l, m in enumerate(masses)]
    mom = sum(m * np.stack([W3 @ (Vd * f[l]).T for Vd in V]) for l, m in enumerate(masses))
    en = sum(m * (W3 @ (E * f[l]).T) for l, m in enumerate(masses))
    return np.array(mass), mom, en


@pytest.mark.parametrize("order", [1, 2])
def test_run0d_two_species_vs_oracle(spb, orc, order):
    N, Lv, dt, steps, masses = 8, 5.0, 0.01, 3, [1.0, 2.0]
    v, eta = orc.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    Ws = _pair_tables(spb, op, N, Lv, masses)
    rng = np.random.default_rng(3)
    f0 = np.stack([np.stack([orc.maxwellian(v, 1.0, rng.uniform(-0.5, 0.5, 3), rng.uniform(0.7, 1.3)).ravel()
                             + orc.maxwellian(v, 0.3, rng.uniform(-1, 1, 3), 0.6).ravel() for _ in range(1)])
                   for _ in masses])  # (Ns, 1, N^3)
    run = spb.Run.ensemble_species(op, f0, order=order, dt=dt)
    ref = orc.run_0d_species(N, v, eta, Ws, f0[:, 0], masses, dt, steps, order=order)
    m0, p0, e0 = _moments(N, v, f0[:, 0], masses)
    for t in range(steps):
        run.step(1)
        g = run.state()[:, 0]
        assert np.abs(g - ref[t]).max() < 1e-12 * np.abs(ref[t]).max(), t
    m1, p1, e1 = _moments(N, v, g, masses)
    assert np.abs(m1 - m0).max() < 1e-12 * m0.max() and abs(e1 - e0) < 1e-12 * e0
    assert np.abs(p1 - p0).max() < 1e-12 * e0
    run.close()
    op.close()


@pytest.mark.parametrize("order,bc", [(1, 1), (2, 0), (2, 1)])
def test_run1d_two_species_vs_oracle(spb, orc, order, bc):
    """Transport per species with its own mass at the diffuse walls, per-cell
    ComputeQ of every pair, joint projection, Euler or Heun; graph path on."""
    N, Lv, nX, dt, Kn, steps, masses = 8, 5.0, 6, 0.002, 0.5, 2, [1.0, 2.0]
    v, eta = orc.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    Ws = _pair_tables(spb, op, N, Lv, masses)
    x, dx = orc.uniform_mesh(nX, 1.0, order)
    f0 = np.stack([np.stack([orc.maxwellian_1d(v, 1.0 + 0.1 * c + 0.05 * l, 0.3 - 0.1 * c, 1.0 + 0.1 * l).ravel()
                             for c in range(nX)]) for l in range(2)])
    run = spb.Run.mesh_species(op, f0, x, dx, order=order, dt=dt, Kn=Kn, bc=bc)
    run.use_graph(True)
    ref = orc.run_1d_species(N, v, eta, Ws, f0, masses, x, dx, dt, steps, order, Kn, bc)
    for t in range(steps):
        run.step(1)
        g = run.state()
        assert np.abs(g - ref[t]).max() < 1e-12 * np.abs(ref[t]).max(), t
    run.close()
    op.close()


def test_boltz_two_species_files(spb, orc, tmp_path):
    """Two named species through spbte_boltz_main: their HS tables generated
    and stored under the reference's names (src/weights.c:70), one moment file
    per species (src/output.c:146-150), against the oracle's two-species run
    (moments with species 0's mass, as src/output.c:171-177 computes them)."""
    N, Lv, dt, nT = 8, 5.0, 0.01, 3
    d = tmp_path / "two"
    for sub in ("input/species", "Data", "Weights"):
        (d / sub).mkdir(parents=True)
    (d / "input" / "species" / "A").write_text("name\nA\nmass\n1.0\nd_ref\n2.0\nStop\n")
    (d / "input" / "species" / "B").write_text("name\nB\nmass\n2.0\nd_ref\n2.0\nStop\n")
    text = "".join(f"{k}\n{val}\n" for k, val in dict(
        N=N, L_v=Lv, Knudsen=1.0, Lambda=1.0, Time_step=dt, Number_of_time_steps=nT, Space_order=2,
        Data_writing_frequency=1, Init_field=2, SpaceInhom=0).items()) + "num_species\n2\nA\nB\nStop\n"
    (d / "input" / "two.in").write_text(text)
    (d / "input" / "two.out").write_text(open(os.path.join(GOLDEN, "BKW8", "BKW8.test.out")).read())
    spb.boltz("two.in", "two.out", workdir=d)
    names = sorted(os.listdir(d / "Weights"))
    assert names == sorted(f"N8_isotropic_L_v5_HS_{a}_{i}_{b}_{j}.wts" for i, a in enumerate("AB")
                           for j, b in enumerate("AB"))
    v, eta = orc.grid_0d(N, Lv)
    Ws = [None] * 4
    for i, a in enumerate("AB"):
        for j, b in enumerate("AB"):
            Ws[j * 2 + i] = orc.weights_load(str(d / "Weights" / f"N8_isotropic_L_v5_HS_{a}_{i}_{b}_{j}.wts"), N)
    f0 = orc.bkw_initial(N, v).ravel()
    states = orc.run_0d_species(N, v, eta, Ws, [f0, f0], [1.0, 2.0], dt, nT, order=2)
    for l, a in enumerate("AB"):
        rows = np.array([[float(t) for t in ln.split()] for ln in open(d / "Data" / f"moments_two.in_{a}")
                         if not ln.startswith("#")])
        assert rows.shape == (nT + 1, 14)
        ref = np.array([orc.moments_row_0d(N, v, f0, 0.0)] +
                       [orc.moments_row_0d(N, v, states[t][l], dt * (t + 1)) for t in range(nT)])
        m = np.abs(ref) > 1e-12
        assert (np.abs(rows - ref)[m] / np.abs(ref)[m]).max() < 1e-6, a


def test_abi_table_cache_holds_a_three_species_step(spb, orc, capfd):
    """A three-species mixture walks its 9 pair tables every step
    (exec/boltz.c:193-198): the reference interface keeps all of them on the
    GPU (no eviction warning, so no re-upload) and every ComputeQ matches the
    oracle with its own table, in two rounds."""
    from conftest import rel_l2
    N, Lv = 8, 5.0
    v, eta = orc.grid_0d(N, Lv)
    spb.initialize_coll(N, Lv, v, eta)
    tables = [orc.weights_iso(N, eta, Lv, 1.0, m_i=1.0 + i % 3, m_j=1.0 + i // 3) for i in range(9)]
    rows = [spb.WeightRows(W) for W in tables]   # one conv_weights pointer per pair, kept alive
    f = orc.bkw_initial(N, v).ravel()
    g = orc.maxwellian(v, 0.9, np.array([0.2, 0.0, 0.0]), 1.1).ravel()
    capfd.readouterr()
    for _ in range(2):
        for W, R in zip(tables, rows):
            assert rel_l2(spb.ComputeQ(f, g, R), orc.compute_q(N, v, eta, W, f, g)) < 1e-12
    assert "evicting" not in capfd.readouterr().err
    spb.dealloc_coll()


@pytest.mark.parametrize("masses", [(1.0,), (1.0, 2.0, 0.5)])
def test_projection_kernels_against_oracle(spb, orc, masses):
    """The two projection paths of spbte_conserve_species against the oracle's
    conserveAllMoments (src/conserve.c:207-264): one workgroup per group
    (k_project, an operator with no partial-sum scratch yet) and the split
    (x-plane, group) kernels k_proj_part/k_proj_apply (after an evaluation
    has sized the scratch), for one and three species; each <= 1e-12 relative
    to the oracle and the projected arrays conserve to 1e-12."""
    from conftest import rel_l2
    N, Lv, ngroups = 8, 5.0, 3
    Ns = len(masses)
    v, eta = orc.grid_0d(N, Lv)
    n3 = N ** 3
    rng = np.random.default_rng(0x5EED0007)
    Q = rng.standard_normal((ngroups, Ns * Ns, n3)) * 1e-2
    cons = orc.Conserver(N, v, masses)
    ref = np.stack([cons(Q[g]) for g in range(ngroups)])
    op = spb.Operator(N, Lv, v, eta)
    op.set_species(masses)
    one_wg = op.conserve_species(Q, Ns)
    op.generate_weights_iso(1.0, 2.0, 2.0, masses[0], masses[0])
    f = np.abs(rng.standard_normal((4, n3)))
    op.compute_q(f)  # sizes the operator's scratch for 4 cells
    split = op.conserve_species(Q, Ns)
    op.close()
    w = np.ones(N); w[0] = w[-1] = 0.5
    W3 = ((w[:, None, None] * w[None, :, None] * w[None, None, :]) * (v[1] - v[0]) ** 3).ravel()
    V = [np.repeat(v, N * N), np.tile(np.repeat(v, N), N), np.tile(v, N * N)]
    E = 0.5 * sum(Vd ** 2 for Vd in V)
    for got in (one_wg, split):
        for g in range(ngroups):
            assert rel_l2(got[g], ref[g]) < 1e-12
            # each species' mass, then the total momentum and energy, relative
            # to the same sums of |terms|
            qs = [sum(got[g][n * Ns + m] for n in range(Ns)) for m in range(Ns)]
            for m in range(Ns):
                assert abs(W3 @ qs[m]) <= 1e-12 * (W3 @ np.abs(qs[m]))
            for row in V + [E]:
                tot = sum(masses[m] * (W3 * row) @ qs[m] for m in range(Ns))
                mag = sum(masses[m] * (W3 * np.abs(row)) @ np.abs(qs[m]) for m in range(Ns))
                assert abs(tot) <= 1e-12 * mag
