"""GPU checks of the learned surrogate: the BF16 MFMA forward and backward
kernels against the CPU restatement (oracle/net_oracle.py, which rounds the same
GEMM operands to BF16), training on the GPU, and the surrogate Q through the
FP64 projection.

Tolerances: outputs 2e-3 and gradients 2e-2 relative L2 per parameter block:
the GPU and the restatement round identical FP32 values to BF16, so they differ
only where an FP32 sum of a different order crosses a BF16 rounding boundary
(one BF16 ulp, 2^-8 relative, on that element)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _net_pair(spb, N, Kb, C, r, L, seed):
    from oracle import net_oracle as no
    v = np.linspace(-5, 5, N)
    eta = np.linspace(-3, 3, N)
    op = spb.Operator(N, 5.0, v, eta)
    net = spb.Net(op, Kb, width=C, rank=r, layers=L, seed=seed)
    net.set_loss_floor(0.0)  # the restatement's loss is the plain relative error
    rng = np.random.default_rng(seed)
    off, n = no.layout(Kb ** 3, C, r, L)
    assert n == net.npar and net.offsets() == [off[k][0] for k in off]
    p = (rng.standard_normal(n) * 0.3).astype(np.float32)
    net.set_params(p)
    return op, net, no.NetOracle(Kb ** 3, C, r, L, p), off, rng


@pytest.mark.parametrize("Kb,C,r,L,B", [(4, 16, 8, 2, 8), (8, 32, 16, 1, 16)])
def test_net_forward_and_gradient_match_oracle(spb, Kb, C, r, L, B):
    op, net, ref, off, rng = _net_pair(spb, 8, Kb, C, r, L, seed=Kb + L)
    M = Kb ** 3
    X = rng.standard_normal((M, B, 4)).astype(np.float32)
    T = rng.standard_normal((M, B, 2)).astype(np.float32)
    Y = net.forward_features(X)
    Yr, _ = ref.forward(X)
    assert rel_l2(Y, Yr) < 2e-3
    loss, g = net.grad_features(X, T)
    lr, gr = ref.loss_grad(X, T)
    assert abs(loss - lr) < 2e-3 * lr
    for name, (o, sh) in off.items():
        k = int(np.prod(sh))
        assert rel_l2(g[o:o + k], gr[o:o + k]) < 2e-2, name
    net.close()
    op.close()


def test_net_trains_and_predicts_conserving(spb, orc):
    """A small surrogate (N=16, 12^3 modes) trained on the GPU on exact Q of
    two-Maxwellian cells: the loss falls tenfold, and the predicted Q passes
    through the FP64 projection (mass, momentum, energy conserved)."""
    from spectralbte_amd import workloads
    N, Lv = 16, 6.0
    v, eta = workloads.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    f = workloads.two_maxwellian_ensemble(N, Lv, 64, seed=11)
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    net = spb.Net(op, 12, width=32, rank=32, layers=2, maxpreserve=True, seed=3)
    net.calibrate(d_f.ptr, d_Q.ptr, 64)
    net.reserve(64)
    net.add(d_f.ptr, d_Q.ptr, 64)
    assert net.dataset_size() == 64
    e0 = net.eval(np.arange(64))
    losses = net.train(300, 16, 2e-3)
    e1 = net.eval(np.arange(64))
    assert np.all(np.isfinite(losses))
    assert np.mean(e1 ** 2) < 0.1 * np.mean(e0 ** 2), (np.mean(e0), np.mean(e1))
    Q = net.predict(f)
    w = np.ones(N); w[0] = w[-1] = 0.5
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    pw = (w[I] * w[J] * w[K] * (v[1] - v[0]) ** 3).ravel()
    Cm = np.stack([pw, pw * v[I].ravel(), pw * v[J].ravel(), pw * v[K].ravel(), pw * (v[I] ** 2 + v[J] ** 2 + v[K] ** 2).ravel()])
    assert np.all(np.isfinite(Q))
    assert np.abs(Q @ Cm.T).max() < 1e-12 * (np.abs(Q) @ np.abs(Cm.T)).max()
    net.close()
    op.close()


def test_net_collide_per_group_fallback(spb, orc):
    """spbte_net_collide on 64 shock-profile cells (N=16, L_v=8, hard spheres) in
    groups of 4: every checked cell keeps the exact Q; a group whose checked
    cell's error exceeds tol is evaluated exactly, cell by cell equal to the
    exact operator; the other groups keep the network's prediction bit for bit;
    the count of exact cells is the checked cells plus the fallback groups'
    other cells.  tol = 0 sends every group to the exact operator, a huge tol
    none.  The fraction of cells the network serves at tol 1e-3 (error floored
    at 0.1 of the largest checked |Q|) is printed."""
    from spectralbte_amd import workloads
    N, Lv, ncell, ce = 16, 8.0, 64, 4
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    f = np.concatenate([workloads.shock_profile(N, Lv, 32, mach=m, blend_cells=8.0) for m in (1.2, 1.4)])
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    net = spb.Net(op, 12, width=32, rank=32, layers=2, maxpreserve=False, seed=5)
    net.calibrate(d_f.ptr, d_Q.ptr, ncell)
    net.reserve(ncell)
    net.add(d_f.ptr, d_Q.ptr, ncell)
    net.train(200, 16, 2e-3)
    Qe = op.compute_q(f, conserve=True)
    Qn = net.predict(f, conserve=True)
    for phase in (0, 3):
        for tol, floor_rel in ((1e-3, 0.1), (0.0, 0.1), (1e30, 0.0)):
            n_exact, worst = net.collide_dev(d_f.ptr, d_Q.ptr, ncell, "fp64", check_every=ce, phase=phase, tol=tol,
                                             floor_rel=floor_rel)
            op.sync()
            Q = d_Q.download()
            checked = [g * ce + phase for g in range(ncell // ce)]
            qn = np.linalg.norm(Qe[checked], axis=1)
            fl = floor_rel * qn.max()
            errs = np.linalg.norm(Qn[checked] - Qe[checked], axis=1) / np.maximum(qn, fl)
            assert abs(worst - errs.max()) <= 1e-9 * max(errs.max(), 1e-300), (worst, errs.max())
            exact_cells = 0
            for g, c in enumerate(checked):
                cells = range(g * ce, (g + 1) * ce)
                assert rel_l2(Q[c], Qe[c]) < 1e-12
                if errs[g] > tol:
                    exact_cells += ce
                    for x in cells:
                        assert rel_l2(Q[x], Qe[x]) < 1e-12, (tol, g, x)
                else:
                    exact_cells += 1
                    for x in cells:
                        if x != c:
                            assert np.array_equal(Q[x], Qn[x]), (tol, g, x)
            assert n_exact == exact_cells, (tol, n_exact, exact_cells)
            if tol == 0.0:
                assert n_exact == ncell
            if tol > 1e29:
                assert n_exact == ncell // ce
            if tol == 1e-3:
                print(f"phase {phase}: the network serves {ncell - n_exact} of {ncell} cells at tol 1e-3 "
                      f"(floor 0.1); largest checked error {worst:.3g}")
    net.close()
    op.close()
