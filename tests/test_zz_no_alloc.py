"""spbte_compute_q_noalloc (the allocation-free variant of spbte_compute_q)
through the C ABI: its output equals spbte_compute_q's on the same batch, bit
for bit (same kernels and tile), on two calls in a row with the second batch
smaller; batches beyond the fixed workspace and tables that are not resident
are refused before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = {"fp64": 0, "fp32": 1, "fp16x2": 2, "fp32a": 3}
SPBTE_ERR_ARG = -1


def _noalloc(spb, op, d_f, d_Q, ncell, form, flags):
    return spb.lib().spbte_compute_q_noalloc(op.h, d_f.ptr, None, d_Q.ptr, ncell, FORMS[form], flags)


def _allocating(spb, op, d_f, d_Q, ncell, form, flags):
    return spb.lib().spbte_compute_q(op.h, d_f.ptr, None, d_Q.ptr, ncell, FORMS[form], flags)


@pytest.mark.parametrize("N,Lv,sizes,form,mp", [(16, 5.0, (6, 3), "fp64", True), (16, 5.0, (5, 2), "fp32a", False),
                                                (32, 6.0, (2, 1), "fp32a", True), (32, 6.0, (3, 1), "fp64", False)])
def test_noalloc_matches_compute_q(spb, N, Lv, sizes, form, mp):
    from spectralbte_amd import workloads
    v, eta = workloads.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    f = workloads.two_maxwellian_ensemble(N, Lv, max(sizes), seed=7 + N)
    flags = spb.SELF | spb.CONSERVE | (spb.MAXPRESERVE if mp else 0)
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    d_R = spb.DeviceArray(f.shape)
    for n in sizes:
        print(f"N={N} form={form} maxpreserve={mp} ncell={n} tile={op.get_tile(n, form, mp)}", flush=True)
        # the allocating call builds the table layout this batch's tile streams
        assert _allocating(spb, op, d_f, d_R, n, form, flags) == 0
        op.sync()
        ref = d_R.download()[:n]
        d_Q.upload(np.full(f.shape, np.nan))
        rc = _noalloc(spb, op, d_f, d_Q, n, form, flags)
        assert rc == 0, spb.lib().spbte_last_error()
        op.sync()
        got = d_Q.download()
        assert np.array_equal(got[:n], ref), (n, float(np.abs(got[:n] - ref).max()))
        assert np.all(np.isnan(got[n:]))   # nothing beyond the batch is written
    op.close()


def test_noalloc_refusals(spb):
    """A batch larger than the workspace, and a form whose table was never
    built, return SPBTE_ERR_ARG with a message; Q is left untouched."""
    from spectralbte_amd import workloads
    N, Lv = 32, 6.0
    v, eta = workloads.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    ncell = 64   # 64 cells at N=32 need far more than the fixed workspace
    f = workloads.two_maxwellian_ensemble(N, Lv, ncell, seed=3)
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    d_Q.upload(np.zeros(f.shape))
    flags = spb.SELF | spb.MAXPRESERVE
    assert _allocating(spb, op, d_f, d_Q, ncell, "fp64", flags) == 0   # tables resident
    op.sync()
    d_Q.upload(np.zeros(f.shape))
    assert _noalloc(spb, op, d_f, d_Q, ncell, "fp64", flags) == SPBTE_ERR_ARG
    assert b"workspace" in spb.lib().spbte_last_error()
    assert _noalloc(spb, op, d_f, d_Q, 1, "fp16x2", flags) == SPBTE_ERR_ARG
    assert b"not resident" in spb.lib().spbte_last_error()
    op.sync()
    assert not np.any(d_Q.download())
    assert _noalloc(spb, op, d_f, d_Q, 0, "fp64", flags) == 0
    op.close()
