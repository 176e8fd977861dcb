"""The C-ABI library: it loads, it exports exactly the functions include/*.h
declares, and its host-only paths (weight-file loading, ghat/gHat3) agree with
the oracle.  No GPU compute here."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.add(m.group(1))
    return names


def exported_functions(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[2] for ln in out.splitlines() if len(ln.split()) == 3 and ln.split()[1] == "T"}


def test_library_exports_match_headers(spb):
    spb.lib()
    decl = declared_functions()
    assert {"ComputeQ", "ComputeQ_maxPreserve", "conserveAllMoments", "initialize_weights", "spbte_compute_q"} <= decl
    assert exported_functions(spb.LIB_PATH) == decl


def test_reference_abi_signatures(spb):
    """Every reference entry point resolves with the reference's arity
    (src/collisions.h:27-44, src/conserve.h:6-18, src/weights.h:12-49)."""
    L = spb.lib()
    for name in ["initialize_coll", "dealloc_coll", "ComputeQ", "ComputeQ_maxPreserve", "fft3D", "ifft3D",
                 "initialize_conservation", "initialize_conservation_fast", "dealloc_conservation", "factorMatrixIntoLU",
                 "solveWithCCt", "conserveAllMoments", "createCCtAndPivot", "alloc_weights", "initialize_weights",
                 "generate_conv_weights_iso", "dealloc_weights", "ghat", "gHat3"]:
        assert hasattr(L, name)


def test_species_struct_layout(spb):
    """src/species.h:11-26 passed by value: same size and offsets as the C header."""
    S = spb.Species
    assert ctypes.sizeof(S) == 8 * 3 + 8 * 8 + 8 * 2 + 80
    assert S.mass.offset == 32 and S.d_ref.offset == 48 and S.name.offset == 104


def test_no_gpu_is_an_error_not_a_fallback(spb):
    """Without a visible GPU the batched interface refuses to run (no CPU path)."""
    if spb.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(spb.SpbteError):
        spb.Operator(8, 5.0, np.linspace(-5, 5, 8), np.linspace(-3, 3, 8))


def test_initialize_weights_loads_reference_file_and_ghat3(spb, orc, tmp_path):
    """initialize_weights (src/weights.c:33-88) finds Weights/N8_isotropic_L_v5_lambda0.wts
    in the working directory and loads it into the caller's rows without a GPU;
    gHat3 (src/weights.c:181-262) then matches the oracle's integral."""
    wd = tmp_path / "run"
    (wd / "Weights").mkdir(parents=True)
    shutil.copy(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), wd / "Weights")
    v, eta = orc.grid_0d(8, 5.0)
    L = spb.lib()
    cw = ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))()
    L.alloc_weights(8, ctypes.byref(cw), 1)
    eta_c = np.ascontiguousarray(eta)
    L.initialize_weights.restype = None
    L.initialize_weights.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p, spb.Species, spb.Species]
    cwd = os.getcwd()
    os.chdir(wd)
    try:
        sp = spb.default_species()
        L.initialize_weights(8, eta_c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5.0, 0.0, 0, 0,
                             ctypes.cast(cw[0], ctypes.c_void_p), sp, sp)
    finally:
        os.chdir(cwd)
    rows = cw[0]
    Wf = orc.weights_load(os.path.join(GOLDEN, "BKW8", "N8_isotropic_L_v5_lambda0.wts"), 8)
    for r in (0, 77, 511):
        got = np.ctypeslib.as_array(rows[r], shape=(512,))
        assert np.array_equal(got, Wf[r])
    # gHat3 for a few (xi, zeta): prefactor * I(mu|zeta|, |xi|, |xi - mu zeta|)
    P = 16 * np.pi ** 2 * (eta[1] - eta[0]) ** 3 / (2 * np.pi) ** 1.5 / (4 * np.pi)
    for (l, m, n), (i, j, k) in [((0, 1, 2), (3, 4, 5)), ((7, 7, 7), (0, 0, 0)), ((4, 4, 4), (4, 4, 4))]:
        xi = np.array([eta[l], eta[m], eta[n]]); ze = np.array([eta[i], eta[j], eta[k]])
        ref = P * orc.ghat_integral(0.5 * np.linalg.norm(ze), np.linalg.norm(xi), np.linalg.norm(xi - 0.5 * ze), 5.0, 0.0)
        got = L.gHat3(*xi, *ze)
        assert abs(got - ref) <= 1e-14 * max(abs(ref), 1e-300) + 1e-18
        w = np.ones(8); w[0] = w[-1] = 0.5
        r_idx = (i * 8 + j) * 8 + k; c_idx = (l * 8 + m) * 8 + n
        assert abs(w[l] * w[m] * w[n] * got - Wf[r_idx, c_idx]) <= 1e-13 * np.abs(Wf).max()
    L.dealloc_weights(8, rows)


def test_run_and_net_entry_points_reject_bad_arguments(spb):
    """Argument checks of the batched runs, transport and surrogate entry points
    happen before any device call, so they are errors (not crashes) without a GPU."""
    import ctypes
    L = spb.lib()
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)
    x = (ctypes.c_double * 8)()
    assert L.spbte_run0d_create(ctypes.byref(h), None, 4, 2, 0.01, 1.0, 0, 1) == -1        # no operator
    assert L.spbte_run1d_create(ctypes.byref(h), None, 4, 1, None, None, 0.01, 1.0, 1, 1.0, 2.0, 0) == -1  # no mesh
    assert L.spbte_run1d_create(ctypes.byref(h), None, 4, 1, ctypes.cast(x, dp), ctypes.cast(x, dp), 0.01, 1.0, 7,
                                1.0, 2.0, 0) == -1                                          # bad boundary
    assert L.spbte_run1d_create(ctypes.byref(h), None, 4, 2, ctypes.cast(x, dp), ctypes.cast(x, dp), 0.01, 1.0, 2,
                                1.0, 2.0, 0) == -7                                          # periodic at order 2
    assert L.spbte_run_step(None, 1) == -1
    assert L.spbte_run_set_surrogate(None, None, 8, 1e-3, 0.0) == -1
    assert L.spbte_advect(None, None, None, 4, 1, None, None, 0.01, 1, 1.0, 2.0) == -1
    assert L.spbte_net_create(ctypes.byref(h), None, 12, 64, 128, 4, 0, 1) == -1
    assert L.spbte_net_collide(None, None, None, 4, 0, 8, 0, 1e-3, 0.0, None, None) == -1
    assert L.spbte_net_train(None, 10, 8, 1e-3, 1, None) == -1
    assert "bad" in L.spbte_last_error().decode() or "null" in L.spbte_last_error().decode()


def test_operator_entry_points_reject_null_operator(spb):
    """The operator's compute, transform, projection and timing entry points
    return SPBTE_ERR_ARG for a null operator or negative counts instead of
    touching it."""
    import ctypes
    L = spb.lib()
    assert L.spbte_compute_q(None, None, None, None, 1, 0, 0) == -1
    assert L.spbte_qhat(None, None, None, None, 1, 0) == -1
    assert L.spbte_transform(None, None, None, 1, 0) == -1
    assert L.spbte_conserve(None, None, 1) == -1
    assert L.spbte_conserve_pairs(None, None, 1) == -1
    assert L.spbte_conserve_species(None, None, 1) == -1
    assert L.spbte_compute_q(None, None, None, None, -1, 0, 0) == -1
    t = ctypes.c_double()
    assert L.spbte_time_compute_q(None, None, None, None, 1, 0, 0, 3, ctypes.byref(t), ctypes.byref(t)) == -1
    assert "bad arguments" in L.spbte_last_error().decode()


def test_output_flags_file(spb, tmp_path):
    """read_flags of src/output.c:417-443 on the reference's own flag files (all
    columns on), and a file that switches two columns off."""
    for case in ("BKW8/BKW8.test.out", "heat_transport/heat_transport.test.out"):
        assert spb.read_output_flags(os.path.join(GOLDEN, case)) == spb.OUT_ALL
    p = tmp_path / "flags.out"
    p.write_text("# flags\ndensity\n1\nvelocity\n0\ntemperature\n1\nslice\n0\nStop\n")
    assert spb.read_output_flags(str(p)) == spb.OUT_ALL & ~spb.OUT_VELOCITY & ~spb.OUT_SLICE
    p.write_text("density\n1\n")
    with pytest.raises(spb.SpbteError):
        spb.read_output_flags(str(p))


def _aniso_call(spb, N, eta, Lv, lam, weight_flag, glance, wd):
    L = spb.lib()
    cw = ctypes.POINTER(ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))()
    L.alloc_weights(N, ctypes.byref(cw), 1)
    L.initialize_weights_AnIso.restype = None
    L.initialize_weights_AnIso.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]
    eta_c = np.ascontiguousarray(eta)
    cwd = os.getcwd()
    os.chdir(wd)
    try:
        L.initialize_weights_AnIso(N, eta_c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), Lv, lam, weight_flag,
                                   ctypes.cast(cw[0], ctypes.c_void_p), glance)
    finally:
        os.chdir(cwd)
    return cw[0], eta_c


def test_initialize_weights_aniso_loads_glance_file(spb, orc, tmp_path):
    """With glance > 0, initialize_weights_AnIso loads
    Weights/N4_AnIso_L_v5_lambda1_glance0.01.wts (src/aniso_weights.c:68,
    :71-81) without a GPU."""
    N, Lv, lam = 4, 5.0, 1.0
    v, eta = orc.grid_0d(N, Lv)
    W = np.random.default_rng(4).standard_normal((64, 64))
    (tmp_path / "Weights").mkdir()
    W.tofile(tmp_path / "Weights" / "N4_AnIso_L_v5_lambda1_glance0.01.wts")
    rows, _ = _aniso_call(spb, N, eta, Lv, lam, 0, 0.01, tmp_path)
    for r in (0, 33, 63):
        assert np.array_equal(np.ctypeslib.as_array(rows[r], shape=(64,)), W[r])
    spb.lib().dealloc_weights(N, rows)


def test_initialize_weights_aniso_loads_landau_file_and_ghat3l(spb, orc, tmp_path):
    """initialize_weights_AnIso (src/aniso_weights.c:37-81) with glance 0 finds
    Weights/N4_Aniso_L_v5_lambda1_Landau.wts and loads it into the caller's rows
    without a GPU (the third argument is L_v, as exec/boltz.c:154 passes it);
    gHat3L (:371-398) then matches the oracle's entry."""
    N, Lv, lam = 4, 5.0, 1.0
    v, eta = orc.grid_0d(N, Lv)
    W = orc.weights_landau(N, eta, Lv, lam)
    (tmp_path / "Weights").mkdir()
    W.tofile(tmp_path / "Weights" / "N4_Aniso_L_v5_lambda1_Landau.wts")
    rows, eta_c = _aniso_call(spb, N, eta, Lv, lam, 0, 0.0, tmp_path)
    for r in (0, 21, 63):
        assert np.array_equal(np.ctypeslib.as_array(rows[r], shape=(64,)), W[r])
    L = spb.lib()
    L.gHat3L.restype = ctypes.c_double
    L.gHat3L.argtypes = [ctypes.c_double] * 6
    for (i, j, k), (l, m, n) in [((0, 1, 2), (3, 0, 1)), ((2, 2, 2), (1, 3, 0)), ((3, 3, 3), (2, 2, 2))]:
        got = L.gHat3L(eta[i], eta[j], eta[k], eta[l], eta[m], eta[n])
        ref = orc.landau_entry((eta[i], eta[j], eta[k]), (eta[l], eta[m], eta[n]), Lv, lam)
        assert abs(got - ref) <= 1e-14 * abs(ref)
        assert abs(got - W[(i * N + j) * N + k, (l * N + m) * N + n]) <= 1e-14 * abs(ref)
    L.dealloc_weights(N, rows)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 (or a multi-rank launch) exits
    with a message and status 2 before touching a device."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for args, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})):
        r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), *args], capture_output=True, text=True,
                           timeout=120, env=dict(os.environ, **env))
        assert r.returncode == 2
        assert "one GPU" in r.stderr and r.stdout == ""
