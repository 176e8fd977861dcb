"""spectralbte_amd -- MI355X collision operator of SpectralBTE, Python host side.

The product is ``libspectralbte_amd.so`` (C host code + HIP kernels for gfx950),
whose C ABI is declared in ``include/``.  This module loads it with ctypes and
offers two views of it:

* the reference's own interface (``initialize_coll``, ``ComputeQ``,
  ``ComputeQ_maxPreserve``, ``fft3D``, ``initialize_conservation``,
  ``conserveAllMoments``, ``initialize_weights`` ...), with numpy arrays in place
  of the C pointers -- the names, argument meanings and error behaviour of
  src/collisions.h, src/conserve.h and src/weights.h;
* ``Operator``: the batched, device-resident interface of ``include/spbte.h``.

There is no CPU fallback: if the library is missing, ``lib()`` raises; if no GPU
is visible, operator creation raises ``SpbteError``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspectralbte_amd.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

FORMS = {"fp64": 0, "fp32": 1, "fp16x2": 2, "fp32a": 3}
MAXPRESERVE, CONSERVE, SELF = 1, 2, 4

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class SpbteError(RuntimeError):
    pass


def build(force=False):
    """Compile the library for gfx950 with the Makefile beside this file."""
    args = ["make", "-s", "-C", _HERE, "-j8"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


def lib():
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpbteError(f"{LIB_PATH} is missing: run spectralbte_amd.build() (the operator has no CPU path)")
        L = ctypes.CDLL(LIB_PATH)
        I, D, S = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            # reference interface
            "initialize_coll": (None, [I, D, _dp, _dp]),
            "dealloc_coll": (None, []),
            "ComputeQ": (None, [_dp, _dp, _dp, ctypes.POINTER(_dp)]),
            "ComputeQ_maxPreserve": (None, [_dp, _dp, _dp, ctypes.POINTER(_dp)]),
            "fft3D": (None, [_dp, _dp, I]),
            "ifft3D": (None, [_dp, _dp]),
            "initialize_conservation": (None, [I, D, _dp, _vp, I]),
            "initialize_conservation_fast": (None, [I, D, _dp]),
            "dealloc_conservation": (None, []),
            "factorMatrixIntoLU": (None, [I, D]),
            "solveWithCCt": (None, [I, _dp]),
            "conserveAllMoments": (None, [ctypes.POINTER(_dp)]),
            "createCCtAndPivot": (None, []),
            "alloc_weights": (None, [I, _vp, I]),
            "generate_conv_weights_iso": (None, [ctypes.POINTER(_dp)]),
            "dealloc_weights": (None, [I, ctypes.POINTER(_dp)]),
            "ghat": (D, [D, _vp]),
            "gHat3": (D, [D, D, D, D, D, D]),
            # batched interface
            "spbte_last_error": (ctypes.c_char_p, []),
            "spbte_device_count": (I, [ip]),
            "spbte_malloc": (I, [ctypes.POINTER(_vp), S]),
            "spbte_free": (I, [_vp]),
            "spbte_memcpy_htod": (I, [_vp, _vp, S]),
            "spbte_memcpy_dtoh": (I, [_vp, _vp, S]),
            "spbte_memcpy_dtod": (I, [_vp, _vp, S]),
            "spbte_device_copy": (I, [_vp, _vp, S]),
            "spbte_valu_peak": (I, [I, _dp, _dp]),
            "spbte_op_synchronize": (I, [_vp]),
            "spbte_op_create": (I, [ctypes.POINTER(_vp), I, D, _dp, _dp]),
            "spbte_op_destroy": (None, [_vp]),
            "spbte_op_set_species": (I, [_vp, I, _dp]),
            "spbte_op_set_weights": (I, [_vp, _dp]),
            "spbte_op_set_weights_rows": (I, [_vp, ctypes.POINTER(_dp)]),
            "spbte_op_load_weights_file": (I, [_vp, ctypes.c_char_p]),
            "spbte_op_generate_weights_iso": (I, [_vp, D, D, D, D, D]),
            "spbte_op_get_weights": (I, [_vp, _dp]),
            "spbte_compute_q": (I, [_vp, _vp, _vp, _vp, I, I, I]),
            "spbte_compute_q_noalloc": (I, [_vp, _vp, _vp, _vp, I, I, I]),
            "spbte_conserve": (I, [_vp, _vp, I]),
            "spbte_conserve_species": (I, [_vp, _vp, I]),
            "spbte_transform": (I, [_vp, _vp, _vp, I, I]),
            "spbte_qhat": (I, [_vp, _vp, _vp, _vp, I, I]),
            "spbte_set_tile": (I, [_vp, I, I, I, I, I, I]),
            "spbte_autotune": (I, [_vp, I, I, I, ip, ip, ip, ip]),
            "spbte_get_tile": (I, [_vp, I, I, I, ip, ip, ip, ip]),
            "spbte_time_compute_q": (I, [_vp, _vp, _vp, _vp, I, I, I, I, _dp, _dp]),
            "spbte_run0d_create": (I, [ctypes.POINTER(_vp), _vp, I, I, D, D, I, I]),
            "spbte_run1d_create": (I, [ctypes.POINTER(_vp), _vp, I, I, _dp, _dp, D, D, I, D, D, I]),
            "spbte_run_destroy": (None, [_vp]),
            "spbte_run_set_state": (I, [_vp, _dp]),
            "spbte_run_get_state": (I, [_vp, _dp]),
            "spbte_run_state": (I, [_vp, ctypes.POINTER(_vp)]),
            "spbte_run_use_graph": (I, [_vp, I]),
            "spbte_run_step": (I, [_vp, I]),
            "spbte_run_moments": (I, [_vp, _dp]),
            "spbte_advect": (I, [_vp, _vp, _vp, I, I, _dp, _dp, D, I, D, D]),
            "spbte_net_create": (I, [ctypes.POINTER(_vp), _vp, I, I, I, I, I, ctypes.c_ulonglong]),
            "spbte_net_destroy": (None, [_vp]),
            "spbte_net_num_params": (ctypes.c_long, [_vp]),
            "spbte_net_layout": (I, [_vp, ctypes.POINTER(ctypes.c_long)]),
            "spbte_net_get_params": (I, [_vp, _fp]),
            "spbte_net_set_params": (I, [_vp, _fp]),
            "spbte_net_set_scales": (I, [_vp, D, D]),
            "spbte_net_get_scales": (I, [_vp, _dp, _dp]),
            "spbte_net_calibrate": (I, [_vp, _vp, _vp, I]),
            "spbte_net_dataset_reserve": (I, [_vp, ctypes.c_long]),
            "spbte_net_dataset_add": (I, [_vp, _vp, _vp, I]),
            "spbte_net_dataset_size": (ctypes.c_long, [_vp]),
            "spbte_net_train": (I, [_vp, I, I, ctypes.c_float, ctypes.c_ulonglong, _fp]),
            "spbte_net_grad": (I, [_vp, ip, I, _fp, _dp]),
            "spbte_net_eval": (I, [_vp, ip, I, _fp]),
            "spbte_net_forward_features": (I, [_vp, _fp, I, _fp]),
            "spbte_net_grad_features": (I, [_vp, _fp, _fp, I, _fp, _dp]),
            "spbte_net_predict": (I, [_vp, _vp, _vp, I, I]),
            "spbte_net_collide": (I, [_vp, _vp, _vp, I, I, I, I, D, D, ip, _dp]),
            "spbte_net_set_loss_floor": (I, [_vp, D]),
            "spbte_run_set_surrogate": (I, [_vp, _vp, I, D, D]),
            "spbte_read_output_flags": (I, [ctypes.c_char_p, ip]),
            "spbte_run_open_output": (I, [_vp, ctypes.c_char_p, I, I]),
            "spbte_run_write_output": (I, [_vp, D]),
            "spbte_run_close_output": (I, [_vp]),
            "spbte_run_store_restart": (I, [_vp, ctypes.c_char_p, ctypes.c_char_p, I]),
            "spbte_run_load_restart": (I, [_vp, ctypes.c_char_p, ctypes.c_char_p, ip]),
            "spbte_run_surrogate_stats": (I, [_vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                                              ctypes.POINTER(ctypes.c_long), _dp]),
            "spbte_run_set_forcing": (I, [_vp, D]),
            "spbte_op_store_weights_file": (I, [_vp, ctypes.c_char_p]),
            "spbte_op_generate_weights_landau": (I, [_vp, D]),
            "spbte_op_generate_weights_glance": (I, [_vp, D, D]),
            "spbte_op_select_table": (I, [_vp, I]),
            "spbte_conserve_pairs": (I, [_vp, _vp, I]),
            "spbte_op_get_species": (I, [_vp, ip, _dp]),
            "spbte_run0d_create_species": (I, [ctypes.POINTER(_vp), _vp, I, I, I, D, D, I, I]),
            "spbte_run1d_create_species": (I, [ctypes.POINTER(_vp), _vp, I, I, I, _dp, _dp, D, D, I, D, D, I]),
            "spbte_run_species_state": (I, [_vp, I, ctypes.POINTER(_vp)]),
            "spbte_run_open_output_species": (I, [_vp, I, ctypes.c_char_p, I, I]),
            "spbte_run_store_restart_species": (I, [_vp, I, ctypes.c_char_p]),
            "spbte_run_load_restart_species": (I, [_vp, I, ctypes.c_char_p]),
            "spbte_weightgen_main": (I, [I, ctypes.POINTER(ctypes.c_char_p)]),
            # front end (spbte_boltz.h)
            "spbte_read_input": (I, [ctypes.c_char_p, _vp]),
            "spbte_species_defaults": (None, [_vp, I]),
            "spbte_read_species": (I, [ctypes.c_char_p, _vp]),
            "spbte_read_mesh": (I, [ctypes.c_char_p, I, ip, _dp, _dp, _dp]),
            "spbte_velocity_grid": (None, [I, D, I, _dp, _dp]),
            "spbte_initial_state": (I, [_vp, _vp, _dp, I, _dp]),
            "spbte_boltz_main": (I, [I, ctypes.POINTER(ctypes.c_char_p)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SpbteError(f"spbte error {rc}: {lib().spbte_last_error().decode(errors='replace')}")


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def valu_peak(fp64=False):
    """(TFLOP/s, clock GHz) of independent FMA chains on every CU (spbte_valu_peak)."""
    t, g = ctypes.c_double(), ctypes.c_double()
    _check(lib().spbte_valu_peak(int(bool(fp64)), ctypes.byref(t), ctypes.byref(g)))
    return t.value, g.value


def device_count():
    n = ctypes.c_int(0)
    rc = lib().spbte_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


# ------------------------------------------------------------ device arrays
class DeviceArray:
    """A float64 device buffer owned by Python (spbte_malloc / spbte_free)."""

    def __init__(self, shape, data=None):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = _vp()
        _check(lib().spbte_malloc(ctypes.byref(self.ptr), max(self.nbytes, 8)))
        if data is not None:
            self.upload(data)

    @classmethod
    def from_host(cls, arr):
        arr = _c(arr)
        return cls(arr.shape, arr)

    def upload(self, arr):
        arr = _c(arr)
        assert arr.nbytes == self.nbytes, (arr.shape, self.shape)
        _check(lib().spbte_memcpy_htod(self.ptr, arr.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=np.float64)
        _check(lib().spbte_memcpy_dtoh(out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def offset(self, elems):
        return _vp(self.ptr.value + 8 * int(elems))

    def free(self):
        if self.ptr:
            lib().spbte_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------- batched operator
class Operator:
    """spbte_op: one velocity grid, its weight table and the batched kernels."""

    def __init__(self, N, L_v, v, eta):
        self.N = int(N)
        self.n3 = self.N ** 3
        self.v = _c(v)
        self.eta = _c(eta)
        self.h = _vp()
        _check(lib().spbte_op_create(ctypes.byref(self.h), self.N, float(L_v), _p(self.v), _p(self.eta)))

    def close(self):
        if self.h:
            lib().spbte_op_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_species(self, masses):
        m = _c(masses)
        _check(lib().spbte_op_set_species(self.h, len(m), _p(m)))

    def select_table(self, slot):
        """Table slot for the table calls and evaluations (pair (l, m) of Ns
        species in slot m*Ns + l)."""
        _check(lib().spbte_op_select_table(self.h, int(slot)))

    def set_weights(self, W):
        W = _c(W)
        assert W.size == self.n3 * self.n3
        _check(lib().spbte_op_set_weights(self.h, _p(W)))

    def load_weights_file(self, path):
        _check(lib().spbte_op_load_weights_file(self.h, str(path).encode()))

    def generate_weights_iso(self, lam, d_i=2.0, d_j=2.0, m_i=1.0, m_j=1.0):
        _check(lib().spbte_op_generate_weights_iso(self.h, lam, d_i, d_j, m_i, m_j))

    def generate_weights_landau(self, lam):
        """The anisotropic generator's glance = 0 (Landau) table on this grid."""
        _check(lib().spbte_op_generate_weights_landau(self.h, float(lam)))

    def generate_weights_glance(self, lam, glance):
        """The anisotropic generator's glancing-collision (glance > 0) table on this grid."""
        _check(lib().spbte_op_generate_weights_glance(self.h, float(lam), float(glance)))

    def store_weights_file(self, path):
        _check(lib().spbte_op_store_weights_file(self.h, os.fsencode(path)))

    def get_weights(self):
        W = np.empty((self.n3, self.n3))
        _check(lib().spbte_op_get_weights(self.h, _p(W)))
        return W

    def sync(self):
        _check(lib().spbte_op_synchronize(self.h))

    # device-pointer level
    def compute_q_dev(self, d_f, d_g, d_Q, ncell, form="fp64", flags=0):
        _check(lib().spbte_compute_q(self.h, d_f, d_g, d_Q, int(ncell), FORMS[form], int(flags)))

    # numpy level
    def compute_q(self, f, g=None, form="fp64", maxpreserve=False, conserve=False):
        """f, g: (ncell, N^3) or (N,N,N).  Returns Q with f's shape."""
        shape = np.shape(f)
        f = _c(f).reshape(-1, self.n3)
        ncell = f.shape[0]
        df = DeviceArray.from_host(f)
        dg = DeviceArray.from_host(_c(g).reshape(-1, self.n3)) if g is not None else None
        dQ = DeviceArray((ncell, self.n3))
        flags = (MAXPRESERVE if maxpreserve else 0) | (CONSERVE if conserve else 0) | (SELF if g is None else 0)
        self.compute_q_dev(df.ptr, dg.ptr if dg is not None else None, dQ.ptr, ncell, form, flags)
        self.sync()
        return dQ.download().reshape(shape)

    def conserve(self, Q):
        shape = np.shape(Q)
        Q = _c(Q).reshape(-1, self.n3)
        dQ = DeviceArray.from_host(Q)
        _check(lib().spbte_conserve(self.h, dQ.ptr, Q.shape[0]))
        self.sync()
        return dQ.download().reshape(shape)

    def conserve_species(self, Q, Ns):
        """Q: (ngroups, Ns*Ns, N^3)."""
        shape = np.shape(Q)
        Q = _c(Q).reshape(-1, Ns * Ns * self.n3)
        dQ = DeviceArray.from_host(Q)
        _check(lib().spbte_conserve_species(self.h, dQ.ptr, Q.shape[0]))
        self.sync()
        return dQ.download().reshape(shape)

    def transform(self, x, invert=False):
        """x: complex (ncell, N^3) or (N,N,N).  Returns complex with x's shape."""
        shape = np.shape(x)
        x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1, self.n3)
        d = DeviceArray.from_host(x.view(np.float64))
        _check(lib().spbte_transform(self.h, d.ptr, d.ptr, x.shape[0], int(bool(invert))))
        self.sync()
        return d.download().view(np.complex128).reshape(shape)

    def qhat(self, Fhat, Ghat, form="fp64"):
        shape = np.shape(Fhat)
        F = np.ascontiguousarray(Fhat, dtype=np.complex128).reshape(-1, self.n3)
        G = np.ascontiguousarray(Ghat, dtype=np.complex128).reshape(-1, self.n3)
        dF = DeviceArray.from_host(F.view(np.float64))
        dG = DeviceArray.from_host(G.view(np.float64))
        dQ = DeviceArray(dF.shape)
        _check(lib().spbte_qhat(self.h, dF.ptr, dG.ptr, dQ.ptr, F.shape[0], FORMS[form]))
        self.sync()
        return dQ.download().view(np.complex128).reshape(shape)

    def set_tile(self, form, maxpreserve, tr, lines, tb, kind=0):
        """Pin a tile (kind 0: tiled kernel; kind k >= 1: streaming kernel with the summed
        x-planes in k slices); tr=0 unpins."""
        _check(lib().spbte_set_tile(self.h, FORMS[form], int(maxpreserve), int(kind), tr, lines, tb))

    def get_tile(self, ncell, form, maxpreserve=False):
        """(kind, tr, lines, tb) a batch of ncell cells would use."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().spbte_get_tile(self.h, ncell, FORMS[form], int(maxpreserve), *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, ncell, form, maxpreserve=False):
        """Time every tile on this batch shape, pin the fastest; returns (kind, tr, lines, tb)."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().spbte_autotune(self.h, ncell, FORMS[form], int(maxpreserve), *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def time_compute_q(self, d_f, d_g, d_Q, ncell, form="fp64", flags=0, iters=3):
        tot, ker = ctypes.c_double(), ctypes.c_double()
        _check(lib().spbte_time_compute_q(self.h, d_f, d_g, d_Q, int(ncell), FORMS[form], int(flags), int(iters),
                                          ctypes.byref(tot), ctypes.byref(ker)))
        return tot.value, ker.value


# ------------------------------------------------------ batched time stepping
BC_COPY, BC_DIFFUSE, BC_PERIODIC, BC_DIFFUSE_LEFT = 0, 1, 2, 3
OUT_DENSITY, OUT_VELOCITY, OUT_TEMPERATURE, OUT_PRESSURE, OUT_ENTROPY, OUT_SLICE = 1, 2, 4, 8, 16, 32
OUT_ALL = 63


def read_output_flags(path):
    """The column flags of a reference output-flags file (src/output.c:417-443)."""
    f = ctypes.c_int(0)
    _check(lib().spbte_read_output_flags(os.fsencode(path), ctypes.byref(f)))
    return f.value


class Run:
    """spbte_run: the state of a batch of cells in HBM, advanced by whole steps
    (exec/boltz.c:193-241 for 0D, :285-354 for 1D).  Create with ``Run.ensemble``
    or ``Run.mesh``."""

    def __init__(self, op, handle, ncell, ns=1):
        self.op, self.h, self.ncell, self.ns = op, handle, ncell, ns
        self.n3 = op.N ** 3

    @classmethod
    def ensemble(cls, op, f0, order=2, dt=0.01, Kn=1.0, form="fp64", maxpreserve=True):
        f0 = _c(f0).reshape(-1, op.N ** 3)
        h = _vp()
        _check(lib().spbte_run0d_create(ctypes.byref(h), op.h, f0.shape[0], order, dt, Kn, FORMS[form],
                                        int(maxpreserve)))
        r = cls(op, h, f0.shape[0])
        r.set_state(f0)
        return r

    @classmethod
    def mesh(cls, op, f0, x, dx, order=1, dt=0.01, Kn=1.0, bc=BC_DIFFUSE, T0=1.0, T1=2.0, form="fp64"):
        """f0: (nX, N^3) interior cells; x, dx: nX + 2*order entries with ghosts."""
        f0 = _c(f0).reshape(-1, op.N ** 3)
        x, dx = _c(x), _c(dx)
        assert x.size == f0.shape[0] + 2 * order and dx.size == x.size
        h = _vp()
        _check(lib().spbte_run1d_create(ctypes.byref(h), op.h, f0.shape[0], order, _p(x), _p(dx), dt, Kn, bc, T0, T1,
                                        FORMS[form]))
        r = cls(op, h, f0.shape[0])
        r.set_state(f0)
        return r

    @classmethod
    def ensemble_species(cls, op, f0, order=2, dt=0.01, Kn=1.0, form="fp64", maxpreserve=True):
        """Ns species (op.set_species, pair tables in slots m*Ns + l): f0 is
        (Ns, ncell, N^3)."""
        f0 = _c(f0)
        ns, ncell = f0.shape[0], f0.shape[1]
        h = _vp()
        _check(lib().spbte_run0d_create_species(ctypes.byref(h), op.h, ns, ncell, order, dt, Kn, FORMS[form],
                                                int(maxpreserve)))
        r = cls(op, h, ncell, ns)
        r.set_state(f0)
        return r

    @classmethod
    def mesh_species(cls, op, f0, x, dx, order=1, dt=0.01, Kn=1.0, bc=BC_DIFFUSE, T0=1.0, T1=2.0, form="fp64"):
        """Ns species on a 1D mesh: f0 is (Ns, nX, N^3)."""
        f0 = _c(f0)
        ns, nX = f0.shape[0], f0.shape[1]
        x, dx = _c(x), _c(dx)
        assert x.size == nX + 2 * order and dx.size == x.size
        h = _vp()
        _check(lib().spbte_run1d_create_species(ctypes.byref(h), op.h, ns, nX, order, _p(x), _p(dx), dt, Kn, bc, T0,
                                                T1, FORMS[form]))
        r = cls(op, h, nX, ns)
        r.set_state(f0)
        return r

    def set_state(self, f):
        f = _c(f)
        assert f.size == self.ns * self.ncell * self.n3
        _check(lib().spbte_run_set_state(self.h, _p(f)))

    def state(self):
        """(ncell, N^3), or (Ns, ncell, N^3) for several species."""
        out = np.empty((self.ns, self.ncell, self.n3))
        _check(lib().spbte_run_get_state(self.h, _p(out)))
        return out[0] if self.ns == 1 else out

    def state_ptr(self):
        p = _vp()
        _check(lib().spbte_run_state(self.h, ctypes.byref(p)))
        return p

    def use_graph(self, on=True):
        _check(lib().spbte_run_use_graph(self.h, int(on)))

    def step(self, nsteps=1, sync=True):
        _check(lib().spbte_run_step(self.h, int(nsteps)))
        if sync:
            self.op.sync()

    def set_forcing(self, Ma):
        """Poiseuille body force along v_y (1D order-2 runs, Init_field 5)."""
        _check(lib().spbte_run_set_forcing(self.h, float(Ma)))

    def set_surrogate(self, net, check_every=8, tol=1e-3, floor_rel=0.0):
        _check(lib().spbte_run_set_surrogate(self.h, net.h if net is not None else None, int(check_every), float(tol),
                                             float(floor_rel)))

    def surrogate_stats(self):
        """(cells, cells evaluated exactly, stages that fell back, largest sampled error)."""
        a, b, c, d = ctypes.c_long(), ctypes.c_long(), ctypes.c_long(), ctypes.c_double()
        _check(lib().spbte_run_surrogate_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                               ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def open_output(self, path, flags=OUT_ALL, append=False):
        _check(lib().spbte_run_open_output(self.h, os.fsencode(path), int(flags), int(append)))

    def write_output(self, time):
        _check(lib().spbte_run_write_output(self.h, float(time)))

    def close_output(self):
        _check(lib().spbte_run_close_output(self.h))

    def store_restart(self, pdf_path, time_path=None, step=0):
        _check(lib().spbte_run_store_restart(self.h, os.fsencode(pdf_path),
                                             os.fsencode(time_path) if time_path else None, int(step)))

    def load_restart(self, pdf_path, time_path=None):
        step = ctypes.c_int(0)
        _check(lib().spbte_run_load_restart(self.h, os.fsencode(pdf_path),
                                            os.fsencode(time_path) if time_path else None, ctypes.byref(step)))
        return step.value

    def moments(self):
        """(ncell, 5): rho, u_x, u_y, u_z, T per cell (unit mass); (Ns, ncell, 5) for several species."""
        out = np.empty((self.ns, self.ncell, 5))
        _check(lib().spbte_run_moments(self.h, _p(out)))
        return out[0] if self.ns == 1 else out

    def open_output_species(self, l, path, flags=OUT_ALL, append=False):
        _check(lib().spbte_run_open_output_species(self.h, int(l), os.fsencode(path), int(flags), int(append)))

    def store_restart_species(self, l, pdf_path):
        _check(lib().spbte_run_store_restart_species(self.h, int(l), os.fsencode(pdf_path)))

    def load_restart_species(self, l, pdf_path):
        _check(lib().spbte_run_load_restart_species(self.h, int(l), os.fsencode(pdf_path)))

    def close(self):
        if self.h:
            lib().spbte_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advect(op, f, x, dx, order, dt, bc, T0=1.0, T1=2.0):
    """One transport stage on the GPU (upwindOne or advectTwo): f is
    (nX + 2*order, N^3) with ghost cells; returns (f with its ghosts as the
    stage left them, fc)."""
    f = _c(f)
    nX = f.shape[0] - 2 * order
    d_f = DeviceArray.from_host(f)
    d_fc = DeviceArray(f.shape, np.zeros_like(f))
    _check(lib().spbte_advect(op.h, d_f.ptr, d_fc.ptr, nX, order, _p(_c(x)), _p(_c(dx)), dt, bc, T0, T1))
    return d_f.download(), d_fc.download()


# ------------------------------------------------------------ learned surrogate
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(_fp)


class Net:
    """spbte_net: the spectral mixing surrogate of the collision operator of
    ``op`` on the kept block of Kb^3 modes (see include/spbte.h)."""

    def __init__(self, op, Kb, width=64, rank=128, layers=4, maxpreserve=False, seed=1):
        self.op = op
        self.Kb, self.C, self.r, self.L = Kb, width, rank, layers
        self.M = Kb ** 3
        self.h = _vp()
        _check(lib().spbte_net_create(ctypes.byref(self.h), op.h, Kb, width, rank, layers, int(maxpreserve), seed))
        self.npar = lib().spbte_net_num_params(self.h)

    def params(self):
        out = np.empty(self.npar, dtype=np.float32)
        _check(lib().spbte_net_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.npar
        _check(lib().spbte_net_set_params(self.h, _fptr(p)))

    def offsets(self):
        o = (ctypes.c_long * (3 + 6 * self.L + 2))()
        k = lib().spbte_net_layout(self.h, o)
        return list(o)[:k]

    def scales(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().spbte_net_get_scales(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_scales(self, fscale, tscale):
        _check(lib().spbte_net_set_scales(self.h, fscale, tscale))

    def calibrate(self, d_f, d_Q, ncell):
        _check(lib().spbte_net_calibrate(self.h, d_f, d_Q, int(ncell)))

    def reserve(self, nsamples):
        _check(lib().spbte_net_dataset_reserve(self.h, int(nsamples)))

    def add(self, d_f, d_Q, ncell):
        _check(lib().spbte_net_dataset_add(self.h, d_f, d_Q, int(ncell)))

    def dataset_size(self):
        return lib().spbte_net_dataset_size(self.h)

    def train(self, nsteps, batch, lr, seed=7):
        losses = np.zeros(nsteps, dtype=np.float32)
        _check(lib().spbte_net_train(self.h, int(nsteps), int(batch), float(lr), seed, _fptr(losses)))
        return losses

    def grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.empty(self.npar, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().spbte_net_grad(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), idx.size, _fptr(g),
                                    ctypes.byref(loss)))
        return loss.value, g

    def eval(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        e = np.empty(idx.size, dtype=np.float32)
        _check(lib().spbte_net_eval(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), idx.size, _fptr(e)))
        return e

    def forward_features(self, X):
        X = _f32(X)
        B = X.shape[1]
        Y = np.empty((self.M, B, 2), dtype=np.float32)
        _check(lib().spbte_net_forward_features(self.h, _fptr(X), B, _fptr(Y)))
        return Y

    def grad_features(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.empty(self.npar, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().spbte_net_grad_features(self.h, _fptr(X), _fptr(T), X.shape[1], _fptr(g), ctypes.byref(loss)))
        return loss.value, g

    def predict_dev(self, d_f, d_Q, ncell, conserve=True):
        _check(lib().spbte_net_predict(self.h, d_f, d_Q, int(ncell), int(conserve)))

    def collide_dev(self, d_f, d_Q, ncell, form="fp64", check_every=8, phase=0, tol=1e-3, floor_rel=0.0):
        """Surrogate-assisted Q with the sampled exact check; returns (cells evaluated exactly, max sampled error)."""
        ne, err = ctypes.c_int(), ctypes.c_double()
        _check(lib().spbte_net_collide(self.h, d_f, d_Q, int(ncell), FORMS[form], int(check_every), int(phase),
                                       float(tol), float(floor_rel), ctypes.byref(ne), ctypes.byref(err)))
        return ne.value, err.value

    def set_loss_floor(self, rms_floor):
        _check(lib().spbte_net_set_loss_floor(self.h, float(rms_floor)))

    def predict(self, f, conserve=True):
        f = _c(f).reshape(-1, self.op.N ** 3)
        d_f = DeviceArray.from_host(f)
        d_Q = DeviceArray(f.shape)
        self.predict_dev(d_f.ptr, d_Q.ptr, f.shape[0], conserve)
        self.op.sync()
        return d_Q.download()

    def close(self):
        if self.h:
            lib().spbte_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------ reference interface (numpy)
class Species(ctypes.Structure):
    """Binary image of the reference's species struct (src/species.h:11-26)."""
    _fields_ = [("id", ctypes.c_size_t), ("num_levels", ctypes.c_size_t), ("lev_id", ctypes.c_void_p),
                ("Rgas", ctypes.c_double), ("mass", ctypes.c_double), ("mm", ctypes.c_double),
                ("d_ref", ctypes.c_double), ("T_ref", ctypes.c_double), ("mu_ref", ctypes.c_double),
                ("omega", ctypes.c_double), ("E0", ctypes.c_double), ("Ei", ctypes.c_void_p),
                ("gi", ctypes.c_void_p), ("name", ctypes.c_char * 80)]


def default_species(i=0):
    """src/species.c:28-43 defaults."""
    s = Species()
    s.id = i
    s.Rgas = 1.0
    s.mass = 1.0
    s.mm = 1.0
    s.d_ref = 2.0
    s.T_ref = 1.0
    s.mu_ref = 1.0
    s.E0 = 0.0
    s.name = b"default"
    return s


class WeightRows:
    """A contiguous (N^3, N^3) numpy table seen as the reference's double** rows."""

    def __init__(self, W):
        self.W = _c(W)
        n3 = self.W.shape[0]
        base = self.W.ctypes.data
        self.rows = (_dp * n3)(*[ctypes.cast(base + 8 * n3 * r, _dp) for r in range(n3)])

    @property
    def ptr(self):
        return ctypes.cast(self.rows, ctypes.POINTER(_dp))


class _RefState:
    v = None
    eta = None
    cons_v = None
    species = None


_ref = _RefState()


def initialize_coll(N, L_v, v, eta):
    """src/collisions.h:27.  v and eta are kept alive here (the C side borrows them)."""
    _ref.v = _c(v)
    _ref.eta = _c(eta)
    lib().initialize_coll(int(N), float(L_v), _p(_ref.v), _p(_ref.eta))


def dealloc_coll():
    lib().dealloc_coll()


def ComputeQ(f, g, W):
    """src/collisions.h:38 with a WeightRows (or contiguous table)."""
    rows = W if isinstance(W, WeightRows) else WeightRows(W)
    f = _c(f)
    g = f if g is f else _c(g)
    Q = np.zeros_like(f)
    lib().ComputeQ(_p(f), _p(g), _p(Q), rows.ptr)
    return Q


def ComputeQ_maxPreserve(f, g, W):
    rows = W if isinstance(W, WeightRows) else WeightRows(W)
    f = _c(f)
    g = f if g is f else _c(g)
    Q = np.zeros_like(f)
    lib().ComputeQ_maxPreserve(_p(f), _p(g), _p(Q), rows.ptr)
    return Q


def fft3D(x, invert):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    out = np.empty_like(a)
    lib().fft3D(_p(a.view(np.float64)), _p(out.view(np.float64)), int(invert))
    return out


def initialize_conservation(N, h_v, v, masses=(1.0,)):
    _ref.cons_v = _c(v)
    sp = (Species * len(masses))()
    for i, m in enumerate(masses):
        sp[i] = default_species(i)
        sp[i].mass = m
    _ref.species = sp
    lib().initialize_conservation(int(N), float(h_v), _p(_ref.cons_v), ctypes.cast(sp, _vp), len(masses))


def dealloc_conservation():
    lib().dealloc_conservation()


def conserveAllMoments(Qs):
    """Qs: list of Ns*Ns arrays (projected copies are returned)."""
    arrs = [_c(q).copy() for q in Qs]
    ptrs = (_dp * len(arrs))(*[_p(a) for a in arrs])
    lib().conserveAllMoments(ctypes.cast(ptrs, ctypes.POINTER(_dp)))
    return arrs


def solveWithCCt(b):
    b = _c(b).copy()
    lib().solveWithCCt(len(b), _p(b))
    return b


# ------------------------------------------------ boltz_ front end (spbte_boltz.h)
class Input(ctypes.Structure):
    """Binary image of spbte_input: the keywords of a reference input file."""
    _fields_ = [("N", ctypes.c_int), ("L_v", ctypes.c_double), ("Kn", ctypes.c_double), ("lambda_", ctypes.c_double),
                ("dt", ctypes.c_double), ("nT", ctypes.c_int), ("order", ctypes.c_int), ("data_freq", ctypes.c_int),
                ("restart", ctypes.c_int), ("restart_time", ctypes.c_double), ("init_flag", ctypes.c_int),
                ("bc_flag", ctypes.c_int), ("inhom", ctypes.c_int), ("weight_flag", ctypes.c_int),
                ("aniso", ctypes.c_int), ("mesh_file", ctypes.c_char * 80), ("num_species", ctypes.c_int),
                ("species_names", (ctypes.c_char * 80) * 8)]


def read_input(path):
    """The keywords of a reference input file (src/input.c:10-183)."""
    inp = Input()
    _check(lib().spbte_read_input(os.fsencode(path), ctypes.byref(inp)))
    return inp


def read_species(path):
    """A species file (src/species.c:64-159) over the defaults of :28-43."""
    sp = Species()
    lib().spbte_species_defaults(ctypes.byref(sp), 0)
    _check(lib().spbte_read_species(os.fsencode(path), ctypes.byref(sp)))
    return sp


def read_mesh(path, order):
    """(x, dx, dx_min) of a mesh file (src/mesh_setup.c:9-200), nX + 2*order entries."""
    n = ctypes.c_int(0)
    _check(lib().spbte_read_mesh(os.fsencode(path), int(order), ctypes.byref(n), None, None, None))
    x, dx, dmin = np.empty(n.value + 2 * order), np.empty(n.value + 2 * order), ctypes.c_double()
    _check(lib().spbte_read_mesh(os.fsencode(path), int(order), ctypes.byref(n), _p(x), _p(dx), ctypes.byref(dmin)))
    return x, dx, dmin.value


def velocity_grid(N, L_v, inhom):
    """(v, eta) of initialize_hom (inhom=0) or initialize_inhom (inhom=1)."""
    v, eta = np.empty(N), np.empty(N)
    lib().spbte_velocity_grid(int(N), float(L_v), int(inhom), _p(v), _p(eta))
    return v, eta


def initial_state(inp, species, v, nX=1):
    """The Init_field state of an Input: (N^3,) for 0D, (nX, N^3) for 1D."""
    n3 = inp.N ** 3
    f = np.empty((nX if inp.inhom else 1) * n3)
    _check(lib().spbte_initial_state(ctypes.byref(inp), ctypes.byref(species), _p(_c(v)), int(nX), _p(f)))
    return f.reshape(nX, n3) if inp.inhom else f


def boltz(input_name, flags_name, workdir="."):
    """Runs spbte_boltz_main (exec/boltz.c) in workdir, which holds input/,
    Weights/ and Data/ as for the reference's boltz_.  Raises on failure."""
    prev = os.getcwd()
    argv = (ctypes.c_char_p * 3)(b"boltz_amd", os.fsencode(input_name), os.fsencode(flags_name))
    os.chdir(workdir)
    try:
        rc = lib().spbte_boltz_main(3, argv)
    finally:
        os.chdir(prev)
    if rc != 0:
        raise SpbteError(lib().spbte_last_error().decode())


def weightgen(N, glance, beta, lam, L_v, workdir="."):
    """Runs spbte_weightgen_main (exec/MPIWeightGenerator.c) in workdir, which
    holds Weights/.  Raises on failure."""
    prev = os.getcwd()
    args = [b"weightgen_amd"] + [str(a).encode() for a in (N, glance, beta, lam, L_v)]
    argv = (ctypes.c_char_p * len(args))(*args)
    os.chdir(workdir)
    try:
        rc = lib().spbte_weightgen_main(len(args), argv)
    finally:
        os.chdir(prev)
    if rc != 0:
        raise SpbteError(lib().spbte_last_error().decode())
