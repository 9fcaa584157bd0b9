"""ctypes loader for the FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps oracle/liboracle.so (the checker), liboracle_fast.so
(the timed OpenMP baseline) and, when present, oracle/_ref/libref_driver.so
(the reference engine compiled from its own sources).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_U64 = C.c_uint64
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

PROX_FN = C.CFUNCTYPE(None, _DP, _DP, _DP, _I, _I, C.c_void_p)

_libs = {}


def _bind(lib):
    sig = {
        "or_new": (_P, [_I]),
        "or_free": (None, [_P]),
        "or_add_factor": (None, [_P, _P, _P, _I, _IP]),
        "or_alloc_state": (None, [_P]),
        "or_set_rho_alpha": (None, [_P, _D, _D]),
        "or_update_XM": (None, [_P]), "or_update_X": (None, [_P]), "or_update_M": (None, [_P]),
        "or_update_Z": (None, [_P]), "or_update_UN": (None, [_P]), "or_update_U": (None, [_P]),
        "or_update_N": (None, [_P]),
        "or_iterate": (None, [_P, _I, _DP]),
        "or_iterate_omp": (_D, [_P, _I, _I]),
        "or_iterate_par": (None, [_P, _I, _I, _DP]),
        "or_max_alpha_radius": (_D, [_D, _D, _D]),
        "or_splitmix64": (_U64, [C.POINTER(_U64)]),
        "or_uniform01": (_D, [C.POINTER(_U64)]),
        "or_build_square": (_P, [_I, _U64, _D, _D, _D]),
        "or_build_triangle": (_P, [_I, _U64, _D, _D]),
        "or_residuals": (None, [_P, _DP, _DP]),
        "or_warm_start": (None, [_P, _DP, _DP, _I]),
        "or_square_quality": (None, [_P, _I, _DP, _DP, _DP]),
        "or_nvar": (_I, [_P]), "or_nedge": (_I, [_P]), "or_nfac": (_I, [_P]), "or_ndims": (_I, [_P]),
        "or_ptr": (_DP, [_P, _I]),
        "or_edge_var": (_IP, [_P]), "or_fac_first": (_IP, [_P]), "or_fac_deg": (_IP, [_P]),
        "or_var_deg": (_I, [_P, _I]), "or_var_edges": (_IP, [_P, _I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    for name in ("or_prox_pair2", "or_prox_wall2", "or_prox_radius2",
                 "or_prox_pair3", "or_prox_box3", "or_prox_radius3"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [_DP, _DP, _DP, _I, _I, C.c_void_p]
    return lib


def load(fast=False):
    """The oracle library; fast=True gives the -Ofast OpenMP build used for timing."""
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        _libs[name] = _bind(C.CDLL(path))
    return _libs[name]


def build_native(outdir):
    """Compiles the oracle with the reference's build flags (-Ofast) plus this
    host's instruction set (-march=native) and OpenMP into outdir, and loads
    it: the timed CPU baseline, built on the machine it runs on."""
    import subprocess
    path = os.path.join(outdir, "liboracle_native.so")
    subprocess.run(["gcc", "-std=c99", "-Ofast", "-march=native", "-fopenmp", "-fPIC", "-shared", "-o", path,
                    os.path.join(HERE, "paradmm_oracle.c"), "-lm"], check=True)
    return _bind(C.CDLL(path))


def load_ref():
    """The reference engine compiled from source (this container only), or None."""
    path = os.path.join(HERE, "_ref", "libref_driver.so")
    if not os.path.exists(path):
        return None
    if "ref" not in _libs:
        lib = C.CDLL(path)
        lib.ref_square_run.restype = _I
        lib.ref_square_run.argtypes = [_I, _U64, _D, _D, _D, _I, _DP]
        lib.ref_triangle_run.restype = _I
        lib.ref_triangle_run.argtypes = [_I, _U64, _D, _D, _I, _DP]
        lib.ref_square_run_omp.restype = _I
        lib.ref_square_run_omp.argtypes = [_I, _U64, _D, _D, _D, _I, _I, _I, _DP]
        lib.ref_triangle_run_omp.restype = _I
        lib.ref_triangle_run_omp.argtypes = [_I, _U64, _D, _D, _I, _I, _I, _DP]
        lib.ref_square_warm_start.restype = _I
        lib.ref_square_warm_start.argtypes = [_I, _D, _D, _DP, _DP, _I, _DP, _DP, _DP, _DP, _DP]
        lib.ref_triangle_warm_start.restype = _I
        lib.ref_triangle_warm_start.argtypes = [_I, _DP, _DP, _I, _DP, _DP, _DP, _DP, _DP]
        for name in ("ref_prox_pair2", "ref_prox_wall2", "ref_prox_radius2"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [_DP, _DP, _DP, _I, _I, C.c_void_p]
        _libs["ref"] = lib
    return _libs["ref"]


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def call_prox(fn, x, n, rhos, nvars, ndims, params=None):
    """Calls a prox with the reference signature on numpy float64 arrays."""
    p = None
    if params is not None:
        p = params.ctypes.data_as(C.c_void_p)
    fn(dptr(x), dptr(n), dptr(rhos), nvars, ndims, p)


class Graph:
    """An oracle graph (factor-major edges, reference builder semantics)."""

    def __init__(self, handle, lib):
        self.h = handle
        self.lib = lib

    @classmethod
    def square(cls, ncircles, seed, rho=1.0, alpha=1.0, a_exp=1.0, fast=False):
        lib = load(fast)
        return cls(lib.or_build_square(ncircles, seed, rho, alpha, a_exp), lib)

    @classmethod
    def triangle(cls, ncircles, seed, lo=-1.0, hi=2.0):
        lib = load()
        return cls(lib.or_build_triangle(ncircles, seed, lo, hi), lib)

    def close(self):
        if self.h:
            self.lib.or_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nvar(self):
        return self.lib.or_nvar(self.h)

    @property
    def nedge(self):
        return self.lib.or_nedge(self.h)

    @property
    def nfac(self):
        return self.lib.or_nfac(self.h)

    @property
    def ndims(self):
        return self.lib.or_ndims(self.h)

    _WHICH = {"X": 0, "M": 1, "N": 2, "U": 3, "Z": 4, "RHO": 5, "ALPHA": 6, "Zprev": 7}

    def array(self, name):
        """A numpy view of one state array (shares memory with the oracle)."""
        w = self._WHICH[name]
        if name in ("Z", "Zprev"):
            n = self.nvar * self.ndims
        elif name in ("RHO", "ALPHA"):
            n = self.nedge
        else:
            n = self.nedge * self.ndims
        p = self.lib.or_ptr(self.h, w)
        return np.ctypeslib.as_array(p, shape=(n,))

    def edge_var(self):
        return np.ctypeslib.as_array(self.lib.or_edge_var(self.h), shape=(self.nedge,)).copy()

    def iterate(self, iters, record=False):
        if record:
            hist = np.zeros((iters, self.nvar * self.ndims))
            self.lib.or_iterate(self.h, iters, dptr(hist))
            return hist
        self.lib.or_iterate(self.h, iters, None)
        return None

    def iterate_omp(self, iters, nthreads):
        return self.lib.or_iterate_omp(self.h, iters, nthreads)

    def iterate_par(self, iters, nthreads=0, record=False):
        """iterate() through the OpenMP passes (same results; nthreads 0:
        the OpenMP default)."""
        if record:
            hist = np.zeros((iters, self.nvar * self.ndims))
            self.lib.or_iterate_par(self.h, iters, nthreads, dptr(hist))
            return hist
        self.lib.or_iterate_par(self.h, iters, nthreads, None)
        return None

    def residuals(self):
        p, d = C.c_double(), C.c_double()
        self.lib.or_residuals(self.h, C.byref(p), C.byref(d))
        return p.value, d.value

    def square_quality(self, ncircles):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.lib.or_square_quality(self.h, ncircles, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value
