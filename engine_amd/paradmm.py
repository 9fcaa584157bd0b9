"""Python mirror of the parADMM C API (paradmm.h + paradmm_gpu.h) over ctypes.

The product is libparadmm.so (C ABI, HIP kernels for gfx950).  This module
only binds it: the `graph` struct of core/paradmm.h:33-63, the legacy entry
points (startG, addNode, initialize_*, update*) and the MI355X extension
(paradmm_gpu_*).  It raises at import time if the library is not built; there
is no fallback.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PARADMM_LIB") or os.path.join(HERE, "libparadmm.so")

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_FP = C.POINTER(C.c_float)

PROX_FN = C.CFUNCTYPE(None, _DP, _DP, _DP, C.c_int, C.c_int, C.c_void_p)

KIND_PAIR_2D, KIND_WALL_2D, KIND_RADIUS_2D, KIND_PAIR_3D, KIND_BOX_3D, KIND_RADIUS_3D = 1, 2, 3, 4, 5, 6


class GraphStruct(C.Structure):
    """`graph` of core/paradmm.h:33-63, field for field."""
    _fields_ = [
        ("num_opts", C.c_int),
        ("num_vars", C.c_int),
        ("num_dims", C.c_int),
        ("num_edges", C.c_int),
        ("list_of_operators_functions_pointers", C.POINTER(C.c_void_p)),
        ("list_of_operators_param_pointers", C.POINTER(C.c_void_p)),
        ("list_of_operators_param_sizes", _IP),
        ("operatorix_to_numedges", _IP),
        ("operatorix_to_edgelistix", C.POINTER(_IP)),
        ("varix_to_numedges", _IP),
        ("varix_to_edgelistix", C.POINTER(_IP)),
        ("edgeix_to_varlistix", _IP),
        ("X", _DP), ("M", _DP), ("N", _DP), ("U", _DP),
        ("Z", _DP),
        ("RHOS", _DP), ("ALPHAS", _DP),
    ]


class SolveOpts(C.Structure):
    """paradmm_solve_opts of paradmm_gpu.h"""
    _fields_ = [("max_iters", C.c_int), ("check_every", C.c_int), ("tol_primal", C.c_double),
                ("tol_dual", C.c_double), ("tol_feas", C.c_double), ("policy_every", C.c_int)]


class SolveResult(C.Structure):
    """paradmm_solve_result of paradmm_gpu.h"""
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("primal_rms", C.c_double),
                ("dual_rms", C.c_double), ("max_overlap", C.c_double), ("max_box", C.c_double),
                ("objective", C.c_double), ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchConfig(C.Structure):
    """paradmm_batch_config of paradmm_gpu.h"""
    _fields_ = [("tile", C.c_int), ("blocks_per_cu", C.c_int), ("barrier", C.c_int), ("ms_per_iter", C.c_double),
                ("contact_tile", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ParadmmError(RuntimeError):
    pass


_lib = None

# Every live handle, so that they are released while the HIP runtime (and a
# profiler attached to it) is still up: at interpreter exit an atexit hook
# closes batches, then policies, then graphs, and __del__ makes no library
# call once the interpreter is finalising.
_live = {"batch": weakref.WeakSet(), "mlp": weakref.WeakSet(), "graph": weakref.WeakSet()}


def release_all():
    """Closes every live Batch, RhoMLP and Graph (batches first: they use graphs)."""
    for kind in ("batch", "mlp", "graph"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(release_all)


def load():
    """Loads libparadmm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C engine_amd` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)  # RTLD_LOCAL: the API names also exist in the reference build
    G = C.POINTER(GraphStruct)
    V = C.c_void_p
    sig = {
        "abs_d": (C.c_double, [C.c_double]),
        "quicksort": (None, [_IP, C.c_int]),
        "get_random": (C.c_double, [C.c_double, C.c_double]),
        "startG": (None, [G, C.c_int]),
        "closeG": (None, [G]),
        "printG": (None, [G]),
        "initialize_RHOS_APHAS": (None, [G, C.c_double, C.c_double]),
        "initialize_X_N_Z_M_U_zeros": (None, [G]),
        "initialize_X_N_Z_M_U_rand": (None, [G] + [C.c_double] * 10),
        "initialize_X_N_Z_M_U_optimal": (None, [G, _DP, _DP, C.c_int]),
        "addNode": (None, [G, V, V, C.c_int, _IP]),
        "addNodeWithParamSize": (None, [G, V, V, C.c_int, C.c_int, _IP]),
        "update_outgoing_rhos": (None, [G]),
        "updateXM": (None, [G]), "updateX": (None, [G]), "updateM": (None, [G]),
        "updateZ": (None, [G]), "updateUN": (None, [G]), "updateU": (None, [G]),
        "updateN": (None, [G]), "updateG": (None, [G]),
        "updateXOpenMP": (None, [G]), "updateMOpenMP": (None, [G]), "updateZOpenMP": (None, [G]),
        "updateUOpenMP": (None, [G]), "updateNOpenMP": (None, [G]),
        "updateGOpenMP": (None, [G, C.c_int, C.c_int]),
        "paradmm_last_error": (C.c_char_p, []),
        "paradmm_register_prox_kind": (C.c_int, [V, C.c_int]),
        "paradmm_prox_kind_of": (C.c_int, [V]),
        "paradmm_packing_square": (C.c_int, [G, C.c_int, C.c_double]),
        "paradmm_packing_init_square": (C.c_int, [G, C.c_uint64]),
        "paradmm_init_rand_seeded": (C.c_int, [G, C.c_uint64] + [C.c_double] * 10),
        "paradmm_gpu_prepare": (C.c_int, [G]),
        "paradmm_gpu_iterate": (C.c_int, [G, C.c_int]),
        "paradmm_gpu_sync_to_host": (C.c_int, [G]),
        "paradmm_gpu_sync_from_host": (C.c_int, [G]),
        "paradmm_gpu_read": (C.c_int, [G, C.c_int, _DP]),
        "paradmm_gpu_release": (C.c_int, [G]),
        "paradmm_gpu_layout": (C.c_int, [G]),
        "paradmm_gpu_contact_info": (C.c_int, [G, _IP, _IP, C.POINTER(C.c_long)]),
        "paradmm_gpu_residuals": (C.c_int, [G, _DP, _DP]),
        "paradmm_gpu_time_iterations": (C.c_int, [G, C.c_int, _DP]),
        "paradmm_gpu_profile_kernels": (C.c_int, [G, C.c_int, C.c_int, C.c_char_p, _DP, _IP, _IP]),
        "paradmm_gpu_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_int, _DP]),
        "paradmm_gpu_solve": (C.c_int, [G, C.POINTER(SolveOpts), C.POINTER(SolveResult)]),
        "paradmm_gpu_square_quality": (C.c_int, [G, _DP, _DP, _DP]),
        "paradmm_rho_mlp_create": (C.c_int, [C.POINTER(C.c_void_p), _FP, C.c_int, C.c_uint64]),
        "paradmm_rho_mlp_destroy": (None, [V]),
        "paradmm_gpu_set_rho_policy_mode": (C.c_int, [G, C.c_int, C.c_int]),
        "paradmm_add_nodes": (C.c_int, [G, V, V, C.c_size_t, C.c_int, C.c_int, _IP]),
        "paradmm_save_state": (C.c_int, [G, C.c_char_p]),
        "paradmm_load_state": (C.c_int, [G, C.c_char_p]),
        "paradmm_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(G), C.c_int]),
        "paradmm_batch_destroy": (None, [V]),
        "paradmm_batch_configure": (C.c_int, [V, C.c_int, C.c_int, C.c_int]),
        "paradmm_batch_autotune": (C.c_int, [V, C.c_int, C.POINTER(BatchConfig), C.POINTER(BatchConfig), C.c_int,
                                             _IP]),
        "paradmm_batch_solve": (C.c_int, [V, C.POINTER(SolveOpts), C.POINTER(SolveResult), _DP]),
        "paradmm_batch_sync_to_host": (C.c_int, [V]),
        "paradmm_batch_set_rho_policy": (C.c_int, [V, V]),
        "paradmm_batch_set_rho_policy_mode": (C.c_int, [V, C.c_int, C.c_int]),
        "paradmm_batch_info": (C.c_int, [V, _IP, _IP, _IP, C.POINTER(C.c_long), _IP]),
        "paradmm_batch_configure_contact": (C.c_int, [V, C.c_int, C.c_int, C.c_int]),
        "paradmm_batch_contact_info": (C.c_int, [V, _IP, _IP, _IP, C.POINTER(C.c_long), C.POINTER(C.c_long), _DP]),
        "paradmm_rho_mlp_get_params": (C.c_int, [V, _FP, C.c_int]),
        "paradmm_rho_mlp_forward": (C.c_int, [V, _FP, C.c_int, _FP]),
        "paradmm_rho_mlp_gradient": (C.c_int, [V, _FP, _FP, C.c_int, _FP, _DP]),
        "paradmm_rho_mlp_train": (C.c_int, [V, _FP, _FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, _DP]),
        "paradmm_rho_mlp_layout_probe": (C.c_int, [_FP]),
        "paradmm_gpu_rho_features": (C.c_int, [G, C.c_float, _FP]),
        "paradmm_gpu_set_rho_policy": (C.c_int, [G, V]),
        "paradmm_gpu_apply_rho_policy": (C.c_int, [G, V, C.c_float]),
        "paradmm_gpu_scale_rho": (C.c_int, [G, C.c_double]),
        "paradmm_gpu_state_save": (C.c_int, [G]),
        "paradmm_gpu_state_restore": (C.c_int, [G]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names the C headers declare (checked against the library's exports by the tests)."""
    return [
        "abs_d", "quicksort", "get_random", "realloc2", "startG", "closeG", "printG",
        "initialize_RHOS_APHAS", "initialize_X_N_Z_M_U_zeros", "initialize_X_N_Z_M_U_rand",
        "initialize_X_N_Z_M_U_optimal", "addNode", "addNodeWithParamSize", "update_outgoing_rhos",
        "updateXM", "updateX", "updateM", "updateZ", "updateUN", "updateU", "updateN", "updateG",
        "updateXOpenMP", "updateMOpenMP", "updateZOpenMP", "updateUOpenMP", "updateNOpenMP", "updateGOpenMP",
    ]


def prox_address(name):
    """Address of one of the library's built-in prox operators (a device kind token)."""
    lib = load()
    return C.cast(getattr(lib, name), C.c_void_p).value


def copy_bandwidth(nbytes=1 << 30, iters=20):
    """GB/s of a plain device copy (read + write counted)."""
    gbs = C.c_double()
    check(load().paradmm_gpu_copy_bandwidth(nbytes, iters, C.byref(gbs)), "paradmm_gpu_copy_bandwidth")
    return gbs.value


def _fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


class RhoMLP:
    """The learned rho policy network (paradmm_rho_mlp of paradmm_gpu.h)."""
    NPARAMS = 1313

    def __init__(self, params=None, seed=0):
        self.lib = load()
        h = C.c_void_p()
        if params is None:
            check(self.lib.paradmm_rho_mlp_create(C.byref(h), None, 0, seed), "paradmm_rho_mlp_create")
        else:
            p = np.ascontiguousarray(params, dtype=np.float32)
            check(self.lib.paradmm_rho_mlp_create(C.byref(h), _fptr(p), p.size, seed), "paradmm_rho_mlp_create")
        self.h = h
        _live["mlp"].add(self)

    def params(self):
        p = np.zeros(self.NPARAMS, dtype=np.float32)
        check(self.lib.paradmm_rho_mlp_get_params(self.h, _fptr(p), p.size), "get_params")
        return p

    def forward(self, feat):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        y = np.zeros(f.shape[0], dtype=np.float32)
        check(self.lib.paradmm_rho_mlp_forward(self.h, _fptr(f), f.shape[0], _fptr(y)), "forward")
        return y

    def gradient(self, feat, target):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.NPARAMS, dtype=np.float32)
        loss = C.c_double()
        check(self.lib.paradmm_rho_mlp_gradient(self.h, _fptr(f), _fptr(t), f.shape[0], _fptr(g), C.byref(loss)),
              "gradient")
        return loss.value, g

    def train(self, feat, target, epochs, batch=4096, lr=1e-3, seed=0):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        check(self.lib.paradmm_rho_mlp_train(self.h, _fptr(f), _fptr(t), f.shape[0], epochs, batch, lr, seed,
                                             C.byref(loss)), "train")
        return loss.value

    def close(self):
        if self.h:
            self.lib.paradmm_rho_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def mfma_layout_probe():
    out = np.zeros(256, dtype=np.float32)
    check(load().paradmm_rho_mlp_layout_probe(_fptr(out)), "layout probe")
    return out.reshape(16, 16)


def check(rc, what=""):
    if rc != 0:
        msg = load().paradmm_last_error().decode(errors="replace")
        raise ParadmmError(f"{what}: {msg}" if what else msg)


class Graph:
    """Owns one `graph` struct of libparadmm.  Arrays are numpy views on the
    library-owned host arrays (edge-major AoS, as in the reference)."""

    def __init__(self, ndims):
        self.lib = load()
        self.g = GraphStruct()
        self.ptr = C.pointer(self.g)
        self.lib.startG(self.ptr, ndims)
        self._keep = []  # parameter blocks referenced by factors
        _live["graph"].add(self)

    # ---- construction
    def add_factor(self, fn_addr, variables, params=None):
        vars_arr = (C.c_int * len(variables))(*variables)
        p = None
        if params is not None:
            buf = (C.c_double * len(params))(*params)
            self._keep.append(buf)
            p = C.cast(buf, C.c_void_p)
        self.lib.addNode(self.ptr, C.c_void_p(fn_addr), p, len(variables), vars_arr)

    def packing_square(self, ncircles, a_exp=1.0):
        check(self.lib.paradmm_packing_square(self.ptr, ncircles, a_exp), "paradmm_packing_square")

    def add_factors(self, fn_addr, variables, params=None):
        """Many factors of one operator and degree (paradmm_add_nodes):
        variables is a (count, nvars) int array; params an optional
        (count, k) float64 array, one row per factor."""
        v = np.ascontiguousarray(variables, dtype=np.int32)
        p, stride = None, 0
        if params is not None:
            pa = np.ascontiguousarray(params, dtype=np.float64)
            self._keep.append(pa)
            p, stride = pa.ctypes.data_as(C.c_void_p), pa.strides[0] if pa.ndim > 1 else 0
        check(self.lib.paradmm_add_nodes(self.ptr, fn_addr, p, stride, v.shape[0], v.shape[1],
                                         v.ctypes.data_as(_IP)), "paradmm_add_nodes")

    def save_state(self, path):
        check(self.lib.paradmm_save_state(self.ptr, os.fsencode(path)), "paradmm_save_state")

    def load_state(self, path):
        check(self.lib.paradmm_load_state(self.ptr, os.fsencode(path)), "paradmm_load_state")

    def init_rho_alpha(self, rho, alpha):
        self.lib.initialize_RHOS_APHAS(self.ptr, rho, alpha)

    def init_zeros(self):
        self.lib.initialize_X_N_Z_M_U_zeros(self.ptr)

    def init_rand_seeded(self, seed, X=(0.0, 1.0), N=(0.0, 1.0), Z=(0.0, 1.0), M=(0.0, 1.0), U=(0.0, 1.0)):
        """initialize_X_N_Z_M_U_rand with a seed (paradmm_init_rand_seeded);
        each argument is a (low, high) bound pair."""
        check(self.lib.paradmm_init_rand_seeded(self.ptr, seed, X[0], X[1], N[0], N[1], Z[0], Z[1], M[0], M[1],
                                                U[0], U[1]), "paradmm_init_rand_seeded")

    def init_square(self, seed):
        check(self.lib.paradmm_packing_init_square(self.ptr, seed), "paradmm_packing_init_square")

    # ---- views
    @property
    def nedge(self):
        return self.g.num_edges

    @property
    def nvar(self):
        return self.g.num_vars

    @property
    def nfac(self):
        return self.g.num_opts

    @property
    def ndims(self):
        return self.g.num_dims

    def array(self, name):
        p = getattr(self.g, name)
        if not p:
            raise ParadmmError(f"{name} is not allocated")
        if name == "Z":
            n = self.nvar * self.ndims
        elif name in ("RHOS", "ALPHAS"):
            n = self.nedge
        else:
            n = self.nedge * self.ndims
        return np.ctypeslib.as_array(p, shape=(n,))

    def edge_var(self):
        return np.ctypeslib.as_array(self.g.edgeix_to_varlistix, shape=(self.nedge,)).copy()

    # ---- legacy passes
    def legacy(self, name, *args):
        getattr(self.lib, name)(self.ptr, *args)

    # ---- device
    def prepare(self):
        check(self.lib.paradmm_gpu_prepare(self.ptr), "paradmm_gpu_prepare")

    def iterate(self, iters):
        check(self.lib.paradmm_gpu_iterate(self.ptr, iters), "paradmm_gpu_iterate")

    def sync_to_host(self):
        check(self.lib.paradmm_gpu_sync_to_host(self.ptr), "paradmm_gpu_sync_to_host")

    def sync_from_host(self):
        check(self.lib.paradmm_gpu_sync_from_host(self.ptr), "paradmm_gpu_sync_from_host")

    _READ = {"Z": 0, "X": 1, "U": 2, "RHO": 3}

    def read(self, name):
        """One device array as float64 (paradmm_gpu_read): Z (V, D); X, U
        (E, D) in edge order; RHO (E,).  The host arrays are not touched."""
        w = self._READ[name]
        shape = (self.nvar, self.ndims) if name == "Z" else (self.nedge,) if name == "RHO" else (self.nedge, self.ndims)
        out = np.empty(shape, dtype=np.float64)
        check(self.lib.paradmm_gpu_read(self.ptr, w, out.ctypes.data_as(_DP)), "paradmm_gpu_read")
        return out

    def layout(self):
        return self.lib.paradmm_gpu_layout(self.ptr)

    def contact_info(self):
        """(in the contact state, contact tile side, pairs listed by the last iteration)"""
        a, b, c = C.c_int(), C.c_int(), C.c_long()
        check(self.lib.paradmm_gpu_contact_info(self.ptr, C.byref(a), C.byref(b), C.byref(c)), "contact_info")
        return bool(a.value), b.value, c.value

    def residuals(self):
        p, d = C.c_double(), C.c_double()
        check(self.lib.paradmm_gpu_residuals(self.ptr, C.byref(p), C.byref(d)), "paradmm_gpu_residuals")
        return p.value, d.value

    def time_iterations(self, iters):
        ms = C.c_double()
        check(self.lib.paradmm_gpu_time_iterations(self.ptr, iters, C.byref(ms)), "paradmm_gpu_time_iterations")
        return ms.value

    def solve(self, max_iters, tol=1e-4, tol_feas=1e-4, check_every=50, policy_every=0, tol_dual=None):
        """Iterates to tolerance on the GPU; returns the paradmm_solve_result as
        a dict.  tol bounds the primal RMS residual, and the dual one unless
        tol_dual is given."""
        o = SolveOpts(max_iters, check_every, tol, tol if tol_dual is None else tol_dual, tol_feas, policy_every)
        r = SolveResult()
        check(self.lib.paradmm_gpu_solve(self.ptr, C.byref(o), C.byref(r)), "paradmm_gpu_solve")
        return r.as_dict()

    def rho_features(self, iter_frac=0.0):
        f = np.zeros((self.nedge, 6), dtype=np.float32)
        check(self.lib.paradmm_gpu_rho_features(self.ptr, iter_frac, _fptr(f)), "rho_features")
        return f

    def set_rho_policy(self, mlp):
        check(self.lib.paradmm_gpu_set_rho_policy(self.ptr, mlp.h if mlp is not None else None), "set_rho_policy")

    def set_rho_policy_mode(self, mode, average=False):
        """0 per edge, 1 per-type mean, 2 one mean over all edges, 3 that
        mean plus the edges the network raises above it (True/False map to
        1/0); average: mean over the applications so far."""
        check(self.lib.paradmm_gpu_set_rho_policy_mode(self.ptr, int(mode), 1 if average else 0),
              "set_rho_policy_mode")

    def apply_rho_policy(self, mlp, iter_frac):
        check(self.lib.paradmm_gpu_apply_rho_policy(self.ptr, mlp.h, iter_frac), "apply_rho_policy")

    def scale_rho(self, factor):
        check(self.lib.paradmm_gpu_scale_rho(self.ptr, factor), "scale_rho")

    def state_save(self):
        check(self.lib.paradmm_gpu_state_save(self.ptr), "state_save")

    def state_restore(self):
        check(self.lib.paradmm_gpu_state_restore(self.ptr), "state_restore")

    def square_quality(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.lib.paradmm_gpu_square_quality(self.ptr, C.byref(a), C.byref(b), C.byref(c)),
              "paradmm_gpu_square_quality")
        return a.value, b.value, c.value

    def profile_kernels(self, iters, max_kernels=16):
        """{kernel name: (mean ms, launches)} over iters eagerly launched iterations."""
        names = C.create_string_buffer(64 * max_kernels)
        ms = (C.c_double * max_kernels)()
        cnt = (C.c_int * max_kernels)()
        nk = C.c_int()
        check(self.lib.paradmm_gpu_profile_kernels(self.ptr, iters, max_kernels, names, ms, cnt, C.byref(nk)),
              "paradmm_gpu_profile_kernels")
        raw = names.raw
        out = {}
        for j in range(nk.value):
            nm = raw[64 * j:64 * (j + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (ms[j], cnt[j])
        return out

    def close(self):
        if self.ptr is not None:
            self.lib.closeG(self.ptr)
            self.ptr = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def packing_square(ncircles, seed, rho=1.0, alpha=1.0, a_exp=1.0):
    """Builds and initialises the unit-square packing problem (dim 3)."""
    g = Graph(3)
    g.packing_square(ncircles, a_exp)
    g.init_rho_alpha(rho, alpha)
    g.init_square(seed)
    return g


class Batch:
    """Many unit-square packings solved in one persistent launch
    (paradmm_batch_* of paradmm_gpu.h).  Keeps its Graph objects alive."""

    def __init__(self, graphs):
        self.lib = load()
        self.graphs = list(graphs)
        arr = (C.POINTER(GraphStruct) * len(self.graphs))(*[g.ptr for g in self.graphs])
        h = C.c_void_p()
        check(self.lib.paradmm_batch_create(C.byref(h), arr, len(self.graphs)), "paradmm_batch_create")
        self.h = h
        _live["batch"].add(self)

    def configure(self, tile, blocks_per_cu=1, barrier=1):
        check(self.lib.paradmm_batch_configure(self.h, tile, blocks_per_cu, barrier), "paradmm_batch_configure")

    def configure_contact(self, contact_tile, blocks_per_cu=1, barrier=1):
        """Contact-state kernel of the plain fixed-rho batch: contact tile 32
        or 64, or 0 to keep the per-edge kernel."""
        check(self.lib.paradmm_batch_configure_contact(self.h, contact_tile, blocks_per_cu, barrier),
              "paradmm_batch_configure_contact")

    def contact_info(self):
        t, b, gr = C.c_int(), C.c_int(), C.c_int()
        nt, lp = C.c_long(), C.c_long()
        ph = (C.c_double * 4)()
        check(self.lib.paradmm_batch_contact_info(self.h, C.byref(t), C.byref(b), C.byref(gr), C.byref(nt),
                                                  C.byref(lp), ph), "paradmm_batch_contact_info")
        return {"contact_tile": t.value, "blocks_per_cu": b.value, "grid": gr.value, "contact_tiles": nt.value,
                "listed_pairs": lp.value,
                "phase_ms": {"after_upload": ph[0], "enter": ph[1], "iterations": ph[2], "leave": ph[3]}}

    def autotune(self, iters=20):
        chosen = BatchConfig()
        tried = (BatchConfig * 32)()
        n = C.c_int()
        check(self.lib.paradmm_batch_autotune(self.h, iters, C.byref(chosen), tried, 32, C.byref(n)),
              "paradmm_batch_autotune")
        return chosen.as_dict(), [tried[i].as_dict() for i in range(n.value)]

    def set_rho_policy(self, mlp):
        check(self.lib.paradmm_batch_set_rho_policy(self.h, mlp.h if mlp is not None else None),
              "paradmm_batch_set_rho_policy")
        self._policy = mlp

    def set_rho_policy_mode(self, mode, average=False):
        """2: one rho per instance; 3: plus the edges the network raises;
        average: running mean of each instance's estimates."""
        check(self.lib.paradmm_batch_set_rho_policy_mode(self.h, int(mode), 1 if average else 0),
              "paradmm_batch_set_rho_policy_mode")

    def solve(self, max_iters, tol=1e-4, tol_feas=1e-4, check_every=50, policy_every=0):
        o = SolveOpts(max_iters, check_every, tol, tol, tol_feas, policy_every)
        res = (SolveResult * len(self.graphs))()
        ms = C.c_double()
        check(self.lib.paradmm_batch_solve(self.h, C.byref(o), res, C.byref(ms)), "paradmm_batch_solve")
        return [r.as_dict() for r in res], ms.value

    def sync_to_host(self):
        check(self.lib.paradmm_batch_sync_to_host(self.h), "paradmm_batch_sync_to_host")

    def info(self):
        t, b, gr, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        sl = C.c_long()
        check(self.lib.paradmm_batch_info(self.h, C.byref(t), C.byref(b), C.byref(gr), C.byref(sl), C.byref(nt)),
              "paradmm_batch_info")
        return {"tile": t.value, "blocks_per_cu": b.value, "grid": gr.value, "slots": sl.value, "tiles": nt.value}

    def close(self):
        if self.h:
            self.lib.paradmm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def packing_triangle(ncircles, seed, lo=-1.0, hi=2.0):
    """The reference demo's packing (packing/host.c:256-343): circles of a
    position and a radius variable (dim 2) in the triangle of walls
    y >= 0, y <= sqrt(3) x and y <= sqrt(3) (1 - x); one radius factor per
    circle (exponent 2), three wall factors per circle, one pair factor per
    pair, created in host.c's order with the library's dim-2 operators;
    X, N, Z, M, U seeded uniform in [lo, hi] (paradmm_init_rand_seeded: the
    reference draws them from rand()), rho = 10 and alpha = 0.1 on every
    edge, rho = 5 on the radius factors' edges."""
    s = np.sin(np.pi / 3)
    walls = np.array([[0., -1., 0., 0.], [-s, 0.5, 0., 0.], [s, 0.5, 1., 0.]])
    g = Graph(2)
    c = np.arange(ncircles, dtype=np.int32)
    g.add_factors(prox_address("paradmm_prox_max_radius_2d"), (2 * c + 1).reshape(-1, 1), np.array([2.0]))
    for k in range(3):
        g.add_factors(prox_address("paradmm_prox_disk_plane_2d"), np.stack([2 * c, 2 * c + 1], 1), walls[k])
    i, j = np.triu_indices(ncircles, 1)
    g.add_factors(prox_address("paradmm_prox_disk_collision_2d"), np.stack([2 * i, 2 * i + 1, 2 * j, 2 * j + 1], 1))
    g.init_rand_seeded(seed, (lo, hi), (lo, hi), (lo, hi), (lo, hi), (lo, hi))
    g.init_rho_alpha(10.0, 0.1)
    g.array("RHOS")[:ncircles] = 5.0
    return g


def batch_sizes(count, seed, lo=16, hi=512):
    """Instance sizes N_i ~ U{lo..hi} drawn from the batch seed (splitmix64)."""
    out = []
    x = seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        out.append(lo + int(z % (hi - lo + 1)))
    return out
