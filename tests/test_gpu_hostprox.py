"""Factors whose callback has no device kind run on the host, in FP64,
inside the GPU iteration (gather of N, X, rho; callback; scatter of X, rho).
Checked against the FP64 oracle running the very same callbacks."""
import ctypes as C

import numpy as np
import pytest

from engine_amd import paradmm as P
from gpu_helpers import max_rel
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)


def _anchor(x, n, rhos, nvars, ndims, params):
    """Quadratic pull of every edge toward a point a with weight w:
    x = (rho n + w a) / (rho + w); params = double[4] (w, a0, a1, a2)."""
    p = C.cast(params, DP)
    w = p[0]
    for s in range(nvars):
        r = rhos[s]
        for k in range(ndims):
            x[s * ndims + k] = (r * n[s * ndims + k] + w * p[1 + k]) / (r + w)


ANCHOR = P.PROX_FN(_anchor)


def _scale_rho(x, n, rhos, nvars, ndims, params):
    """Identity prox that also doubles its edges' rho (the reference lets a
    callback write rhos, core/paradmm.c:590)."""
    for s in range(nvars):
        for k in range(ndims):
            x[s * ndims + k] = n[s * ndims + k]
        rhos[s] = 2.0 * rhos[s]


SCALE_RHO = P.PROX_FN(_scale_rho)


def _build(nc, seed, rho):
    """Unit-square packing of nc circles plus one anchor factor per circle
    (a host callback), in the library and in the oracle, same factor order."""
    rng = np.random.default_rng(seed)
    anchors = [np.array([0.5, rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), 0.05]) for _ in range(nc)]
    g = P.Graph(3)
    g.packing_square(nc, 1.0)
    for i in range(nc):
        g._keep.append(anchors[i])
        g.lib.addNode(g.ptr, C.cast(ANCHOR, C.c_void_p), anchors[i].ctypes.data_as(C.c_void_p), 1,
                      (C.c_int * 1)(i))
    g.init_rho_alpha(rho, 1.0)
    g.init_square(seed)
    lib = po.load()
    h = lib.or_new(3)
    a_exp = np.array([1.0])
    keep = [a_exp]
    for i in range(nc):
        lib.or_add_factor(h, C.cast(lib.or_prox_radius3, C.c_void_p), a_exp.ctypes.data_as(C.c_void_p), 1,
                          (C.c_int * 1)(i))
    for i in range(nc):
        lib.or_add_factor(h, C.cast(lib.or_prox_box3, C.c_void_p), None, 1, (C.c_int * 1)(i))
    for i in range(nc - 1):
        for j in range(i + 1, nc):
            lib.or_add_factor(h, C.cast(lib.or_prox_pair3, C.c_void_p), None, 2, (C.c_int * 2)(i, j))
    for i in range(nc):
        lib.or_add_factor(h, C.cast(ANCHOR, C.c_void_p), anchors[i].ctypes.data_as(C.c_void_p), 1,
                          (C.c_int * 1)(i))
    lib.or_alloc_state(h)
    lib.or_set_rho_alpha(h, rho, 1.0)
    o = po.Graph(h, lib)
    o._keep = keep + anchors
    for name in ("X", "M", "N", "U", "Z"):
        o.array(name)[:] = g.array(name)
    o.array("Zprev")[:] = g.array("Z")
    return g, o


@pytest.mark.parametrize("nc,rho", [(9, 5.0), (30, 20.0)])
def test_host_callback_factors_match_oracle(nc, rho):
    """Untethered for 5 iterations (fp32 state and FP64 trajectories part
    later), then two-iteration steps from 15 oracle states."""
    g, o = _build(nc, 3, rho)
    g.prepare()
    assert g.layout() == 1  # a host callback rules out the dense pair tiles
    g.iterate(1)
    o.iterate(1)
    g.sync_to_host()
    assert max_rel(g.array("Z"), o.array("Z")) < 1e-5
    g.iterate(4)
    o.iterate(4)
    g.sync_to_host()
    assert max_rel(g.array("Z"), o.array("Z")) < 1e-5
    errs = []
    names = ("X", "M", "N", "U", "Z")
    for step in range(15):
        for n in names:
            g.array(n)[:] = o.array(n)
        g.sync_from_host()
        g.iterate(2)
        g.sync_to_host()
        snap = {n: o.array(n).copy() for n in names}
        o.iterate(2)
        errs.append(max_rel(g.array("Z"), o.array("Z")))
        for n in names:
            o.array(n)[:] = snap[n]
        o.iterate(3)
    errs = np.array(errs)
    print(f"N={nc} rho={rho}: two-step Z error median {np.median(errs):.2e} max {errs.max():.2e}")
    assert np.median(errs) < 2e-6 and np.mean(errs < 1e-5) >= 0.9
    # legacy per-pass calls run the callbacks too
    for n in names:
        g.array(n)[:] = o.array(n)
    g.legacy("updateXM")
    g.legacy("updateZ")
    g.legacy("updateUN")
    o.lib.or_update_XM(o.h)
    o.lib.or_update_Z(o.h)
    o.lib.or_update_UN(o.h)
    assert max_rel(g.array("Z"), o.array("Z")) < 1e-5


def test_host_callback_may_write_rho():
    g = P.Graph(3)
    g.packing_square(6, 1.0)
    for i in range(6):
        g.lib.addNode(g.ptr, C.cast(SCALE_RHO, C.c_void_p), None, 1, (C.c_int * 1)(i))
    g.init_rho_alpha(3.0, 1.0)
    g.init_square(1)
    g.iterate(2)
    g.sync_to_host()
    rho = g.array("RHOS")
    E = g.nedge
    assert np.all(rho[:E - 6] == 3.0)
    assert np.allclose(rho[E - 6:], 12.0)  # doubled in each of the two X passes
