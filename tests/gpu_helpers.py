"""Shared builders for the GPU parity tests: the same problem in libparadmm
(through its C ABI) and in the FP64 oracle."""
import numpy as np

from engine_amd import paradmm as P
from oracle import pyoracle as po

STATE = ("X", "M", "N", "U", "Z")
OR_NAME = {"X": "X", "M": "M", "N": "N", "U": "U", "Z": "Z", "RHOS": "RHO", "ALPHAS": "ALPHA"}


def triangle_lib(nc, seed, lo=-1.0, hi=2.0):
    """The reference demo's triangle packing (packing/host.c:256-343) built
    with the library's dim-2 operators, init copied from the oracle build."""
    s = np.sin(np.pi / 3)
    walls = [[0., -1., 0., 0.], [-s, 0.5, 0., 0.], [s, 0.5, 1., 0.]]
    g = P.Graph(2)
    rad = P.prox_address("paradmm_prox_max_radius_2d")
    wall = P.prox_address("paradmm_prox_disk_plane_2d")
    pair = P.prox_address("paradmm_prox_disk_collision_2d")
    for i in range(nc):
        g.add_factor(rad, [2 * i + 1], [2.0])
    for k in range(3):
        for j in range(nc):
            g.add_factor(wall, [2 * j, 2 * j + 1], walls[k])
    for i in range(nc - 1):
        for j in range(i + 1, nc):
            g.add_factor(pair, [2 * i, 2 * i + 1, 2 * j, 2 * j + 1])
    g.init_zeros()
    g.init_rho_alpha(10.0, 0.1)
    o = po.Graph.triangle(nc, seed, lo, hi)
    copy_state(o, g)
    return g, o


def square_pair(nc, seed, rho=1.0, alpha=1.0, a=1.0):
    g = P.packing_square(nc, seed, rho, alpha, a)
    o = po.Graph.square(nc, seed, rho, alpha, a)
    return g, o


def copy_state(o, g, names=("X", "M", "N", "U", "Z", "RHOS", "ALPHAS")):
    """oracle -> library host arrays"""
    for n in names:
        g.array(n)[:] = o.array(OR_NAME[n])
    g.sync_from_host() if g.layout() else None


def randomize(o, g, seed, lo=-1.0, hi=2.0, rho_range=(0.5, 20.0)):
    rng = np.random.default_rng(seed)
    for n in ("X", "M", "N", "U", "Z"):
        a = o.array(n)
        a[:] = rng.uniform(lo, hi, size=a.shape)
    o.array("RHO")[:] = rng.uniform(*rho_range, size=o.nedge)
    o.array("ALPHA")[:] = rng.uniform(0.05, 1.0, size=o.nedge)
    for n in ("X", "M", "N", "U", "Z", "RHOS", "ALPHAS"):
        g.array(n)[:] = o.array(OR_NAME[n])


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def trajectory_error(hist_gpu, hist_ref):
    """per-iteration norm-wise relative error max_v |dZ| / max_v |Z_ref|"""
    num = np.max(np.abs(hist_gpu - hist_ref), axis=1)
    den = np.maximum(np.max(np.abs(hist_ref), axis=1), 1e-300)
    return num / den
