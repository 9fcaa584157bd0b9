"""Pins the FP64 oracle against the reference.

  * golden fixtures written by tests/golden/make_golden.py from the reference
    compiled from its own sources (core/paradmm.c, packing/host.c);
  * where oracle/_ref is built (this container only), the reference engine live.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

TOL = 1e-12


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("fixture", ["square20_seed0.npz", "square20_seed0_rho100.npz"])
def test_square_trajectory_matches_reference(oracle, fixture):
    """Oracle engine (paradmm.c:569-741 restated) vs reference engine, same dim-3 ops."""
    f = load(fixture)
    g = oracle.Graph.square(int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    h = g.iterate(200, record=True)
    assert np.max(np.abs(h - f["z"])) <= TOL
    g.iterate(int(f["iters_final"]) - 200)
    assert np.max(np.abs(g.array("Z") - f["z_final"])) <= TOL


def test_square_converged_state_matches_reference(oracle):
    """12000 iterations at rho = 100 (converged): oracle and reference agree."""
    f = load("square20_seed0_rho100_converged.npz")
    g = oracle.Graph.square(int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    g.iterate(int(f["iters_final"]))
    assert np.max(np.abs(g.array("Z") - f["z_final"])) <= 1e-10


def test_triangle_trajectory_matches_reference(oracle):
    """The reference demo problem (packing/host.c factor order, its own dim-2 ops)."""
    f = load("triangle6_seed1.npz")
    g = oracle.Graph.triangle(int(f["n"]), int(f["seed"]), float(f["lo"]), float(f["hi"]))
    h = g.iterate(f["z"].shape[0], record=True)
    assert np.max(np.abs(h - f["z"])) <= 1e-12 * max(1.0, np.max(np.abs(f["z"])))


def test_dim2_operators_match_reference_kat(oracle):
    """packing/host.c operators: disk_collision, disk_plane, max radius (a = 1, 2)."""
    f = load("prox2_kat.npz")
    lib = oracle.load()
    for i in range(f["pair_n"].shape[0]):
        x = np.zeros(8)
        oracle.call_prox(lib.or_prox_pair2, x, f["pair_n"][i].copy(), f["pair_rho"][i].copy(), 4, 2)
        assert np.allclose(x, f["pair_x"][i], rtol=1e-13, atol=1e-13)
    for i in range(f["wall_n"].shape[0]):
        x = np.zeros(4)
        oracle.call_prox(lib.or_prox_wall2, x, f["wall_n"][i].copy(), f["wall_rho"][i].copy(), 2, 2,
                         f["walls"][f["wall_k"][i]].copy())
        assert np.allclose(x, f["wall_x"][i], rtol=1e-13, atol=1e-13)
    for i in range(f["rad_n"].shape[0]):
        x = np.zeros(2)
        oracle.call_prox(lib.or_prox_radius2, x, f["rad_n"][i].copy(), f["rad_rho"][i].copy(), 1, 2,
                         np.array([f["rad_a"][i]]))
        assert np.allclose(x, f["rad_x"][i], rtol=1e-13, atol=1e-13, equal_nan=True)
    # the KAT really exercises the active branches
    assert np.mean(np.any(f["pair_x"] != f["pair_n"], axis=1)) > 0.2
    assert np.mean(np.any(f["wall_x"] != f["wall_n"], axis=1)) > 0.2


@pytest.mark.parametrize("nc,seed,rho,alpha", [(7, 3, 1.0, 1.0), (11, 5, 10.0, 0.1), (16, 2, 100.0, 1.0)])
def test_square_live_reference(oracle, nc, seed, rho, alpha):
    ref = oracle.load_ref()
    if ref is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    iters = 300
    hr = np.zeros((iters, nc * 3))
    ref.ref_square_run(nc, seed, rho, alpha, 1.0, iters, oracle.dptr(hr))
    g = oracle.Graph.square(nc, seed, rho, alpha, 1.0)
    h = g.iterate(iters, record=True)
    assert np.max(np.abs(h - hr)) <= TOL


@pytest.mark.parametrize("nc,seed", [(3, 4), (9, 8)])
def test_triangle_live_reference(oracle, nc, seed):
    ref = oracle.load_ref()
    if ref is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    iters = 150
    hr = np.zeros((iters, nc * 4))
    ref.ref_triangle_run(nc, seed, -1.0, 2.0, iters, oracle.dptr(hr))
    g = oracle.Graph.triangle(nc, seed)
    h = g.iterate(iters, record=True)
    assert np.max(np.abs(h - hr)) <= 1e-12 * max(1.0, np.max(np.abs(hr)))
