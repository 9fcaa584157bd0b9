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


def _warm_oracle(oracle, f, kind):
    if kind == "square":
        g = oracle.Graph.square(int(f["n"]), 0, float(f["rho"]), 1.0, float(f["a"]))
    else:
        g = oracle.Graph.triangle(int(f["n"]), 0)
    zi = np.ascontiguousarray(f["zinit"])
    ui = np.ascontiguousarray(f["uinit"])
    g.lib.or_warm_start(g.h, oracle.dptr(zi), oracle.dptr(ui), int(f["maxiter"]))
    return g


@pytest.mark.parametrize("fixture,kind", [("warm_square9.npz", "square"), ("warm_triangle7.npz", "triangle")])
def test_warm_start_matches_reference(oracle, fixture, kind):
    """or_warm_start against initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375)
    of the reference compiled from source: every array it writes."""
    f = load(fixture)
    g = _warm_oracle(oracle, f, kind)
    for k in "XMNUZ":
        scale = max(1.0, float(np.max(np.abs(f[k]))))
        assert np.max(np.abs(g.array(k) - f[k])) <= TOL * scale, k
    # the sweeps really moved U away from Uinit and X stayed the broadcast Z
    assert np.mean(f["U"] != f["uinit"]) > 0.5


@pytest.mark.parametrize("nc,rho,maxiter", [(6, 3.0, 1), (11, 50.0, 5)])
def test_warm_start_live_reference(oracle, nc, rho, maxiter):
    ref = oracle.load_ref()
    if ref is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    rng = np.random.default_rng(nc)
    g = oracle.Graph.square(nc, 1, rho, 1.0, 1.0)
    zi = np.ascontiguousarray(np.column_stack([rng.uniform(0, 1, nc), rng.uniform(0, 1, nc),
                                               rng.uniform(0, 0.3, nc)]).ravel())
    ui = rng.uniform(-0.1, 0.1, 3 * g.nedge)
    out = {k: np.zeros(3 * g.nedge) for k in "XMNU"}
    out["Z"] = np.zeros(3 * nc)
    ref.ref_square_warm_start(nc, rho, 1.0, oracle.dptr(zi), oracle.dptr(ui), maxiter,
                              *[oracle.dptr(out[k]) for k in "XMNUZ"])
    g.lib.or_warm_start(g.h, oracle.dptr(zi), oracle.dptr(ui), maxiter)
    for k in "XMNUZ":
        assert np.max(np.abs(g.array(k) - out[k])) <= TOL, k


def test_omp_passes_match_reference_square(oracle):
    """or_iterate_omp (the timed CPU baseline's loop) against the reference's
    five update*OpenMP passes and its one-region updateGOpenMP: the -O2
    checker build to 1e-12 after 20 iterations, the -Ofast baseline build
    (reassociated sums, so its rounding differs and grows with the
    iterations) to 1e-12 after 5."""
    f = load("omp_square40.npz")
    n, seed, rho, alpha = int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"])
    assert np.max(np.abs(f["z_passes"][-1] - f["z_region"])) <= TOL
    g = oracle.Graph.square(n, seed, rho, alpha, 1.0)
    g.iterate_omp(int(f["iters"]), 4)
    assert np.max(np.abs(g.array("Z") - f["z_passes"][-1])) <= TOL
    # the serial oracle gives the same Z: the schedule does not change results
    s = oracle.Graph.square(n, seed, rho, alpha, 1.0)
    h = s.iterate(int(f["iters"]), record=True)
    assert np.max(np.abs(h - f["z_passes"])) <= TOL
    fast = oracle.Graph.square(n, seed, rho, alpha, 1.0, fast=True)
    fast.iterate_omp(5, 4)
    assert np.max(np.abs(fast.array("Z") - f["z_passes"][4])) <= TOL


def test_omp_passes_match_reference_triangle(oracle):
    f = load("omp_triangle12.npz")
    g = oracle.Graph.triangle(int(f["n"]), int(f["seed"]), float(f["lo"]), float(f["hi"]))
    g.iterate_omp(int(f["iters"]), 4)
    scale = max(1.0, float(np.max(np.abs(f["z_region"]))))
    assert np.max(np.abs(g.array("Z") - f["z_passes"][-1])) <= TOL * scale
    assert np.max(np.abs(g.array("Z") - f["z_region"])) <= TOL * scale


@pytest.mark.parametrize("nc,seed,rho,alpha,threads", [(7, 3, 1.0, 1.0, 3), (60, 9, 1.0, 1.0, 8)])
def test_omp_live_reference(oracle, nc, seed, rho, alpha, threads):
    ref = oracle.load_ref()
    if ref is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    iters = 30
    for mode in (1, 2):
        hr = np.zeros((iters, nc * 3))
        ref.ref_square_run_omp(nc, seed, rho, alpha, 1.0, iters, mode, threads, oracle.dptr(hr))
        zr = hr[-1] if mode == 1 else hr[0]
        g = oracle.Graph.square(nc, seed, rho, alpha, 1.0)
        g.iterate_omp(iters, threads)
        assert np.max(np.abs(g.array("Z") - zr)) <= TOL, mode
        fast = oracle.Graph.square(nc, seed, rho, alpha, 1.0, fast=True)
        hr5 = np.zeros((5, nc * 3))
        ref.ref_square_run_omp(nc, seed, rho, alpha, 1.0, 5, 1, threads, oracle.dptr(hr5))
        fast.iterate_omp(5, threads)
        assert np.max(np.abs(fast.array("Z") - hr5[-1])) <= TOL


def test_parallel_iterate_equals_serial(oracle):
    """or_iterate_par (the OpenMP passes, used by the checks at full size)
    gives the serial oracle's bits: Z history, U and Zprev."""
    a = oracle.Graph.square(300, 4, 1.0, 1.0, 1.0)
    b = oracle.Graph.square(300, 4, 1.0, 1.0, 1.0)
    ha = a.iterate(30, record=True)
    hb = b.iterate_par(30, 4, record=True)
    assert np.array_equal(ha, hb)
    for k in ("X", "U", "N", "M", "Zprev"):
        assert np.array_equal(a.array(k), b.array(k)), k
