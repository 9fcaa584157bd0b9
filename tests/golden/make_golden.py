"""Writes the golden fixtures of tests/golden/ from the reference compiled from
its own sources (oracle/_ref, built by `make -C oracle` where the reference
is present).  Run once in a container that has the reference; the fixtures are
committed and travel, the reference does not.

Fixtures (numpy .npz, float64):
  square20_seed0.npz   unit-square packing N=20 (dim 3), rho=1, alpha=1, a=1:
                       reference engine (core/paradmm.c updateXM/updateZ/updateUN)
                       with the oracle's dim-3 operators; Z after iterations 1..200
                       and after 5000.
  square20_seed0_rho100.npz   same at rho=100 (the converging setting).
  square20_seed0_rho100_converged.npz   Z of the rho=100 run after 12000
                       iterations (converged: tolerance 1e-4 is met at 8450).
  triangle6_seed1.npz  the reference demo's triangle packing (packing/host.c
                       factor order and operators) N=6, seeded init in [-1,2];
                       Z after iterations 1..200.
  prox2_kat.npz        the reference's own dim-2 operators (packing/host.c) on
                       seeded random inputs: inputs and outputs.
  warm_square9.npz     initialize_X_N_Z_M_U_optimal (core/paradmm.c:266-375) of
  warm_triangle7.npz   the reference on the square (oracle dim-3 operators) and
                       triangle (its own operators) graphs: random Zinit, Uinit,
                       maxiter 3; X, M, N, U, Z out.
  omp_square40.npz     the reference's OpenMP passes (the five update*OpenMP
  omp_triangle12.npz   calls, packing/host.c:237-244, on 4 threads) and its
                       one-region updateGOpenMP (core/paradmm.c:949-1092): Z
                       after 20 iterations of each.

`python make_golden.py warm_omp` writes only the last four.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from oracle import pyoracle as po  # noqa: E402


def square(ref, nc, seed, rho, alpha, a, iters):
    nv = nc * 3
    hist = np.zeros((iters, nv))
    ref.ref_square_run(nc, seed, rho, alpha, a, iters, po.dptr(hist))
    return hist


def main():
    ref = po.load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (needs the reference sources): make -C oracle")
    h = square(ref, 20, 0, 1.0, 1.0, 1.0, 5000)
    np.savez_compressed(os.path.join(HERE, "square20_seed0.npz"), z=h[:200], z_final=h[-1],
                        n=20, seed=0, rho=1.0, alpha=1.0, a=1.0, iters_final=5000)
    h = square(ref, 20, 0, 100.0, 1.0, 1.0, 5000)
    np.savez_compressed(os.path.join(HERE, "square20_seed0_rho100.npz"), z=h[:200], z_final=h[-1],
                        n=20, seed=0, rho=100.0, alpha=1.0, a=1.0, iters_final=5000)

    # converged state of the rho = 100 run (tolerance 1e-4 is met at 8450)
    h = square(ref, 20, 0, 100.0, 1.0, 1.0, 12000)
    np.savez_compressed(os.path.join(HERE, "square20_seed0_rho100_converged.npz"), z_final=h[-1],
                        n=20, seed=0, rho=100.0, alpha=1.0, a=1.0, iters_final=12000)

    nc, iters = 6, 200
    hist = np.zeros((iters, 2 * nc * 2))
    ref.ref_triangle_run(nc, 1, -1.0, 2.0, iters, po.dptr(hist))
    np.savez_compressed(os.path.join(HERE, "triangle6_seed1.npz"), z=hist, n=nc, seed=1, lo=-1.0, hi=2.0)

    # known-answer vectors of the reference's dim-2 operators
    rng = np.random.default_rng(12345)
    m = 400
    pair_n = rng.uniform(-1, 2, size=(m, 8))
    pair_n[:, 2:4] = pair_n[:, 0:2] + rng.uniform(-0.5, 0.5, size=(m, 2))  # many overlaps
    pair_rho = rng.uniform(0.5, 20, size=(m, 4))
    pair_x = np.zeros((m, 8))
    for i in range(m):
        x = np.zeros(8)
        po.call_prox(ref.ref_prox_pair2, x, pair_n[i].copy(), pair_rho[i].copy(), 4, 2)
        pair_x[i] = x
    s = np.sin(np.pi / 3)
    walls = np.array([[0., -1., 0., 0.], [-s, 0.5, 0., 0.], [s, 0.5, 1., 0.]])
    wall_n = rng.uniform(-1, 2, size=(m, 4))
    wall_rho = rng.uniform(0.5, 20, size=(m, 2))
    wall_k = rng.integers(0, 3, size=m)
    wall_x = np.zeros((m, 4))
    for i in range(m):
        x = np.zeros(4)
        po.call_prox(ref.ref_prox_wall2, x, wall_n[i].copy(), wall_rho[i].copy(), 2, 2, walls[wall_k[i]].copy())
        wall_x[i] = x
    rad_n = rng.uniform(-1, 2, size=(m, 2))
    rad_rho = rng.uniform(2.5, 20, size=(m, 1))
    rad_a = np.where(rng.uniform(size=m) < 0.5, 1.0, 2.0)
    rad_x = np.zeros((m, 2))
    for i in range(m):
        x = np.zeros(2)
        po.call_prox(ref.ref_prox_radius2, x, rad_n[i].copy(), rad_rho[i].copy(), 1, 2, np.array([rad_a[i]]))
        rad_x[i] = x
    np.savez_compressed(os.path.join(HERE, "prox2_kat.npz"), pair_n=pair_n, pair_rho=pair_rho, pair_x=pair_x,
                        walls=walls, wall_n=wall_n, wall_rho=wall_rho, wall_k=wall_k, wall_x=wall_x,
                        rad_n=rad_n, rad_rho=rad_rho, rad_a=rad_a, rad_x=rad_x)
    write_warm_omp(ref)
    print("golden fixtures written to", HERE)


def write_warm_omp(ref):
    rng = np.random.default_rng(2024)
    nc, rho, maxiter = 9, 3.0, 3
    ne = nc + nc + nc * (nc - 1) // 2 * 2
    zi = np.ascontiguousarray(np.column_stack([rng.uniform(0, 1, nc), rng.uniform(0, 1, nc),
                                               rng.uniform(0, 0.3, nc)]).ravel())
    ui = rng.uniform(-0.1, 0.1, 3 * ne)
    out = {k: np.zeros(3 * ne) for k in "XMNU"}
    out["Z"] = np.zeros(3 * nc)
    ref.ref_square_warm_start(nc, rho, 1.0, po.dptr(zi), po.dptr(ui), maxiter, *[po.dptr(out[k]) for k in "XMNUZ"])
    np.savez_compressed(os.path.join(HERE, "warm_square9.npz"), n=nc, rho=rho, a=1.0, maxiter=maxiter,
                        zinit=zi, uinit=ui, **out)

    nc = 7
    nv, ne = 2 * nc, nc + 3 * nc * 2 + nc * (nc - 1) // 2 * 4
    zi = rng.uniform(-1, 2, 2 * nv)
    ui = rng.uniform(-1, 1, 2 * ne)
    out = {k: np.zeros(2 * ne) for k in "XMNU"}
    out["Z"] = np.zeros(2 * nv)
    ref.ref_triangle_warm_start(nc, po.dptr(zi), po.dptr(ui), maxiter, *[po.dptr(out[k]) for k in "XMNUZ"])
    np.savez_compressed(os.path.join(HERE, "warm_triangle7.npz"), n=nc, maxiter=maxiter, zinit=zi, uinit=ui,
                        **out)

    nc, seed, rho, alpha, iters = 40, 5, 10.0, 0.5, 20
    z5 = np.zeros((iters, 3 * nc))
    ref.ref_square_run_omp(nc, seed, rho, alpha, 1.0, iters, 1, 4, po.dptr(z5))
    zg = np.zeros((1, 3 * nc))
    ref.ref_square_run_omp(nc, seed, rho, alpha, 1.0, iters, 2, 4, po.dptr(zg))
    np.savez_compressed(os.path.join(HERE, "omp_square40.npz"), n=nc, seed=seed, rho=rho, alpha=alpha, a=1.0,
                        iters=iters, z_passes=z5, z_region=zg[0])

    nc, seed = 12, 7
    z5 = np.zeros((iters, 4 * nc))
    ref.ref_triangle_run_omp(nc, seed, -1.0, 2.0, iters, 1, 4, po.dptr(z5))
    zg = np.zeros((1, 4 * nc))
    ref.ref_triangle_run_omp(nc, seed, -1.0, 2.0, iters, 2, 4, po.dptr(zg))
    np.savez_compressed(os.path.join(HERE, "omp_triangle12.npz"), n=nc, seed=seed, lo=-1.0, hi=2.0, iters=iters,
                        z_passes=z5, z_region=zg[0])


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "warm_omp":
        r = po.load_ref()
        if r is None:
            raise SystemExit("oracle/_ref is not built (needs the reference sources): make -C oracle")
        write_warm_omp(r)
    else:
        main()
