#!/usr/bin/env python3
"""Plain-ADMM throughput of the parADMM engine on one MI355X.

A step is one full ADMM iteration (X, M, Z, U, N over every edge) of the
unit-square circle packing, N = 5000 circles by default: 12,507,500 pair
factors, 5000 box and 5000 radius factors, E = 25,005,000 edges of dim 3,
fp32 messages, fixed rho = 1, alpha = 1, seeded synthetic init.  Inputs are
resident in HBM before the timed region; the timed region is K iterations
launched through the engine's captured hipGraph and bracketed by HIP events
on the engine's stream (the library synchronises before and after).  With
one rho and alpha = 1 the engine runs the dense layout's contact state (only
pairs whose prox is active keep per-edge values; every edge is still updated
every iteration), from N = 3200 in its spatial form (circles in Morton
order, contact tiles only where two blocks of circles can touch).

Prints one JSON line: value = edge-updates/s, plus the roofline of the
dominant kernel (its algorithmic bytes per launch / mean launch time from HIP
events recorded by the dispatch itself, the HBM traffic of the committed
rocprofv3 counter profile, and for the contact state its VALU issue rate).
Further legs: the contact state's id-order form and the per-edge path on
the same workload (PARADMM_CONTACT=0, HBM bound), N = 1000 circles, both
contact forms at N = 10000, the 256-instance batch, learned vs fixed rho to
tolerance, and the CPU baseline: the FP64 restatement of the reference's
OpenMP path compiled on this host and timed on all its usable cores and on
one, with the GPU's first iterations checked against the FP64 oracle.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes of one edge-update (one edge through a full X/M/Z/U/N
# iteration, fp32, d = 3): the X pass reads U (12) and rho (4) and writes X
# (12); the Z pass reads X (12), U (12), rho (4), alpha (4) and writes U (12):
# 72 B.  N and M are transients and Z (N x 12 B) stays in cache.
EDGE_UPDATE_BYTES = 72
# What each kernel itself must move per edge it processes (index streams and
# per-factor parameters excluded):
#   generic X pass (per kind): read U, rho; write X                    = 28
#   generic Z pass (fused M, U update): read X, U, rho, alpha; write U = 44
#   dense fused pair-tile pass (X and Z work of a whole iteration in one
#   pass): read X, U, rho, alpha; write X, U                           = 56
#   the same pass inside a captured run with alpha = 1, the X array
#   holding M = X + U (U not stored): read M, rho; write M             = 28
# and with one rho and one alpha on every edge (the variant without those
# streams) 48 and 24.  The contact-state tile kernel (one rho, alpha = 1)
# keeps per-edge values only for the pairs whose prox is active: its bytes
# are not a per-edge constant (the measured traffic is reported instead).
BYTES_PER_EDGE = {
    "disk_collision_3d": 28, "box_3d": 28, "max_radius_3d": 28,
    "zpass_fused": 44,
    "pairtile_fused": 56, "pairtile_m": 28, "unary_fused": 56, "zreduce": 0,
}
UNIFORM_BYTES_PER_EDGE = {"pairtile_fused": 48, "pairtile_m": 24}
# kernels that carry a whole iteration's edge work (their launches count
# E edge-updates each)
FULL_ITERATION_KERNELS = ("pairtile_fused", "pairtile_m", "contact_tile", "scs_rows")
TRAFFIC_PROFILES = {k: os.path.join(ROOT, "profiles", f"{k}_n5000_hbm_traffic.md") for k in FULL_ITERATION_KERNELS}
VALU_PROFILES = {k: os.path.join(ROOT, "profiles", f"{k}_n5000_valu.md") for k in ("contact_tile", "scs_rows")}
# MI355X vector issue: 256 CUs x 4 SIMDs, one wave64 VALU instruction per
# SIMD every 4 cycles at 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def edges_per_launch(name, n):
    pairs = n * (n - 1) // 2
    return {
        "disk_collision_3d": 2 * pairs, "box_3d": n, "max_radius_3d": n,
        "zpass_fused": n * (n + 1), "pairtile_fused": 2 * pairs, "pairtile_m": 2 * pairs, "unary_fused": 2 * n,
    }.get(name, 0)


def _profile_value(path, key, n, seed):
    if path is None or n != 5000 or seed != 0 or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    return None


def measured_traffic(kernel, n, seed):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (copy-kernel calibrated), when it covers this kernel and workload."""
    return _profile_value(TRAFFIC_PROFILES.get(kernel), "traffic_bytes_per_launch", n, seed)


def measured_valu(kernel, n, seed):
    """VALU wave-instructions per launch from the committed rocprofv3
    SQ_INSTS_VALU summary, when it covers this kernel and workload."""
    return _profile_value(VALU_PROFILES.get(kernel), "valu_insts_per_launch", n, seed)


def contact_alg_bytes(n, cs, listed):
    """Bytes one contact-state iteration must move at the least: the contact
    lists (both edges' M, 32 B per listed pair, read, and as many written:
    the list is about the same length from one iteration to the next), the
    list bit masks (cs*cs bits per contact tile, read and written), per
    circle Z and the previous launch's three sums read (36 B), its Z stored
    (12 B), eight sum words cleared (64 B) and this launch's three sums
    added (24 B), and the box and radius edges' X and U read and written
    (2 x 48 B per circle)."""
    nbs = (n + cs - 1) // cs
    nts = nbs * (nbs + 1) // 2
    return 64 * listed + nts * cs * cs // 4 + n * (36 + 12 + 64 + 24) + 96 * n


def roofline_contact(dom_ms, alg_bytes, traffic, valu, copy_gbs, n, edges, listed):
    """Roofline of the contact tile kernel: its own algorithmic bytes per
    launch (contact_alg_bytes) against HBM peak, the measured traffic of the
    committed counter profile, and its VALU issue rate."""
    t = dom_ms * 1e-3
    achieved = alg_bytes / t / 1e9
    r = {
        "bound": "hbm", "kernel": "contact_tile", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "algorithmic_bytes_per_launch": alg_bytes, "listed_pairs": listed,
        "traffic_gbs": (traffic / t / 1e9) if traffic else None,
        "traffic_frac": (traffic / t / 1e9 / HBM_PEAK_GBS) if traffic else None,
        "traffic_frac_of_copy": (traffic / t / 1e9 / copy_gbs) if traffic else None,
        "mean_launch_ms": dom_ms, "copy_kernel_gbs": copy_gbs,
        "pair_tests_per_s": (n * (n - 1) // 2) / t,
        "edge_update_unit": {"bytes_per_edge_update": EDGE_UPDATE_BYTES,
                             "bytes_per_launch": EDGE_UPDATE_BYTES * edges,
                             "equivalent_gbs": EDGE_UPDATE_BYTES * edges / t / 1e9,
                             "note": "the per-edge state an X pass and a Z pass would move for the same "
                                     "edge-updates; the contact state stores per-edge values only for listed "
                                     "pairs, so this is a throughput equivalent, not a bandwidth"},
        "limiter": "latency: all tile wavefronts are resident at once and wait on their first loads "
                   "(Z of the tile's circles, list entries); HBM and VALU are both well below peak",
    }
    if valu:
        rate = valu / t
        r["valu"] = {"insts_per_launch": valu, "achieved": rate, "peak": VALU_PEAK_WAVE_INSTS,
                     "unit": "wave64 VALU instructions/s", "frac": rate / VALU_PEAK_WAVE_INSTS}
    return r


def spatial_alg_bytes(n, listed):
    """Bytes the spatial form's row kernel (k_scs_rows) must move at the
    least: the contact lists (both edges' M, 32 B per listed pair, read and
    as many written), per circle its (Z_k, c) read once (32 B) and its three
    sums added (24 B), and the box and radius edges' X and U read and
    written (96 B).  The per-circle kernel before it (k_scs_circles: Z and
    the pending sums read, Z and (Z_k, c) stored, the sum buffer cleared)
    is timed and reported on its own."""
    return 64 * listed + n * (32 + 24 + 96)


def roofline_spatial(dom_ms, alg_bytes, traffic, valu, copy_gbs, edges, listed, circles_ms):
    """Roofline of the spatial form's row kernel: its own algorithmic bytes
    per launch against HBM peak, the measured traffic of the committed
    counter profile, and its VALU issue rate."""
    t = dom_ms * 1e-3
    achieved = alg_bytes / t / 1e9
    r = {
        "bound": "hbm", "kernel": "scs_rows", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "algorithmic_bytes_per_launch": alg_bytes, "listed_pairs": listed,
        "traffic_gbs": (traffic / t / 1e9) if traffic else None,
        "traffic_frac": (traffic / t / 1e9 / HBM_PEAK_GBS) if traffic else None,
        "traffic_frac_of_copy": (traffic / t / 1e9 / copy_gbs) if traffic else None,
        "mean_launch_ms": dom_ms, "copy_kernel_gbs": copy_gbs,
        "circles_kernel_ms": circles_ms,
        "edge_update_unit": {"bytes_per_edge_update": EDGE_UPDATE_BYTES,
                             "bytes_per_launch": EDGE_UPDATE_BYTES * edges,
                             "equivalent_gbs": EDGE_UPDATE_BYTES * edges / t / 1e9,
                             "note": "the per-edge state an X pass and a Z pass would move for the same "
                                     "edge-updates; the contact state stores per-edge values only for listed "
                                     "pairs, so this is a throughput equivalent, not a bandwidth"},
        "limiter": "latency: each row workgroup lists its running tiles, then its wavefronts take them in "
                   "turn; a tile's loads (its circles' Z, list entries) and its prox rounds are one dependent "
                   "chain; HBM and VALU are both well below peak",
    }
    if valu:
        rate = valu / t
        r["valu"] = {"insts_per_launch": valu, "achieved": rate, "peak": VALU_PEAK_WAVE_INSTS,
                     "unit": "wave64 VALU instructions/s", "frac": rate / VALU_PEAK_WAVE_INSTS}
    return r


def roofline_per_edge(dom, dom_ms, kernel_bytes, traffic, copy_gbs, edges):
    """Roofline of the per-edge tile kernel: the bytes it must move per edge
    (24 B inside captured runs: M read and written; 48 B for the plain fused
    pass) against HBM peak."""
    t = dom_ms * 1e-3
    achieved = kernel_bytes / t / 1e9
    return {
        "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "algorithmic_bytes_per_launch": kernel_bytes,
        "traffic_gbs": (traffic / t / 1e9) if traffic else None,
        "traffic_frac_of_copy": (traffic / t / 1e9 / copy_gbs) if traffic else None,
        "mean_launch_ms": dom_ms, "copy_kernel_gbs": copy_gbs,
        "edge_update_unit_equivalent_gbs": EDGE_UPDATE_BYTES * edges / t / 1e9,
    }


def _physical_cores():
    seen = set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("physical id"):
                    phys = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif not line.strip():
                    if core is not None:
                        seen.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return len(seen) or (os.cpu_count() or 1)


def cpu_baseline(n, gpu_zhist, threads=0):
    """The reference's OpenMP path (the five update*OpenMP passes per
    iteration, packing/host.c:237-244; core/paradmm.c:812-945) as the
    oracle's FP64 restatement, pinned to the reference compiled from source,
    compiled here with gcc -Ofast -march=native -fopenmp (the reference's
    build flags plus this host's ISA) and timed with gettimeofday around the
    iteration loop: N = 5000 (20 iterations) and N = 1000 (1000) on every
    thread this process may use, and a shorter sample of each on 1 thread.
    Also the checker's job of this leg: the GPU's Z over its first
    iterations from the seeded start against the FP64 oracle (-O2 build)."""
    import tempfile

    from oracle import pyoracle as po
    tmp = tempfile.mkdtemp(prefix="pdm_oracle_")
    lib = po.build_native(tmp)
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env_t = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    phys = _physical_cores()
    nt = threads or min(x for x in (avail, phys, env_t or avail) if x > 0)
    runs = []
    for size, iters, th in ((5000, 20, nt), (1000, 1000, nt), (5000, 3, 1), (1000, 100, 1)):
        g = po.Graph(lib.or_build_square(size, 0, 1.0, 1.0, 1.0), lib)
        dt = g.iterate_omp(iters, th)
        runs.append({"n": size, "iterations": iters, "threads": th, "seconds": dt,
                     "edge_updates_per_s": g.nedge * iters / dt})
        print(f"# cpu baseline N={size} {iters} it on {th} threads: {dt:.2f} s, "
              f"{runs[-1]['edge_updates_per_s']:.3e} edge-updates/s", file=sys.stderr, flush=True)
        g.close()
    main = runs[0]
    out = {"value": main["edge_updates_per_s"], "unit": "edge-updates/s", "cores": main["threads"], "kind": "port",
           "sample": f"N=5000 unit-square packing (E=25,005,000), 20 OpenMP iterations of the FP64 restatement "
                     f"(-Ofast -march=native -fopenmp, compiled on this host), {main['seconds']:.2f} s",
           "host_physical_cores": phys, "threads_available": avail, "runs": runs}
    # the GPU run's error against the FP64 oracle over its first iterations
    o = po.Graph.square(n, 0, 1.0, 1.0, 1.0)
    hist = o.iterate_par(gpu_zhist.shape[0], nt, record=True)
    o.close()
    import numpy as np
    num = np.max(np.abs(gpu_zhist - hist), axis=1)
    den = np.maximum(np.max(np.abs(hist), axis=1), 1e-300)
    err = (num / den).tolist()
    oracle_error = {"n": n, "iterations": len(err), "max_rel_Z_error_per_iteration": err, "max": max(err),
                    "note": "norm-wise max |Z_gpu - Z_oracle| / max |Z_oracle| after each iteration from the "
                            "seeded start, the GPU in its default (contact) path, the oracle in FP64"}
    print(f"# FP64 oracle error of the GPU's first {len(err)} iterations at N={n}: max {max(err):.2e}",
          file=sys.stderr, flush=True)
    return out, oracle_error


def batch_rho(n):
    """Fixed rho of one batch instance: the fastest-converging fixed rho grows
    with N (about 35 at N = 40..64, 70 at 100..150, 150 at 300, 150..300 at
    500 on the plain solver), so the batch uses clip(N/2, 35, 300)."""
    return float(min(300.0, max(35.0, 0.5 * n)))


def batch_alg_bytes(sizes, cs, listed):
    """contact_alg_bytes summed over the instances of a batch: the list term
    over the batch's listed pairs, the tile-mask and per-circle terms per
    instance."""
    tot = 64 * listed
    for n in sizes:
        nbs = (n + cs - 1) // cs
        tot += nbs * (nbs + 1) // 2 * cs * cs // 4 + n * (36 + 12 + 64 + 24) + 96 * n
    return tot


def batch_bench(P, count, seed, iters, cap, per_edge=False):
    """The batched workload: `count` unit-square packings with N ~ U{16..512}
    drawn from the batch seed.  The autotuner picks the form (the contact
    state with one launch per iteration, or one cooperative launch) and the
    contact tile side.  Returns the throughput of `iters` iterations without
    retirement (whole solve, and the iterations alone with the contact
    roofline on the batch's own algorithmic bytes), the time-to-tolerance
    solve with per-instance retirement, and the learned policy on the same
    batch against fixed rho = 1.  per_edge: the per-edge kernel (one
    cooperative launch, no contact state) throughout."""
    t0 = time.time()
    sizes = P.batch_sizes(count, seed)
    gs = [P.packing_square(n, 1000 * seed + k, rho=batch_rho(n), alpha=1.0, a_exp=1.0) for k, n in enumerate(sizes)]
    E = sum(g.nedge for g in gs)
    b = P.Batch(gs)
    if per_edge:
        b.configure_contact(0, 0, 0)
    chosen, tried = b.autotune(20)
    t_setup = time.time() - t0
    print(f"# batch: {count} instances, E={E}, setup {t_setup:.1f} s, autotune {chosen} of {tried}",
          file=sys.stderr, flush=True)
    _, ms = b.solve(iters, tol=-1.0, tol_feas=-1.0, check_every=iters + 1)
    thr = E * iters / (ms * 1e-3)
    ci = b.contact_info()
    it_ms = ci["phase_ms"]["iterations"]
    per_iter = it_ms / (iters - 1) if chosen.get("contact_tile") else ms / iters
    thr_it = E / (per_iter * 1e-3)
    print(f"# batch throughput: whole solve {ms / iters:.4f} ms/iter, {thr:.4e} edge-updates/s; iterations alone "
          f"{per_iter:.4f} ms/iter, {thr_it:.4e} edge-updates/s; {ci}", file=sys.stderr, flush=True)
    roof = None
    if chosen.get("contact_tile"):
        alg = batch_alg_bytes(sizes, chosen["contact_tile"], ci["listed_pairs"])
        t = per_iter * 1e-3
        roof = {"bound": "hbm", "kernel": "batch_contact_step", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": alg / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_iteration": alg,
                "listed_pairs": ci["listed_pairs"], "ms_per_iteration": per_iter,
                "edge_update_unit_equivalent_gbs": EDGE_UPDATE_BYTES * E / t / 1e9,
                "limiter": "latency: more work items than fit resident at once, each waiting on its first loads"}
    res, ms2 = b.solve(cap, tol=1e-4, tol_feas=1e-4, check_every=50)
    conv = [r for r in res if r["converged"]]
    done_edge_iters = sum(r["iterations"] * g.nedge for r, g in zip(res, gs))
    print(f"# batch to tolerance: {len(conv)}/{count} converged, launch {ms2:.1f} ms, last retirement "
          f"{max(r['elapsed_ms'] for r in res):.1f} ms", file=sys.stderr, flush=True)
    info = b.info()
    b.close()
    for g in gs:
        g.close()
    # the same batch with the learned rho policy (from rho = 1000, every 50
    # iterations: one rho per instance, the running mean of the network's
    # estimates, plus the edges it raises) against configured fixed rho = 1
    from engine_amd import rho_train as rt
    learned = {}
    for name, rho0, pol in (("fixed_rho1", 1.0, None), ("learned", rt.LEARNED_START_RHO, rt.load_policy())):
        gs = [P.packing_square(n, 1000 * seed + k, rho=rho0, alpha=1.0, a_exp=1.0) for k, n in enumerate(sizes)]
        b = P.Batch(gs)
        b.configure(info["tile"], 1)
        if per_edge:
            b.configure_contact(0, 0, 0)
        elif chosen.get("contact_tile"):
            b.configure_contact(chosen["contact_tile"], chosen["blocks_per_cu"], chosen["barrier"])
        if pol is not None:
            b.set_rho_policy(pol)
            b.set_rho_policy_mode(rt.DEFAULT_MODE, rt.DEFAULT_AVERAGE)
        r2, ms3 = b.solve(cap, tol=1e-4, tol_feas=1e-4, check_every=50, policy_every=50 if pol is not None else 0)
        c2 = [r for r in r2 if r["converged"]]
        learned[name] = {"converged": len(c2), "ms": ms3,
                         "iterations_median_converged": (float(sorted(r["iterations"] for r in c2)[len(c2) // 2])
                                                         if c2 else None)}
        print(f"# batch {name}: {len(c2)}/{count} converged, {ms3:.1f} ms", file=sys.stderr, flush=True)
        b.close()
        if pol is not None:
            pol.close()
        for g in gs:
            g.close()
    return {
        "workload": f"{count} unit-square packings, N ~ U{{16..512}} from batch seed {seed}, rho = clip(N/2, 35, 300), "
                    f"alpha 1",
        "instances": count, "edges": E, "tile": info["tile"], "chosen": chosen,
        "autotune_tried": tried,
        "iters": iters, "ms_per_iter": ms / iters, "edge_updates_per_s": thr,
        "iterations_alone": {"ms_per_iter": per_iter, "edge_updates_per_s": thr_it, "phase_ms": ci["phase_ms"]},
        "roofline": roof,
        "rho_policy_vs_fixed1": learned,
        "to_tol": {"cap": cap, "check_every": 50, "converged": len(conv), "launch_ms": ms2,
                   "edge_updates_per_s": done_edge_iters / (ms2 * 1e-3),
                   "iterations_median": float(sorted(r["iterations"] for r in res)[count // 2])},
    }


TRIANGLE_EDGE_UPDATE_BYTES = 52  # 4 (5 d + 3) at d = 2: the dim-2 counterpart of the 72 B unit


def triangle_bench(P, sizes, seed, iters):
    """The reference demo's own workload (packing/host.c:256-343, N = 2500 as
    its main; packing/gpu.cu runs N = 5000): dim 2, four-edge pair factors,
    three walls per circle, rho 10 (5 on the radius edges), alpha 0.1, on
    the generic layout with the row-tiled pair pass.  Per size: time of
    `iters` iterations, per-kernel means, and the roofline on the 52 B per
    edge-update unit (alpha != 1: every edge carries its own U)."""
    out = []
    for n in sizes:
        t0 = time.time()
        g = P.packing_triangle(n, seed)
        E = g.nedge
        g.iterate(3)
        setup = time.time() - t0
        ms = g.time_iterations(iters) / iters
        kernels = g.profile_kernels(2)
        alg = TRIANGLE_EDGE_UPDATE_BYTES * E
        r = {"n": n, "edges": E, "setup_s": setup, "iters": iters, "ms_per_iter": ms,
             "edge_updates_per_s": E / (ms * 1e-3), "layout": "generic" if g.layout() == 1 else g.layout(),
             "kernels_ms": {k: v[0] for k, v in kernels.items()},
             "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_edge_update": TRIANGLE_EDGE_UPDATE_BYTES,
                          "note": "whole iteration (pair X pass, wall and radius X passes, fused Z pass) against the "
                                  "52 B per edge-update unit"}}
        print(f"# triangle N={n}: E={E}, setup {setup:.1f} s, {ms:.4f} ms/iter, {r['edge_updates_per_s']:.4e} "
              f"edge-updates/s, {r['roofline']['frac']:.2f} of HBM by 52 B/edge-update; kernels {kernels}",
              file=sys.stderr, flush=True)
        out.append(r)
        g.close()
    return out


def learned_bench(P, cases, cap):
    """Time to tolerance (primal, dual, overlap, container <= 1e-4) on one
    GPU: plain fixed rho = 1 (the configured rho), plain fixed rho = 0.4 N
    (the best of the teacher's grid at large N), and the learned policy
    (committed weights applied every 50 iterations from rho = 1000: one rho
    per graph, the running mean of the network's estimates, plus the edges
    the network raises above it)."""
    from engine_amd import rho_train as rt
    m = rt.load_policy()
    out = []
    for n, seed in cases:
        row = {"n": n, "seed": seed, "cap": cap}
        runs = {"fixed_rho1": (1.0, None), f"fixed_rho{0.4 * n:g}": (0.4 * n, None),
                "learned": (rt.LEARNED_START_RHO, m)}
        for name, (rho, pol) in runs.items():
            r = rt.solve_once(n, seed, rho, pol, max_iters=cap, mode=rt.DEFAULT_MODE, average=rt.DEFAULT_AVERAGE)
            row[name] = {"converged": bool(r["converged"]), "iterations": r["iterations"],
                         "ms": r["elapsed_ms"], "objective": r["objective"], "max_overlap": r["max_overlap"]}
        for base in [k for k in runs if k.startswith("fixed")]:
            L, B = row["learned"], row[base]
            if L["converged"] and B["converged"]:
                row[f"speedup_vs_{base}"] = B["ms"] / L["ms"]
            elif L["converged"]:
                row[f"speedup_vs_{base}"] = f">{B['ms'] / L['ms']:.2f} ({base} not converged by the cap)"
            else:
                row[f"speedup_vs_{base}"] = None
        print(f"# learned vs fixed N={n} seed={seed}: " + ", ".join(
            f"{k} {'%d it %.0f ms' % (v['iterations'], v['ms']) if v['converged'] else 'not converged'}"
            for k, v in row.items() if isinstance(v, dict)), file=sys.stderr, flush=True)
        out.append(row)
    m.close()
    return out


def time_leg(P, n, seed, steps, warmup, profile_iters, contact, sparse=None):
    """One single-instance leg: N circles, rho = 1, alpha = 1; the contact
    state (the default for one rho and alpha = 1; sparse: its spatial form
    forced on or off, None: the size rule) or the per-edge path
    (PARADMM_CONTACT=0).  Returns timing, kernel profile and what the
    roofline needs."""
    os.environ["PARADMM_CONTACT"] = "1" if contact else "0"
    if sparse is None:
        os.environ.pop("PARADMM_CONTACT_SPARSE", None)
    else:
        os.environ["PARADMM_CONTACT_SPARSE"] = "1" if sparse else "0"
    t0 = time.time()
    g = P.packing_square(n, seed, rho=1.0, alpha=1.0, a_exp=1.0)
    g.prepare()
    t_setup = time.time() - t0
    E = g.nedge
    g.time_iterations(warmup)
    ms = g.time_iterations(steps)
    prof = g.profile_kernels(profile_iters)
    in_c, cs, listed = g.contact_info()
    p2, d2 = g.residuals()
    layout = {1: "generic", 2: "dense"}.get(g.layout(), "?")
    g.close()
    per_iter = {k: v[0] * v[1] / profile_iters for k, v in prof.items()}
    dom = max(per_iter, key=per_iter.get)
    form = ("spatial" if "scs_rows" in prof else "id-order") if in_c else None
    os.environ.pop("PARADMM_CONTACT_SPARSE", None)
    leg = {"n": n, "edges": E, "setup_s": t_setup, "steps": steps, "ms": ms, "ms_per_step": ms / steps,
           "edge_updates_per_s": E * steps / (ms * 1e-3), "layout": layout,
           "state": "contact" if in_c else "per-edge", "contact_form": form,
           "contact_tile_side": cs, "listed_pairs": listed,
           "kernels_ms": {k: v[0] for k, v in prof.items()}, "dominant": dom, "dominant_ms": prof[dom][0],
           "residuals_last": {"primal2": p2, "dual2": d2}}
    print(f"# N={n} {leg['state']}: setup {t_setup:.1f} s, {steps} iterations {ms / steps:.4f} ms/iter, "
          f"{leg['edge_updates_per_s']:.4e} edge-updates/s; kernels (mean ms x launches): "
          + ", ".join(f"{k} {v[0]:.4f}x{v[1]}" for k, v in prof.items()), file=sys.stderr, flush=True)
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=5000, help="circles")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profile-iters", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-edge", action="store_true", help="skip the per-edge (PARADMM_CONTACT=0) leg")
    ap.add_argument("--no-n1000", action="store_true", help="skip the N = 1000 leg")
    ap.add_argument("--no-id-order", action="store_true",
                    help="skip the id-order contact form leg on the headline workload")
    ap.add_argument("--large-n", type=int, default=10000,
                    help="circles of the leg timing both contact forms at a larger packing (0: skip)")
    ap.add_argument("--batch", type=int, default=256, help="instances of the batched workload (0: skip)")
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-iters", type=int, default=1000)
    ap.add_argument("--batch-cap", type=int, default=20000)
    ap.add_argument("--batch-per-edge", action="store_true",
                    help="the batch on the per-edge kernel instead of the contact state")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-fixed time to tolerance")
    ap.add_argument("--triangle", type=str, default="2500,5000",
                    help="circles of the reference demo's triangle workload legs (empty: skip)")
    ap.add_argument("--host-loop-iters", type=int, default=0,
                    help="also time this many iterations of the legacy per-pass host loop")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the single-instance solve does not shard across GPUs here).", file=sys.stderr)
        sys.exit(2)

    from engine_amd import paradmm as P

    # the headline: the default path (the contact state for one rho and alpha = 1)
    head = time_leg(P, args.n, args.seed, args.steps, args.warmup, args.profile_iters, contact=True)
    E = head["edges"]
    value = head["edge_updates_per_s"]
    copy_gbs = P.copy_bandwidth(1 << 30, 20)
    print(f"# plain copy kernel (1 GiB): {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    dom, dom_ms = head["dominant"], head["dominant_ms"]
    traffic = measured_traffic(dom, args.n, args.seed)
    valu = measured_valu(dom, args.n, args.seed)
    if head["state"] == "contact" and dom == "scs_rows":
        alg = spatial_alg_bytes(args.n, head["listed_pairs"])
        roof = roofline_spatial(dom_ms, alg, traffic, valu, copy_gbs, E, head["listed_pairs"],
                                head["kernels_ms"].get("scs_circles"))
    elif head["state"] == "contact":
        alg = contact_alg_bytes(args.n, head["contact_tile_side"], head["listed_pairs"])
        roof = roofline_contact(dom_ms, alg, traffic, valu, copy_gbs, args.n, E, head["listed_pairs"])
    else:
        kb = UNIFORM_BYTES_PER_EDGE.get(dom, BYTES_PER_EDGE.get(dom, 0)) * E
        roof = roofline_per_edge(dom, dom_ms, kb, traffic, copy_gbs, E)

    # the GPU's Z over its first iterations from the seeded start (checked
    # against the FP64 oracle in the cpu_baseline leg)
    import numpy as np
    os.environ["PARADMM_CONTACT"] = "1"
    g = P.packing_square(args.n, args.seed, rho=1.0, alpha=1.0, a_exp=1.0)
    zhist = np.zeros((5, 3 * args.n))
    for it in range(5):
        g.iterate(1)
        zhist[it] = g.read("Z").ravel()
    host_loop = None
    if args.host_loop_iters > 0:
        # the reference's own loop shape (packing/host.c:355-360): updateXM,
        # updateZ, updateUN through the legacy C API, which keeps the FP64
        # host arrays coherent on every call (PCIe both ways, host-side
        # AoS <-> SoA conversion included)
        g.sync_to_host()
        g.legacy("updateXM")
        t0 = time.time()
        for _ in range(args.host_loop_iters):
            g.legacy("updateXM")
            g.legacy("updateZ")
            g.legacy("updateUN")
        dt = time.time() - t0
        host_loop = {"iterations": args.host_loop_iters, "s_per_iter": dt / args.host_loop_iters,
                     "edge_updates_per_s": E * args.host_loop_iters / dt}
        print(f"# legacy host loop (PCIe-inclusive): {dt / args.host_loop_iters * 1e3:.1f} ms/iter, "
              f"{host_loop['edge_updates_per_s']:.3e} edge-updates/s", file=sys.stderr, flush=True)
    g.close()

    line = {
        "metric": "edge-updates/s (%HBM peak) + time to residual <=1e-4, learned-rho vs fixed-rho",
        "value": value,
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain ADMM unit-square circle packing, N={args.n} circles, E={E} edges of dim 3, "
                        f"fixed rho=1, alpha=1, radius exponent 1, seed {args.seed}",
            "circles": args.n, "edges": E, "layout": head["layout"], "state": head["state"],
            "contact_form": head["contact_form"],
        },
        "roofline": roof,
        "kernels_ms": head["kernels_ms"],
        "residuals_last": head["residuals_last"],
    }
    if not args.no_per_edge:
        pe = time_leg(P, args.n, args.seed, args.steps, args.warmup, args.profile_iters, contact=False)
        kb = UNIFORM_BYTES_PER_EDGE.get(pe["dominant"], BYTES_PER_EDGE.get(pe["dominant"], 0)) * pe["edges"]
        pe["roofline"] = roofline_per_edge(pe["dominant"], pe["dominant_ms"], kb,
                                           measured_traffic(pe["dominant"], args.n, args.seed), copy_gbs, pe["edges"])
        line["per_edge_path"] = pe
    if head["contact_form"] == "spatial" and not args.no_id_order:
        # the contact state's id-order form on the same workload
        io = time_leg(P, args.n, args.seed, args.steps, args.warmup, args.profile_iters, contact=True, sparse=False)
        io["roofline"] = roofline_contact(io["dominant_ms"], contact_alg_bytes(args.n, io["contact_tile_side"],
                                                                              io["listed_pairs"]),
                                          measured_traffic(io["dominant"], args.n, args.seed),
                                          measured_valu(io["dominant"], args.n, args.seed), copy_gbs, args.n,
                                          io["edges"], io["listed_pairs"])
        line["id_order_contact"] = io
    if not args.no_n1000 and args.n != 1000:
        line["n1000"] = time_leg(P, 1000, args.seed, 10000, 200, args.profile_iters, contact=True)
    if args.large_n:
        # the spatial form's O(N) iteration against the id-order form's O(N^2) at a larger packing
        line["large_n"] = [time_leg(P, args.large_n, args.seed, args.steps, args.warmup, args.profile_iters,
                                    contact=True, sparse=sp) for sp in (True, False)]
    os.environ.pop("PARADMM_CONTACT", None)
    if host_loop:
        line["host_loop_pcie_inclusive"] = host_loop
    if args.triangle:
        line["triangle_demo"] = triangle_bench(P, [int(x) for x in args.triangle.split(",")], args.seed, 20)
    if args.batch > 0:
        line["batch"] = batch_bench(P, args.batch, args.batch_seed, args.batch_iters, args.batch_cap,
                                    per_edge=args.batch_per_edge)
    if not args.no_learned:
        line["time_to_tol_learned_vs_fixed"] = learned_bench(P, ((5000, 3), (2000, 3), (2000, 4), (1000, 3), (1000, 4)),
                                                             30000)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], line["oracle_error"] = cpu_baseline(args.n, zhist, args.cpu_threads)
    P.release_all()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
