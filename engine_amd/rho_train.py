"""Training and evaluation of the learned rho policy.

Training data come from the plain solver itself:

1. Teacher.  For each training instance (unit-square packing, N circles,
   seed) the plain fixed-rho solver is run to tolerance (primal and dual RMS
   residual <= tol, max overlap and container violation <= tol) for every rho
   of a grid; the teacher's rho* is the fastest to converge (ties and
   near-ties within 10% broken by the larger objective sum r); if none
   converges, the rho with the smallest final max(primal, dual, overlap).
2. Rows.  From several starting rho (1, 10, rho* and 4 rho*) the solver
   runs the whole iteration budget; at snapshot iterations spread over it
   (dense early, SNAPSHOTS) the per-edge features of
   paradmm_gpu_rho_features are sampled, each with target log10(rho*).  The
   network thus learns which rho to move to from the dynamics it observes,
   wherever it starts and at any fraction of the budget.
3. Training with the library's own bf16-MFMA forward/backward and Adam.

Applied inside paradmm_gpu_solve every `policy_every` iterations, the
network sets per-edge rho; the plain fixed-rho solver stays the fallback and
its feasibility and objective judge every learned result.
"""
import json
import os
import time

import numpy as np

from . import paradmm as P

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "models", "rho_policy.f32")
# how the committed policy is applied (paradmm_gpu_set_rho_policy_mode):
# one rho per graph, the running mean of the network's estimates over the
# applications so far, plus the edges the network raises above it
DEFAULT_MODE, DEFAULT_AVERAGE = 3, True
DEFAULT_MAX_ITERS = 30000
RHO_GRID = (20.0, 35.0, 50.0, 70.0, 100.0, 150.0, 200.0, 300.0, 500.0, 800.0, 1200.0)
# snapshot iterations of the training trajectories: dense early, then spread
# over the whole budget so the iteration-fraction input is covered on [0, 1]
SNAPSHOTS = tuple(range(50, 1001, 50)) + (1500, 2000, 3000, 4000, 6000, 8000, 10000, 13000, 16000, 20000)


# rho of the first policy_every iterations of a learned solve: a start above
# any instance's optimum (a start far below it drives the packing into
# overlaps it does not leave), which the first application then lowers
LEARNED_START_RHO = 1000.0


def solve_once(n, seed, rho, policy=None, max_iters=20000, tol=1e-4, check_every=50, policy_every=50,
               mode=0, average=False):
    """One solve to tolerance; policy None is the plain fixed-rho method.
    mode / average: how the network's outputs are applied
    (paradmm_gpu_set_rho_policy_mode)."""
    g = P.packing_square(n, seed, rho=rho, alpha=1.0)
    g.prepare()
    if policy is not None:
        g.set_rho_policy(policy)
        g.set_rho_policy_mode(mode, average)
    r = g.solve(max_iters, tol=tol, tol_feas=tol, check_every=check_every,
                policy_every=policy_every if policy is not None else 0)
    g.close()
    return r


def best_fixed_rho(n, seed, grid=RHO_GRID, max_iters=20000, tol=1e-4):
    """The teacher: fastest converging fixed rho of the grid for this instance."""
    runs = {rho: solve_once(n, seed, rho, None, max_iters=max_iters, tol=tol) for rho in grid}
    conv = {rho: r for rho, r in runs.items() if r["converged"]}
    if conv:
        fastest = min(r["iterations"] for r in conv.values())
        near = {rho: r for rho, r in conv.items() if r["iterations"] <= 1.1 * fastest}
        best = max(near, key=lambda rho: near[rho]["objective"])
    else:
        best = min(runs, key=lambda rho: max(runs[rho]["primal_rms"], runs[rho]["dual_rms"],
                                              max(runs[rho]["max_overlap"], 0.0)))
    return best, runs


STARTS = (("abs", LEARNED_START_RHO), ("rel", 1.0), ("rel", 0.7), ("rel", 4.0))

# Raised edges.  Late in a solve (after RAISE_FROM of the iteration budget) a
# pair edge whose primal residual |X_e - Z_v| is still above
# RAISE_RESIDUAL is labelled RAISE_FACTOR x rho*: such edges are pairs of
# small circles caught in a two-cycle of the nonconvex pair projection (the
# separation direction flips every iteration, the dual never accumulates and
# Z does not move); a higher rho on those edges alone pulls Z their way and
# ends the cycle.  Every such row of a snapshot is kept (up to RAISE_ROWS),
# RAISE_WEIGHT times.
RAISE_FROM = 5000 / 30000
RAISE_RESIDUAL = 1e-4
RAISE_FACTOR = 10.0
RAISE_ROWS, RAISE_WEIGHT = 2000, 3


def _label(f, rho_star, rng, rows):
    """Rows sampled from one snapshot's features f and their targets:
    log10 rho*, plus log10 RAISE_FACTOR on the raised edges."""
    take = rng.choice(f.shape[0], size=min(rows, f.shape[0]), replace=False)
    if f.shape[0] and f[0, 5] >= RAISE_FROM:
        hot = np.flatnonzero((f[:, 0] > np.log10(RAISE_RESIDUAL) / 5) & (f[:, 3] == 1.0))[:RAISE_ROWS]
        if hot.size:
            take = np.concatenate([take] + [hot] * RAISE_WEIGHT)
    t = np.full(take.size, np.log10(rho_star), dtype=np.float32)
    fr = f[take]
    t[(fr[:, 0] > np.log10(RAISE_RESIDUAL) / 5) & (fr[:, 3] == 1.0) & (fr[:, 5] >= RAISE_FROM)] += np.log10(RAISE_FACTOR)
    return fr, t


def generate(instances, starts=STARTS, snapshots=SNAPSHOTS, max_iters=20000,
             rows_per_snapshot=1000, rng_seed=0, log=print, grid=RHO_GRID, first_weight=1, known=None):
    """Rows of (features, target) for the given (n, seed) instances.  Each
    start is ("abs", rho) or ("rel", multiple of the teacher's rho*).  The
    first start's first three snapshots (the states a learned solve sees at
    its first applications) get first_weight times the rows.  known: an
    earlier teacher record ({"n/seed": {"rho_star": ...}}) whose rho* is
    reused instead of running the grid again."""
    rng = np.random.default_rng(rng_seed)
    feats, targets, teacher = [], [], {}
    for n, seed in instances:
        t0 = time.time()
        key = f"{n}/{seed}"
        if known is not None and key in known:
            rho_star = known[key]["rho_star"]
            teacher[key] = known[key]
        else:
            rho_star, runs = best_fixed_rho(n, seed, grid=grid, max_iters=max_iters)
            teacher[key] = {"rho_star": rho_star,
                            "iterations": {str(k): (v["iterations"] if v["converged"] else None)
                                           for k, v in runs.items()}}
        for si, (kind, v) in enumerate(starts):
            rho0 = v * rho_star if kind == "rel" else v
            g = P.packing_square(int(n), int(seed), rho=rho0, alpha=1.0)
            g.prepare()
            done = 0
            for k, it in enumerate(snapshots):
                g.iterate(it - done)
                done = it
                f = g.rho_features(it / max_iters)
                rows = rows_per_snapshot * (first_weight if si == 0 and k < 3 else 1)
                fr, t = _label(f, rho_star, rng, rows)
                feats.append(fr)
                targets.append(t)
            g.close()
        log(f"# teacher N={n} seed={seed}: rho*={rho_star:g} "
            f"({teacher[f'{n}/{seed}']['iterations']}), {time.time() - t0:.1f} s")
    return np.concatenate(feats).astype(np.float32), np.concatenate(targets).astype(np.float32), teacher


def generate_on_policy(instances, policy, teacher, max_iters=20000, every=50, sample_every=500,
                       rows_per_snapshot=1000, mode=DEFAULT_MODE, average=DEFAULT_AVERAGE, rng_seed=1, log=print):
    """Rows from the states the learned policy itself visits (one round of
    dataset aggregation): the policy is applied every `every` iterations
    from LEARNED_START_RHO as in a learned solve; every `sample_every`
    iterations the features are sampled, labelled with the instance's
    teacher rho* (teacher: {"n/seed": rho*})."""
    rng = np.random.default_rng(rng_seed)
    feats, targets = [], []
    for n, seed in instances:
        rho_star = teacher[f"{n}/{seed}"]
        g = P.packing_square(int(n), int(seed), rho=LEARNED_START_RHO, alpha=1.0)
        g.prepare()
        g.set_rho_policy_mode(mode, average)
        for it in range(every, max_iters + 1, every):
            g.iterate(every)
            if it % sample_every == 0 or it <= 4 * every:
                f = g.rho_features(it / max_iters)
                fr, t = _label(f, rho_star, rng, rows_per_snapshot)
                feats.append(fr)
                targets.append(t)
            g.apply_rho_policy(policy, it / max_iters)
        g.sync_to_host()
        log(f"# on-policy N={n} seed={seed}: rho*={rho_star:g}, final rho {g.array('RHOS')[0]:.1f}")
        g.close()
    return np.concatenate(feats).astype(np.float32), np.concatenate(targets).astype(np.float32)


def save_rows(path, feat, target, teacher=None):
    """Training rows (features [n, 6] float32, targets [n] float32) and the
    teacher record in one .npz (plain arrays; the record as JSON text), so a
    dataset is generated once and reused."""
    np.savez(path, feat=np.asarray(feat, dtype=np.float32), target=np.asarray(target, dtype=np.float32),
             teacher=np.array(json.dumps(teacher or {})))


def load_rows(path):
    """(feat, target, teacher) written by save_rows; nothing is unpickled."""
    with np.load(path, allow_pickle=False) as z:
        return z["feat"], z["target"], json.loads(str(z["teacher"]))


def train(feat, target, epochs=30, batch=4096, lr=2e-3, seed=0):
    m = P.RhoMLP(seed=seed)
    loss = m.train(feat, target, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, loss


def save(m, path=DEFAULT_WEIGHTS, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    m.params().astype("<f4").tofile(path)
    if meta is not None:
        with open(path + ".json", "w") as fh:
            json.dump(meta, fh, indent=1, sort_keys=True)


def load_policy(path=DEFAULT_WEIGHTS):
    p = np.fromfile(path, dtype="<f4")
    if p.size != P.RhoMLP.NPARAMS:
        raise ValueError(f"{path}: expected {P.RhoMLP.NPARAMS} parameters, found {p.size}")
    return P.RhoMLP(params=p)


LEARNED_VARIANTS = {"learned": (0, False), "learned_type": (1, False), "learned_global": (2, False),
                    "learned_global_avg": (2, True), "learned_raise": (3, False), "learned_raise_avg": (3, True)}


def compare(n, seeds, policy, fixed=(1.0, 50.0, 100.0), max_iters=20000, log=print,
            variants=("learned_raise_avg",)):
    """Time to tolerance of fixed rho values and of the learned policy (the
    given application variants, from LEARNED_START_RHO)."""
    out = []
    for seed in seeds:
        row = {"n": n, "seed": seed}
        for rho in fixed:
            row[f"fixed{rho:g}"] = solve_once(n, seed, rho, None, max_iters=max_iters)
        for v in variants:
            mode, avg = LEARNED_VARIANTS[v]
            row[v] = solve_once(n, seed, LEARNED_START_RHO, policy, max_iters=max_iters, mode=mode, average=avg)
        out.append(row)
        log(f"# N={n} seed={seed}: " + ", ".join(
            f"{k} {'%d it %.0f ms' % (v['iterations'], v['elapsed_ms']) if v['converged'] else 'not converged'}"
            f" obj {v['objective']:.4f}" for k, v in row.items() if isinstance(v, dict)))
    return out
