"""Training and evaluation of the learned rho policy.

The data comes from the plain solver itself.  Along a solver trajectory the
state is checkpointed every `window` iterations; from each checkpoint the
solver is rolled forward `window` iterations once per candidate uniform rho
multiplier (0.25x .. 4x, with the scaled dual rescaled) and the multiplier
whose rollout ends with the lowest score

    score = mean over the window's two checks of log10(max(primal, dual))
            + 0.5 * log10(max(max pair overlap, 1e-6))

is the label: the target of every edge at that checkpoint is
log10(rho_e * best multiplier).  The trajectory then continues from the
checkpoint with the best multiplier (a greedy lookahead teacher).  The
network (6 -> 32 -> 32 -> 1, bf16 MFMA) is trained on (edge features,
target) rows with the library's own GPU Adam kernels.

Everything runs through libparadmm's C ABI (paradmm_gpu.h).
"""
import json
import os
import time

import numpy as np

from . import paradmm as P

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "models", "rho_policy.f32")
MULTIPLIERS = (0.25, 0.5, 1.0, 2.0, 4.0)


def _score(g, iters):
    half = iters // 2
    logs = []
    for n in (half, iters - half):
        r = g.solve(n, tol=0.0, tol_feas=-1.0, check_every=n)
        logs.append(np.log10(max(r["primal_rms"], r["dual_rms"], 1e-12)))
    ov, _, _ = g.square_quality()
    return float(np.mean(logs)) + 0.5 * float(np.log10(max(ov, 1e-6)))


def generate(sizes, seeds, rho0s, snapshots=16, window=250, rows_per_snapshot=2000, rng_seed=0, log=print):
    """Rows of (features, target) from greedy counterfactual rollouts."""
    rng = np.random.default_rng(rng_seed)
    feats, targets = [], []
    total = snapshots * window
    for n in sizes:
        for seed in seeds:
            rho0 = float(rho0s[rng.integers(len(rho0s))])
            g = P.packing_square(int(n), int(seed), rho=rho0, alpha=1.0)
            g.prepare()
            rho = rho0
            t0 = time.time()
            for s in range(snapshots):
                it = s * window
                f = g.rho_features(it / total)
                g.state_save()
                scores = []
                for m in MULTIPLIERS:
                    g.state_restore()
                    if m != 1.0:
                        g.scale_rho(m)
                    scores.append(_score(g, window))
                best = MULTIPLIERS[int(np.argmin(scores))]
                take = rng.choice(f.shape[0], size=min(rows_per_snapshot, f.shape[0]), replace=False)
                feats.append(f[take])
                # rho is uniform along these trajectories: feature 2 is log10(rho)/2
                targets.append((f[take, 2] * 2.0 + np.log10(best)).astype(np.float32))
                g.state_restore()
                if best != 1.0:
                    g.scale_rho(best)
                rho *= best
                g.iterate(window)
            log(f"# data N={n} seed={seed} rho0={rho0:g}: final rho {rho:g}, {time.time() - t0:.1f} s")
            g.close()
    return np.concatenate(feats).astype(np.float32), np.concatenate(targets).astype(np.float32)


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


def solve_once(n, seed, rho, policy=None, max_iters=20000, tol=1e-4, check_every=50, policy_every=50):
    """One solve to tolerance; policy None is the plain fixed-rho method."""
    g = P.packing_square(n, seed, rho=rho, alpha=1.0)
    g.prepare()
    if policy is not None:
        g.set_rho_policy(policy)
    r = g.solve(max_iters, tol=tol, tol_feas=tol, check_every=check_every,
                policy_every=policy_every if policy is not None else 0)
    g.close()
    return r
