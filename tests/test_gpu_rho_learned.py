"""What the trained rho network (engine_amd/models/rho_policy.f32) gains over
the plain method without it, on the GPU, judged by the plain method's own
feasibility and objective checks."""
import pytest

from engine_amd import rho_train as rt

pytestmark = pytest.mark.gpu

CAP = 30000


def test_learned_rho_converges_where_configured_fixed_rho_does_not():
    """N = 1000, seed 4: the configured fixed rho = 1 does not reach the
    tolerance within the cap; the learned policy does, to a feasible packing
    whose objective (sum of radii, maximised) is no more than 1% below that of
    fixed rho = 0.4 N (the best of the grid at this size)."""
    n, seed = 1000, 4
    m = rt.load_policy()
    fixed1 = rt.solve_once(n, seed, 1.0, None, max_iters=CAP)
    learned = rt.solve_once(n, seed, rt.LEARNED_START_RHO, m, max_iters=CAP, mode=rt.DEFAULT_MODE,
                            average=rt.DEFAULT_AVERAGE)
    best = rt.solve_once(n, seed, 0.4 * n, None, max_iters=CAP)
    print(f"fixed rho=1: {fixed1['iterations']} it converged={fixed1['converged']}; "
          f"learned: {learned['iterations']} it {learned['elapsed_ms']:.0f} ms; "
          f"fixed rho={0.4 * n:g}: {best['iterations']} it {best['elapsed_ms']:.0f} ms")
    assert fixed1["converged"] == 0
    assert learned["converged"] == 1
    assert learned["max_overlap"] <= 1e-4 and learned["max_box"] <= 1e-4
    assert learned["primal_rms"] <= 1e-4 and learned["dual_rms"] <= 1e-4
    assert best["converged"] == 1
    assert learned["objective"] >= 0.99 * best["objective"]


def test_learned_rho_at_2000_judged_by_the_plain_method():
    """N = 2000, seed 3 (the learned-rho workload): the learned solve reaches
    the tolerance; the FP64 oracle's own quality check on the Z it leaves
    finds no pair overlap and no container violation above 1e-4, and its
    objective is no more than 1% below that of the plain fixed-rho solve
    (rho = 0.4 N) on the same instance, which converges too."""
    from engine_amd import paradmm as P
    from oracle import pyoracle as po
    n, seed = 2000, 3
    m = rt.load_policy()
    g = P.packing_square(n, seed, rho=rt.LEARNED_START_RHO, alpha=1.0)
    g.prepare()
    g.set_rho_policy(m)
    g.set_rho_policy_mode(rt.DEFAULT_MODE, rt.DEFAULT_AVERAGE)
    learned = g.solve(CAP, tol=1e-4, tol_feas=1e-4, check_every=50, policy_every=50)
    z = g.read("Z").ravel()
    g.close()
    m.close()
    plain = rt.solve_once(n, seed, 0.4 * n, None, max_iters=CAP)
    o = po.Graph.square(n, seed)
    o.array("Z")[:] = z
    ov, bx, sr = o.square_quality(n)
    o.close()
    print(f"N={n} seed {seed}: learned {learned['iterations']} it {learned['elapsed_ms']:.0f} ms objective "
          f"{learned['objective']:.4f} (oracle check: overlap {ov:.2e} container {bx:.2e} sum r {sr:.4f}); "
          f"fixed rho={0.4 * n:g} {plain['iterations']} it {plain['elapsed_ms']:.0f} ms objective "
          f"{plain['objective']:.4f}")
    assert learned["converged"] == 1 and plain["converged"] == 1
    assert ov <= 1e-4 and bx <= 1e-4
    assert sr == pytest.approx(learned["objective"], rel=1e-5)
    assert learned["objective"] >= 0.99 * plain["objective"]
