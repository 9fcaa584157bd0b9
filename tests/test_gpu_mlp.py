"""GPU checks of the learned-rho network (bf16 MFMA forward/backward, fp32
accumulation) against the FP64 MLP oracle, and of the policy applied inside
the solver.  Tolerances: the oracle's bf16 mode rounds the same operands the
kernels round, so the remaining difference is fp32 accumulation order plus an
occasional one-ulp bf16 flip of an activation near a rounding boundary."""
import numpy as np
import pytest

from engine_amd import paradmm as P
from gpu_helpers import max_rel
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def test_mfma_layout_probe():
    """v_mfma_f32_16x16x16_bf16 lane layout used by every MLP kernel."""
    d = P.mfma_layout_probe()
    i = np.arange(16)[:, None]
    k = np.arange(16)[None, :]
    A = mo.bf16(i - k * 0.25)
    kk = np.arange(16)[:, None]
    j = np.arange(16)[None, :]
    B = mo.bf16((kk + 2 * j) * 0.125)
    assert np.allclose(d, A @ B, rtol=1e-6, atol=1e-5)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 6)).astype(np.float32)
    X[:, 3] = rng.integers(0, 2, n)
    X[:, 4] = (1 - X[:, 3]) * rng.integers(0, 2, n)
    X[:, 5] = rng.uniform(0, 1, n)
    t = (np.tanh(X[:, 0] - X[:, 1]) + 0.5 * X[:, 3] + 0.3 * X[:, 5]).astype(np.float32)
    return X, t


@pytest.mark.parametrize("n", [1, 17, 1000, 5003])
def test_forward_vs_oracle(n):
    p = mo.init_params(7)
    p += np.random.default_rng(1).normal(0, 0.05, p.shape)
    m = P.RhoMLP(params=p.astype(np.float32))
    X, _ = _data(n, 2)
    y = m.forward(X)
    yo, _ = mo.forward(p.astype(np.float32).astype(np.float64), X.astype(np.float64), exact=False)
    err = np.abs(y - yo)
    assert np.median(err) < 1e-4
    assert np.max(err) < 2e-2 * max(1.0, np.max(np.abs(yo)))


@pytest.mark.parametrize("n", [64, 4096, 10000])
def test_gradient_vs_oracle(n):
    p = (mo.init_params(3) + np.random.default_rng(4).normal(0, 0.05, mo.NPARAMS)).astype(np.float32)
    m = P.RhoMLP(params=p)
    X, t = _data(n, 5)
    loss, g = m.gradient(X, t)
    lo, go = mo.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64), exact=False)
    _, ge = mo.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64), exact=True)
    assert loss == pytest.approx(lo, rel=1e-2)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    rel_exact = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    print(f"n={n}: gradient rel err vs bf16-emulating oracle {rel:.2e}, vs exact FP64 {rel_exact:.2e}")
    assert rel < 1e-2
    assert rel_exact < 5e-2
    # every block of the parameter vector is exercised
    for name, sl in [("W1", slice(0, 192)), ("W2", slice(224, 1248)), ("W3", slice(1280, 1312)), ("b1", slice(192, 224))]:
        assert np.linalg.norm(g[sl] - go[sl]) <= 2e-2 * np.linalg.norm(go[sl]) + 1e-7, name


def test_training_reduces_loss_like_oracle():
    X, t = _data(20000, 9)
    p0 = mo.init_params(11).astype(np.float32)
    m = P.RhoMLP(params=p0)
    l0, _ = m.gradient(X, t)
    final = m.train(X, t, epochs=20, batch=1024, lr=3e-3, seed=1)
    l1, _ = m.gradient(X, t)
    print(f"GPU Adam: loss {l0:.4f} -> {l1:.4f} (last epoch mean {final:.4f})")
    assert l1 < 0.1 * l0


def test_policy_application_matches_network():
    g = P.packing_square(60, 4, rho=5.0, alpha=1.0)
    g.iterate(30)
    p = (mo.init_params(2) * 0.3).astype(np.float32)
    m = P.RhoMLP(params=p)
    feat = g.rho_features(0.25)
    g.sync_to_host()
    # features recomputed on the host from the synced state
    ev = g.edge_var()
    X = g.array("X").reshape(-1, 3)
    Z = g.array("Z").reshape(-1, 3)
    rho = g.array("RHOS").copy()
    U0 = g.array("U").reshape(-1, 3).copy()
    r = np.sqrt(np.sum((X - Z[ev]) ** 2, 1))
    assert np.allclose(feat[:, 0], np.log10(np.maximum(r, 1e-10)) * 0.2, atol=2e-4)
    assert np.allclose(feat[:, 2], np.log10(rho) * 0.5, atol=1e-6)
    n = 60
    typ = np.zeros(g.nedge)
    typ[:n] = 2
    typ[n:2 * n] = 1
    assert np.array_equal(feat[:, 3], (typ == 0).astype(np.float32))
    assert np.array_equal(feat[:, 4], (typ == 2).astype(np.float32))
    assert np.all(feat[:, 5] == np.float32(0.25))
    y = m.forward(feat)
    g.apply_rho_policy(m, 0.25)
    g.sync_to_host()
    rho_new = g.array("RHOS")
    expect = 10.0 ** np.clip(y, -1.0, 4.0)
    assert np.allclose(rho_new, expect, rtol=1e-5)
    U1 = g.array("U").reshape(-1, 3)
    assert np.allclose(U1, U0 * (rho / rho_new)[:, None], rtol=1e-5, atol=1e-7)


def test_policy_per_type_mode_applies_type_means():
    n = 48
    g = P.packing_square(n, 6, rho=20.0, alpha=1.0)
    g.iterate(40)
    p = (mo.init_params(3) * 0.3).astype(np.float32)
    m = P.RhoMLP(params=p)
    feat = g.rho_features(0.1)
    g.sync_to_host()
    rho = g.array("RHOS").copy()
    U0 = g.array("U").reshape(-1, 3).copy()
    y = np.clip(m.forward(feat).astype(np.float64), -1.0, 4.0)
    typ = np.zeros(g.nedge, int)
    typ[:n] = 2
    typ[n:2 * n] = 1
    g.set_rho_policy_mode(True)
    g.apply_rho_policy(m, 0.1)
    g.sync_to_host()
    rho_new = g.array("RHOS")
    for t in (0, 1, 2):
        expect = 10.0 ** y[typ == t].mean()
        assert np.allclose(rho_new[typ == t], expect, rtol=1e-5)
    U1 = g.array("U").reshape(-1, 3)
    assert np.allclose(U1, U0 * (rho / rho_new)[:, None], rtol=1e-5, atol=1e-7)


def test_policy_global_mode_with_average():
    """mode 2: one rho for every edge, 10^(mean of all outputs); with
    averaging, the second application applies the mean of both estimates."""
    n = 40
    g = P.packing_square(n, 2, rho=30.0, alpha=1.0)
    g.iterate(25)
    m = P.RhoMLP(params=(mo.init_params(4) * 0.3).astype(np.float32))
    g.set_rho_policy_mode(2, average=True)
    est = []
    for k in range(2):
        feat = g.rho_features(0.1 * (k + 1))
        est.append(np.clip(m.forward(feat).astype(np.float64), -1.0, 4.0).mean())
        g.apply_rho_policy(m, 0.1 * (k + 1))
        g.sync_to_host()
        rho = g.array("RHOS")
        assert np.allclose(rho, 10.0 ** np.mean(est), rtol=1e-5)
        g.iterate(10)


@pytest.mark.parametrize("layout", ["dense", "generic"])
def test_policy_raise_mode_keeps_edges_the_network_raises(monkeypatch, layout):
    """mode 3: every edge takes 10^(mean of all outputs), except an edge whose
    own output exceeds that mean by more than 0.5 (log10), which keeps
    10^(its output); U is rescaled by rho_old / rho_new.  Both layouts."""
    if layout == "generic":
        monkeypatch.setenv("PARADMM_LAYOUT", "generic")
    n = 40
    g = P.packing_square(n, 2, rho=30.0, alpha=1.0)
    g.iterate(25)
    m = P.RhoMLP(params=(mo.init_params(7) * 1.5).astype(np.float32))
    feat = g.rho_features(0.3)
    g.sync_to_host()
    rho0 = g.array("RHOS").copy()
    U0 = g.array("U").reshape(-1, 3).copy()
    y = np.clip(m.forward(feat).astype(np.float64), -1.0, 4.0)
    mean = y.mean()
    g.set_rho_policy_mode(3)
    g.apply_rho_policy(m, 0.3)
    g.sync_to_host()
    rho = g.array("RHOS")
    up = y - mean > 0.5 + 1e-4
    down = y - mean < 0.5 - 1e-4
    print(f"raised {up.sum()} of {y.size}, mean log10 rho {mean:.3f}")
    assert up.sum() > 0 and down.sum() > 0
    assert np.allclose(rho[up], 10.0 ** y[up], rtol=1e-5)
    assert np.allclose(rho[down], 10.0 ** mean, rtol=1e-5)
    U1 = g.array("U").reshape(-1, 3)
    assert np.allclose(U1, U0 * (rho0 / rho)[:, None], rtol=1e-5, atol=1e-7)
    assert g.layout() == (2 if layout == "dense" else 1)
    # the raised rho is used by the following iterations
    g.iterate(5)
    g.sync_to_host()
    assert np.all(np.isfinite(g.array("Z")))


def test_raised_tiles_match_the_per_edge_variant(monkeypatch):
    """After a mode-3 application with raised edges the dense layout keeps
    its uniform-rho kernel and sends only the tiles holding raised edges
    through the per-edge streams; iterations match the per-edge kernel run
    on the same state (re-uploaded from the host, which selects it).  On
    the per-edge path (the contact state reads every processed pair's own
    rho instead; test_gpu_contact.py)."""
    monkeypatch.setenv("PARADMM_CONTACT", "0")
    n = 70
    m = P.RhoMLP(params=(mo.init_params(7) * 1.5).astype(np.float32))
    gs = []
    for k in range(2):
        g = P.packing_square(n, 5, rho=30.0, alpha=1.0)
        g.iterate(25)
        g.set_rho_policy_mode(3)
        g.apply_rho_policy(m, 0.3)
        g.sync_to_host()
        gs.append(g)
    r = gs[0].array("RHOS")
    assert np.unique(r).size > 1  # some edges raised
    gs[1].sync_from_host()
    for g in gs:
        g.iterate(5)
        g.sync_to_host()
    assert max_rel(gs[0].array("Z"), gs[1].array("Z")) < 1e-6


def test_state_save_restore_and_scale():
    g = P.packing_square(50, 1, rho=10.0)
    g.iterate(10)
    g.state_save()
    g.iterate(7)
    g.sync_to_host()
    z7 = g.array("Z").copy()
    g.state_restore()
    g.iterate(7)
    g.sync_to_host()
    assert np.array_equal(g.array("Z"), z7)
    g.state_restore()
    g.scale_rho(2.0)
    g.sync_to_host()
    assert np.allclose(g.array("RHOS"), 20.0)


def test_solve_reports_quality_and_converges_at_rho100():
    g = P.packing_square(20, 0, rho=100.0)
    r = g.solve(30000, tol=1e-4, tol_feas=1e-4, check_every=50)
    print("solve rho=100:", r)
    assert r["converged"] == 1
    assert r["max_overlap"] <= 1e-4 and r["max_box"] <= 1e-4
    assert r["objective"] > 2.0
    g.sync_to_host()
    from oracle import pyoracle as po
    o = po.Graph.square(20, 0, 100.0)
    o.array("Z")[:] = g.array("Z")
    ov, bx, sr = o.square_quality(20)
    assert abs(ov - r["max_overlap"]) < 1e-6 and abs(sr - r["objective"]) < 1e-5
