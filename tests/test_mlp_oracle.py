"""Pins the FP64 MLP oracle of the learned-rho policy: gradients against
central finite differences, and the bf16 emulation against the exact mode."""
import numpy as np

from oracle import mlp_oracle as mo


def test_param_count():
    assert mo.NPARAMS == 32 * 6 + 32 + 32 * 32 + 32 + 32 + 1 == 1313


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mo.init_params(1)
    p += rng.normal(0, 0.05, p.shape)  # non-zero biases
    X = rng.normal(0, 1, (37, 6))
    t = rng.normal(0, 1, 37)
    _, g = mo.loss_and_grad(p, X, t, exact=True)
    h = 1e-6
    idx = rng.choice(mo.NPARAMS, 120, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mo.loss_and_grad(p + e, X, t)
        lm, _ = mo.loss_and_grad(p - e, X, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265])
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # tie -> even
    assert r[2] == 1.0 + 2 ** -7       # representable
    assert r[3] == 1.0 + 2 ** -6       # tie -> even (up)
    assert r[4] == -2.5
    assert abs(r[5] - 3.140625) < 1e-12


def test_bf16_mode_close_to_exact():
    rng = np.random.default_rng(3)
    p = mo.init_params(4)
    X = rng.normal(0, 1, (256, 6))
    t = rng.normal(0, 1, 256)
    y0, _ = mo.forward(p, X, exact=True)
    y1, _ = mo.forward(p, X, exact=False)
    assert np.max(np.abs(y0 - y1)) < 0.05 * np.max(np.abs(y0))
    _, g0 = mo.loss_and_grad(p, X, t, exact=True)
    _, g1 = mo.loss_and_grad(p, X, t, exact=False)
    assert np.linalg.norm(g0 - g1) < 0.05 * np.linalg.norm(g0)


def test_adam_decreases_loss():
    rng = np.random.default_rng(5)
    p = mo.init_params(6)
    X = rng.normal(0, 1, (512, 6))
    t = np.tanh(X[:, 0] - 0.5 * X[:, 2])
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0, _ = mo.loss_and_grad(p, X, t)
    for s in range(1, 201):
        _, g = mo.loss_and_grad(p, X, t)
        p, m, v = mo.adam_step(p, g, m, v, s, 1e-2)
    l1, _ = mo.loss_and_grad(p, X, t)
    assert l1 < 0.2 * l0
