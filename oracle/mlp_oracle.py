"""FP64 CPU oracle of the learned-rho MLP (TEST INFRASTRUCTURE ONLY).

The network has no reference counterpart (the reference's rho hook,
core/paradmm.c:564-566, is empty).  Architecture 6 -> 32 -> 32 -> 1 with ReLU,
loss 0.5 * mean((y - t)^2).  Two modes:

  exact=True   plain FP64 forward/backward: pinned by central finite
               differences in tests/test_mlp_oracle.py;
  exact=False  the same arithmetic with operands rounded to bf16
               (round-to-nearest-even) exactly where the GPU kernels feed
               MFMA: features, weights, hidden activations and the
               back-propagated pre-activation gradients; accumulation FP64.
               The GPU's fp32-accumulated results are compared to this.

Parameter vector order (1313 floats), shared with the GPU module:
  W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3[1].
"""
import numpy as np

NIN, NH = 6, 32
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round to bf16 (round-to-nearest-even on the fp32 value), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def forward(p, X, exact=True):
    P = unpack(p)
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    Xr = r(X)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    z1 = Xr @ W1.T + P["b1"]
    h1 = np.maximum(z1, 0.0)
    h1r = r(h1)
    z2 = h1r @ W2.T + P["b2"]
    h2 = np.maximum(z2, 0.0)
    h2r = r(h2)
    y = h2r @ W3 + P["b3"][0]
    return y, (Xr, W1, W2, W3, z1, h1r, z2, h2r)


def loss_and_grad(p, X, t, exact=True):
    """0.5 * mean((y - t)^2) and its gradient w.r.t. the parameter vector."""
    t = np.asarray(t, dtype=np.float64)
    B = X.shape[0]
    y, (Xr, W1, W2, W3, z1, h1r, z2, h2r) = forward(p, X, exact)
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    dy = (y - t) / B
    g = {}
    g["W3"] = dy @ h2r
    g["b3"] = np.array([dy.sum()])
    dz2 = np.outer(dy, W3) * (z2 > 0)
    g["b2"] = dz2.sum(0)
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ h1r
    dh1 = dz2r @ W2
    dz1 = dh1 * (z1 > 0)
    g["b1"] = dz1.sum(0)
    dz1r = r(dz1)
    g["W1"] = dz1r.T @ Xr
    loss = 0.5 * np.mean((y - t) ** 2)
    return loss, pack(g)


def init_params(seed):
    """He-normal weights, zero biases (the GPU module draws its own; this is for tests)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / NIN), (NH, NIN)), "b1": np.zeros(NH),
         "W2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
         "W3": rng.normal(0, np.sqrt(1.0 / NH), (NH,)), "b3": np.zeros(1)}
    return pack(d)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
