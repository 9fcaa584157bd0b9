"""GPU parity of the default single-instance path at the full sizes of the
1000- and 5000-circle unit-square workloads (rho = 1, alpha = 1, seed 0),
against the FP64 oracle, through the C ABI.

With one rho and alpha = 1 every single-instance solve runs in the contact
state of the dense layout (the id-order form with contact tiles of 32 at
N = 1000, the spatial form with tiles of 16 at N = 5000).  These tests
check it where it runs:

  * the first untethered iterations from the seeded start, against the
    oracle (<= 1e-5 relative) and against the per-edge path
    (PARADMM_CONTACT=0);
  * teacher-forced steps: the oracle's FP64 state uploaded, four device
    iterations (after-upload pass, contact entry, two contact iterations),
    every array compared with four oracle iterations, and with the per-edge
    path (PARADMM_CONTACT=0) from the same upload;
  * for each size, the first iteration whose Z leaves the oracle by more
    than 1e-5 and the number of pairs whose no-overlap test on the next
    prox input differs from the oracle's (overlap-test flips), printed.

A pair the tile kernel failed to list would leave the contact state with
M = Z and X = N on that pair: its overlap test would differ from the
oracle's (a flip) and its X would be off by the whole projection step.

At rho = 1 the early iterations amplify any perturbation about eightfold
per iteration on the edges of deeply overlapping pairs, in FP64 as well:
the oracle started from the same state rounded to fp32 (the least change
any fp32 engine starts from) ends four iterations later up to 1e-3 away on
a few edges.  So per-edge arrays are judged against that measured
sensitivity (max error <= 30 x it + 1e-5), on their 99.9th percentile
(<= 1e-5), and against the per-edge path (<= 3 x its error); Z, an average
over N + 1 edges, must stay within 1e-5 outright.
"""
import numpy as np
import pytest

from gpu_helpers import OR_NAME, max_rel, square_pair, trajectory_error
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

STATE = ("X", "M", "N", "U", "Z")
THREADS = 16
# (circles, untethered iterations, teacher-forced starts)
CASES = [(1000, 100, (0, 3, 6, 30, 60)), (5000, 20, (0, 3, 9))]


def _pair_tests(nvec, nc):
    """The pair prox's no-overlap test |c1 - c2|^2 < (r1 + r2)^2 on the prox
    input N of every pair (edges 2 nc + 2p and 2 nc + 2p + 1), in FP64."""
    n = np.asarray(nvec).reshape(-1, 3)[2 * nc:]
    a, b = n[0::2], n[1::2]
    d2 = (a[:, 0] - b[:, 0]) ** 2 + (a[:, 1] - b[:, 1]) ** 2
    s = a[:, 2] + b[:, 2]
    return d2 < s * s


def _flips(g, o, nc):
    """(pairs whose test differs, pairs active in the oracle) on the next
    prox input: the device state synchronised to the host against the
    oracle's."""
    g.sync_to_host()
    tg = _pair_tests(g.array("N"), nc)
    to = _pair_tests(o.array("N"), nc)
    return int(np.count_nonzero(tg != to)), int(np.count_nonzero(to))


def _kernels(g):
    """Kernel names of two eager iterations; the tile kernel of the contact
    state's spatial form (the default from N = 3000) counts as the contact
    tile kernel."""
    k = set(g.profile_kernels(2))
    if any(n.startswith("scs_rows") for n in k):
        k.add("contact_tile")
    return k


@pytest.mark.parametrize("nc,iters,starts", CASES, ids=[str(c[0]) for c in CASES])
def test_contact_full_size_untethered(monkeypatch, nc, iters, starts):
    monkeypatch.delenv("PARADMM_LAYOUT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT_TILE", raising=False)
    monkeypatch.delenv("PARADMM_DENSE_TILE", raising=False)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("PARADMM_CONTACT", mode)
        g, _ = square_pair(nc, 0)
        h = np.zeros((iters, 3 * nc))
        for it in range(iters):
            g.iterate(1)
            h[it] = g.read("Z").ravel()
        assert g.layout() == 2
        out[mode] = h
        if mode == "1":
            gc = g
        else:
            g.close()
    _, o = square_pair(nc, 0)
    hist = o.iterate_par(iters, THREADS, record=True)
    ec = trajectory_error(out["1"], hist)
    ep = trajectory_error(out["0"], hist)
    cp = trajectory_error(out["1"], out["0"])
    first = lambda e: int(np.argmax(e > 1e-5)) + 1 if np.any(e > 1e-5) else None  # noqa: E731
    fl, act = _flips(gc, o, nc)
    print(f"\nN={nc} untethered: Z error vs the FP64 oracle, first 5 iterations max contact {ec[:5].max():.2e} "
          f"per-edge {ep[:5].max():.2e}; contact vs per-edge {cp[:5].max():.2e}; first iteration above 1e-5: "
          f"contact {first(ec) or f'>{iters}'} per-edge {first(ep) or f'>{iters}'}; after {iters} iterations "
          f"error {ec[-1]:.2e}, overlap-test flips {fl} of {act} active pairs")
    assert ec[:5].max() <= 1e-5
    assert ep[:5].max() <= 1e-5
    assert cp[:5].max() <= 1e-5
    # the tile kernel ran (not the per-edge path)
    monkeypatch.setenv("PARADMM_CONTACT", "1")
    gc.sync_from_host()
    assert "contact_tile" in _kernels(gc)
    gc.close()
    o.close()


def _abs_err(a, b):
    d = np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))
    return float(d.max()), float(np.quantile(d, 0.999))


@pytest.mark.parametrize("nc,iters,starts", CASES, ids=[str(c[0]) for c in CASES])
def test_contact_full_size_teacher_forced(monkeypatch, nc, iters, starts):
    monkeypatch.delenv("PARADMM_LAYOUT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT_TILE", raising=False)
    steps = 4
    paths = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("PARADMM_CONTACT", mode)
        paths[mode], _ = square_pair(nc, 0)
    _, o = square_pair(nc, 0)
    done = 0
    rows = []
    for k in starts:
        o.iterate_par(k - done, THREADS)
        done = k
        snap = {n: o.array(OR_NAME[n]).copy() for n in STATE}
        zprev = o.array("Zprev").copy()
        err = {}
        for mode, g in paths.items():
            monkeypatch.setenv("PARADMM_CONTACT", mode)
            for n in STATE:
                g.array(n)[:] = snap[n]
            g.sync_from_host()
            g.iterate(steps)
            g.sync_to_host()
        o.iterate_par(steps, THREADS)
        # the oracle from the same state rounded to fp32: the sensitivity
        b = po.Graph.square(nc, 0)
        for n in STATE:
            b.array(OR_NAME[n])[:] = snap[n].astype(np.float32).astype(np.float64)
        b.array("Zprev")[:] = zprev
        b.iterate_par(steps, THREADS)
        sens = {n: _abs_err(b.array(OR_NAME[n]), o.array(OR_NAME[n]))[0] for n in STATE}
        b.close()
        for mode, g in paths.items():
            err[mode] = {n: _abs_err(g.array(n), o.array(OR_NAME[n])) for n in STATE}
        zrel = max_rel(paths["1"].array("Z"), o.array("Z"))
        fl, act = _flips(paths["1"], o, nc)
        rows.append((k, err, sens, zrel, fl, act))
        print(f"\nN={nc} from oracle iteration {k}, {steps} steps: max / 99.9% abs error contact | per-edge | "
              "FP64 sensitivity: " + "; ".join(
                  f"{n} {err['1'][n][0]:.1e}/{err['1'][n][1]:.1e} | {err['0'][n][0]:.1e} | {sens[n]:.1e}"
                  for n in STATE) + f"; Z rel {zrel:.1e}; overlap-test flips {fl} of {act} active pairs")
        for n in STATE:
            o.array(OR_NAME[n])[:] = snap[n]
        o.array("Zprev")[:] = zprev
    for k, err, sens, zrel, fl, act in rows:
        assert zrel <= 1e-5, k
        for n in STATE:
            mx, q = err["1"][n]
            assert q <= 1e-5, (k, n, q)
            assert mx <= 30.0 * sens[n] + 1e-5, (k, n, mx, sens[n])
            assert mx <= 3.0 * err["0"][n][0] + 1e-6, (k, n, mx, err["0"][n][0])
        assert fl <= max(2, act // 1000), k
    for g in paths.values():
        g.close()
    o.close()
