"""GPU parity of the default single-instance path at the full sizes of the
1000- and 5000-circle unit-square workloads (rho = 1, alpha = 1, seed 0),
against the FP64 oracle, through the C ABI.

With one rho and alpha = 1 every single-instance solve runs in the contact
state of the dense layout (contact tiles of 32 at N = 1000, of 64 at
N = 5000).  These tests check it where it runs:

  * the first untethered iterations from the seeded start, against the
    oracle (<= 1e-5 relative) and against the per-edge path
    (PARADMM_CONTACT=0);
  * teacher-forced steps: the oracle's FP64 state uploaded, four device
    iterations (after-upload pass, contact entry, two contact iterations),
    every array compared with four oracle iterations;
  * for each size, the first iteration whose Z leaves the oracle by more
    than 1e-5 and the number of pairs whose no-overlap test on the next
    prox input differs from the oracle's (overlap-test flips), printed.

A pair the tile kernel failed to list would leave the contact state with
M = Z and X = N on that pair, which the teacher-forced X comparison catches.
"""
import numpy as np
import pytest

from gpu_helpers import OR_NAME, max_rel, square_pair, trajectory_error

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
    return set(g.profile_kernels(2))


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
    gc.sync_from_host()
    assert "contact_tile" in _kernels(gc)
    gc.close()
    o.close()


@pytest.mark.parametrize("nc,iters,starts", CASES, ids=[str(c[0]) for c in CASES])
def test_contact_full_size_teacher_forced(monkeypatch, nc, iters, starts):
    monkeypatch.delenv("PARADMM_LAYOUT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT_TILE", raising=False)
    g, o = square_pair(nc, 0)
    done = 0
    rows = []
    for k in starts:
        o.iterate_par(k - done, THREADS)
        done = k
        snap = {n: o.array(OR_NAME[n]).copy() for n in STATE}
        for n in STATE:
            g.array(n)[:] = snap[n]
        g.sync_from_host()
        g.iterate(4)
        g.sync_to_host()
        o.iterate_par(4, THREADS)
        errs = {n: max_rel(g.array(n), o.array(OR_NAME[n])) for n in STATE}
        fl, act = _flips(g, o, nc)
        rows.append((k, errs, fl, act))
        print(f"\nN={nc} from oracle iteration {k}: four-step errors "
              + " ".join(f"{n} {errs[n]:.1e}" for n in STATE) + f"; overlap-test flips {fl} of {act} active pairs")
        for n in STATE:
            o.array(OR_NAME[n])[:] = snap[n]
        if k == starts[-1]:
            break
    for k, errs, fl, act in rows:
        assert errs["Z"] <= 1e-5, k
        for n in ("X", "M", "N"):
            assert errs[n] <= 2e-5, (k, n)
        # U = M - Z cancels: its error is that of M and Z relative to the small U
        assert errs["U"] <= 1e-4, k
        assert fl <= max(2, act // 1000), k
    g.close()
    o.close()
