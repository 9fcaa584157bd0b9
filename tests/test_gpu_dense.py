"""GPU parity of the dense pair-tile layout (all-pairs unit-square packing)
against the FP64 oracle and against the generic layout, through the C ABI.

Sizes cover one diagonal tile (N=20), exact tiles (N=64, 128), ragged edge
tiles (N=70, 150) and the full N=5000 workload (size-independent invariants).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import OR_NAME, max_rel, randomize, square_pair, trajectory_error

pytestmark = pytest.mark.gpu

LEGACY = {
    "updateX": ("or_update_X",), "updateM": ("or_update_M",), "updateZ": ("or_update_Z",),
    "updateU": ("or_update_U",), "updateN": ("or_update_N",), "updateXM": ("or_update_XM",),
    "updateUN": ("or_update_UN",), "updateG": ("or_update_XM", "or_update_Z", "or_update_UN"),
}


@pytest.fixture(autouse=True)
def default_layout(monkeypatch):
    monkeypatch.delenv("PARADMM_LAYOUT", raising=False)


def _compare_all(o, g, tol, names=("X", "M", "N", "U", "Z")):
    for n in names:
        err = max_rel(g.array(n), o.array(OR_NAME[n]))
        assert err < tol, (n, err)


@pytest.mark.parametrize("nc", [20, 70, 130])
@pytest.mark.parametrize("call", sorted(LEGACY))
def test_dense_legacy_single_pass(nc, call):
    g, o = square_pair(nc, 3)
    randomize(o, g, 100 + nc, lo=0.0, hi=1.0)
    getattr(g.lib, call)(g.ptr)
    for fn in LEGACY[call]:
        getattr(o.lib, fn)(o.h)
    assert g.layout() == 2
    _compare_all(o, g, 2e-5)


@pytest.mark.parametrize("fixture", ["square20_seed0.npz", "square20_seed0_rho100.npz"])
def test_dense_trajectory_vs_reference(fixture):
    f = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    g, _ = square_pair(int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    hist = np.zeros_like(f["z"])
    for it in range(hist.shape[0]):
        g.iterate(1)
        g.sync_to_host()
        hist[it] = g.array("Z")
    assert g.layout() == 2
    err = trajectory_error(hist, f["z"])
    print(f"{fixture}: dense max rel Z error over 200 iterations {err.max():.3e}")
    assert err.max() < 1e-5


@pytest.mark.parametrize("nc,rho", [(64, 100.0), (150, 100.0), (150, 1.0)])
def test_dense_teacher_forced_steps(nc, rho):
    """From 40 states along the oracle's trajectory: upload the FP64 state,
    run two fused device iterations (the after-upload kernel, then the
    steady-state kernel) and compare with two oracle iterations.  Isolates
    per-step parity from the fp32/fp64 divergence of a long trajectory."""
    g, o = square_pair(nc, 7, rho=rho)
    errs = []
    for step in range(40):
        for n in ("X", "M", "N", "U", "Z"):
            g.array(n)[:] = o.array(OR_NAME[n])
        g.sync_from_host()
        g.iterate(2)
        g.sync_to_host()
        snap = {n: o.array(OR_NAME[n]).copy() for n in ("X", "M", "N", "U", "Z")}
        o.iterate(2)
        errs.append(max_rel(g.array("Z"), o.array("Z")))
        for n in ("X", "M", "N", "U", "Z"):  # rewind the oracle one iteration
            o.array(OR_NAME[n])[:] = snap[n]
        o.iterate(1)
    errs = np.array(errs)
    print(f"N={nc} rho={rho}: two-step Z error median {np.median(errs):.2e} max {errs.max():.2e}")
    assert g.layout() == 2
    assert np.median(errs) < 2e-6
    assert np.mean(errs < 1e-5) >= 0.95


@pytest.mark.parametrize("nc,rho", [(64, 100.0), (150, 1.0)])
def test_dense_trajectory_divergence_profile(nc, rho):
    """Untethered 200-iteration trajectory: reports where fp32 and FP64 part
    (a flipped overlap test moves one edge by the overlap, then the dynamics
    carry it).  The first iterations must agree to 1e-5."""
    g, o = square_pair(nc, 7, rho=rho)
    hist = o.iterate(200, record=True)
    out = np.zeros_like(hist)
    for it in range(200):
        g.iterate(1)
        g.sync_to_host()
        out[it] = g.array("Z")
    err = trajectory_error(out, hist)
    first = int(np.argmax(err > 1e-5)) if np.any(err > 1e-5) else -1
    print(f"N={nc} rho={rho}: Z error at it 1/10/50/100/200: "
          + " ".join(f"{err[k]:.1e}" for k in (0, 9, 49, 99, 199)) + f"; first > 1e-5 at iteration {first + 1 if first >= 0 else 'none'}")
    assert err[:5].max() < 1e-5


def test_dense_equals_generic_layout(monkeypatch):
    """The two layouts compute the same iteration (fp32 rounding apart)."""
    gd, _ = square_pair(97, 5, rho=10.0, alpha=0.7)
    gd.iterate(3)
    gd.sync_to_host()
    monkeypatch.setenv("PARADMM_LAYOUT", "generic")
    gg, _ = square_pair(97, 5, rho=10.0, alpha=0.7)
    gg.iterate(3)
    gg.sync_to_host()
    assert gd.layout() == 2 and gg.layout() == 1
    # X passes through the projection of near-touching pairs, which amplifies
    # fp32 rounding of N; the consensus and duals agree tighter
    for n, tol in (("X", 1e-3), ("U", 1e-3), ("Z", 1e-5), ("M", 1e-3), ("N", 1e-3)):
        assert max_rel(gd.array(n), gg.array(n)) < tol, n


def test_dense_graph_replay_matches_eager():
    g1, _ = square_pair(77, 21, rho=3.0)
    g2, _ = square_pair(77, 21, rho=3.0)
    g1.iterate(40)
    for _ in range(40):
        g2.iterate(1)
    g1.sync_to_host()
    g2.sync_to_host()
    for n in ("Z", "U", "X"):
        assert np.array_equal(g1.array(n), g2.array(n)), n


def test_dense_residuals_match_host_recomputation():
    g, o = square_pair(90, 2, rho=4.0, alpha=1.0)
    g.iterate(25)
    g.sync_to_host()
    Zp = g.array("Z").reshape(-1, 3).copy()
    g.iterate(1)
    p2, d2 = g.residuals()
    g.sync_to_host()
    ev = g.edge_var()
    X = g.array("X").reshape(-1, 3)
    Z = g.array("Z").reshape(-1, 3)
    rho = g.array("RHOS")
    assert p2 == pytest.approx(np.sum((X - Z[ev]) ** 2), rel=1e-5)
    assert d2 == pytest.approx(np.sum(rho ** 2 * np.sum((Z - Zp)[ev] ** 2, axis=1)), rel=1e-5)
    # residuals again after more iterations (the finish step leaves U current)
    g.iterate(3)
    p2b, _ = g.residuals()
    g.sync_to_host()
    X = g.array("X").reshape(-1, 3)
    Z = g.array("Z").reshape(-1, 3)
    assert p2b == pytest.approx(np.sum((X - Z[ev]) ** 2), rel=1e-5)


def test_dense_warm_start_vs_oracle():
    import ctypes as C
    g, o = square_pair(70, 6, rho=3.0)
    rng = np.random.default_rng(3)
    zinit = rng.uniform(0, 1, size=g.nvar * 3)
    uinit = rng.uniform(-0.1, 0.1, size=g.nedge * 3)
    DP = C.POINTER(C.c_double)
    g.lib.initialize_X_N_Z_M_U_optimal(g.ptr, zinit.ctypes.data_as(DP), uinit.ctypes.data_as(DP), 3)
    o.lib.or_warm_start(o.h, zinit.ctypes.data_as(DP), uinit.ctypes.data_as(DP), 3)
    assert g.layout() == 2
    _compare_all(o, g, 2e-5, names=("X", "M", "N", "U"))


def test_dense_updateGOpenMP_vs_oracle():
    g, o = square_pair(70, 8, rho=50.0)
    g.legacy("updateGOpenMP", 3, 4)
    o.iterate(3)
    assert g.layout() == 2
    _compare_all(o, g, 1e-4)


def test_dense_full_size_invariants():
    """N = 5000 (E = 25,005,000): Z is the rho-weighted mean of M = X + U_old,
    the residuals equal a host recomputation, and two runs agree bitwise."""
    n = 5000
    zs = []
    for run in range(2):
        from engine_amd import paradmm as P
        g = P.packing_square(n, 0, rho=1.0, alpha=1.0)
        assert g.nedge == 25005000
        g.iterate(12)
        g.sync_to_host()
        Zp = g.array("Z").reshape(-1, 3).copy()
        g.iterate(1)
        p2, d2 = g.residuals()
        g.sync_to_host()
        assert g.layout() == 2
        ev = g.edge_var()
        X = g.array("X").reshape(-1, 3)
        U = g.array("U").reshape(-1, 3)
        M = g.array("M").reshape(-1, 3)
        Z = g.array("Z").reshape(-1, 3)
        rho = g.array("RHOS")
        assert np.all(np.isfinite(Z))
        if run == 0:
            num = np.stack([np.bincount(ev, weights=rho * M[:, k], minlength=n) for k in range(3)], axis=1)
            den = np.bincount(ev, weights=rho, minlength=n)
            assert max_rel(Z, num / den[:, None]) < 1e-5
            assert np.allclose(M, X + U - (X - Z[ev]), atol=1e-6)
            assert p2 == pytest.approx(np.sum((X - Z[ev]) ** 2), rel=1e-4)
            assert d2 == pytest.approx(np.sum(rho ** 2 * np.sum((Z - Zp)[ev] ** 2, axis=1)), rel=1e-4)
        zs.append(Z.copy())
        g.close()
    assert np.array_equal(zs[0], zs[1])


@pytest.mark.parametrize("layout", ["dense", "generic"])
def test_coincident_centres_stay_finite(layout, monkeypatch):
    """Two circles on the same centre (box corners clamp circles there): the
    pair operator separates them along +x on the GPU as in the oracle."""
    if layout == "generic":
        monkeypatch.setenv("PARADMM_LAYOUT", "generic")
    g, o = square_pair(6, 2, rho=50.0)
    for a in (o.array("Z"), g.array("Z")):
        a[3:6] = a[0:3]  # circle 1 on circle 0
    ev = g.edge_var()
    for name in ("X", "N", "M"):
        for a, src in ((o.array(name), o.array("Z")), (g.array(name), g.array("Z"))):
            a[:] = src.reshape(-1, 3)[ev].ravel()
    g.sync_from_host() if g.layout() else None
    g.iterate(3)
    g.sync_to_host()
    o.iterate(3)
    assert np.all(np.isfinite(g.array("Z")))
    assert max_rel(g.array("Z"), o.array("Z")) < 1e-5


@pytest.mark.parametrize("layout", ["dense", "generic"])
def test_final_radii_vs_reference(layout, monkeypatch):
    """N=20 at rho=100 run to its fixed point (12000 iterations; the 1e-4
    tolerance is met at 8450): final radii against the reference engine
    compiled from source (golden fixture) within 1e-4; the error against the
    FP64 oracle is printed beside it."""
    if layout == "generic":
        monkeypatch.setenv("PARADMM_LAYOUT", "generic")
    f = np.load(os.path.join(GOLDEN, "square20_seed0_rho100_converged.npz"), allow_pickle=False)
    n, iters = int(f["n"]), int(f["iters_final"])
    g, o = square_pair(n, int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    g.iterate(iters)
    g.sync_to_host()
    o.iterate(iters)
    r = g.array("Z").reshape(-1, 3)[:, 2]
    err_ref = float(np.max(np.abs(r - f["z_final"].reshape(-1, 3)[:, 2])))
    err_or = float(np.max(np.abs(r - o.array("Z").reshape(-1, 3)[:, 2])))
    print(f"{layout}: final radii after {iters} iterations: max |dr| vs reference {err_ref:.2e}, vs FP64 oracle {err_or:.2e}")
    assert g.layout() == (2 if layout == "dense" else 1)
    assert err_ref < 1e-4 and err_or < 1e-4


def test_graph_recorded_again_when_rho_stops_being_uniform():
    """Uniform rho/alpha runs the kernel variant without the per-edge rho and
    alpha streams; after a per-edge change the captured iterations are
    recorded again with the general variant.  Replays equal eager runs."""
    from engine_amd import paradmm as P
    from oracle import mlp_oracle as mo
    m = P.RhoMLP(params=(mo.init_params(5) * 0.3).astype(np.float32))
    zs = []
    for eager in (False, True):
        g = P.packing_square(90, 4, rho=8.0)
        for _ in range(2):
            if eager:
                for _ in range(40):
                    g.iterate(1)
            else:
                g.iterate(40)
            g.apply_rho_policy(m, 0.2)
        g.sync_to_host()
        assert g.layout() == 2
        assert len(np.unique(g.array("RHOS"))) > 1
        zs.append(g.array("Z").copy())
    assert np.array_equal(zs[0], zs[1])


def test_uniform_rho_change_seen_by_captured_iterations():
    """A uniform rho changed after the iterations were captured (scale_rho
    keeps it uniform) is what the replays use: replays equal eager runs, and
    differ from a run without the change."""
    from engine_amd import paradmm as P
    zs = []
    for eager, scale in ((False, 2.0), (True, 2.0), (False, 1.0)):
        g = P.packing_square(90, 4, rho=8.0)
        for k in range(2):
            if eager:
                for _ in range(40):
                    g.iterate(1)
            else:
                g.iterate(40)
            if k == 0:
                g.scale_rho(scale)
        g.sync_to_host()
        assert g.layout() == 2
        zs.append(g.array("Z").copy())
    assert np.array_equal(zs[0], zs[1])
    assert not np.array_equal(zs[0], zs[2])


@pytest.mark.parametrize("layout", ["dense", "generic"])
def test_read_single_arrays_matches_full_sync(monkeypatch, layout):
    """paradmm_gpu_read returns one device array (Z, X, U, rho) equal to what
    a full sync to the host writes, without touching the host arrays."""
    from engine_amd import paradmm as P
    if layout == "generic":
        monkeypatch.setenv("PARADMM_LAYOUT", "generic")
    g = P.packing_square(60, 3, rho=20.0)
    g.iterate(17)
    z0 = g.array("Z").copy()
    reads = {k: g.read(k) for k in ("Z", "X", "U", "RHO")}
    assert np.array_equal(g.array("Z"), z0)  # host untouched
    g.sync_to_host()
    assert np.array_equal(reads["Z"].reshape(-1), g.array("Z"))
    assert np.array_equal(reads["X"].reshape(-1), g.array("X"))
    assert np.array_equal(reads["U"].reshape(-1), g.array("U"))
    assert np.array_equal(reads["RHO"], g.array("RHOS"))
    g.close()


def _per_edge_rho_alpha1(nc, seed):
    g, o = square_pair(nc, seed, rho=10.0)
    rng = np.random.default_rng(seed)
    r = rng.uniform(2.0, 30.0, size=o.nedge)
    o.array("RHO")[:] = r
    g.array("RHOS")[:] = r
    o.array("ALPHA")[:] = 1.0
    g.array("ALPHAS")[:] = 1.0
    return g, o


def test_m_state_runs_with_per_edge_rho_vs_oracle(monkeypatch):
    """alpha = 1 with a different rho on every edge: a captured run of 16
    iterations keeps M = X + U in the X array (the per-edge rho variant)
    and must follow the FP64 oracle; every step starts from the oracle's
    state (uploaded), so only the run itself is compared.  The states are
    taken after 400 iterations: earlier, a pair whose overlap test flips
    between fp32 and FP64 makes any fp32 run part from the oracle (the
    plain kernel as much as this one).  The per-edge path (the contact state
    takes per-edge rho too by default)."""
    monkeypatch.setenv("PARADMM_CONTACT", "0")
    g, o = _per_edge_rho_alpha1(80, 4)
    o.iterate(400)
    errs = []
    for _ in range(3):
        for n in ("X", "M", "N", "U", "Z"):
            g.array(n)[:] = o.array(OR_NAME[n])
        g.sync_from_host()
        g.iterate(16)
        g.sync_to_host()
        o.iterate(16)
        errs.append(max_rel(g.array("Z"), o.array("Z")))
    print("M-state runs with per-edge rho, Z error after 16 iterations:", " ".join(f"{e:.1e}" for e in errs))
    assert g.layout() == 2
    assert max(errs) < 1e-5


def test_m_state_runs_match_eager_with_per_edge_rho(monkeypatch):
    """Bit-for-bit: 32 iterations as two captured runs (the M state, per-edge
    rho) against 32 single iterations (the plain fused kernel), on the
    per-edge path."""
    monkeypatch.setenv("PARADMM_CONTACT", "0")
    g1, _ = _per_edge_rho_alpha1(70, 9)
    g2, _ = _per_edge_rho_alpha1(70, 9)
    g1.iterate(32)
    for _ in range(32):
        g2.iterate(1)
    g1.sync_to_host()
    g2.sync_to_host()
    for n in ("Z", "U", "X"):
        assert np.array_equal(g1.array(n), g2.array(n)), n
