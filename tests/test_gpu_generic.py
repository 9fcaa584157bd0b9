"""GPU parity of the generic (CSR) layout against the FP64 oracle, through
the C ABI.  Tolerances are fp32 vs fp64: single passes 2e-5 of the array's
scale (a handful of fp32 roundings on inputs in [-1, 2]); Z trajectories
1e-5 relative where the dynamics allow it."""
import numpy as np
import pytest

from gpu_helpers import (OR_NAME, copy_state, max_rel, randomize, square_pair, trajectory_error,
                         triangle_lib)

pytestmark = pytest.mark.gpu

PASS_TOL = 2e-5

LEGACY = {
    "updateX": ("or_update_X",), "updateM": ("or_update_M",), "updateZ": ("or_update_Z",),
    "updateU": ("or_update_U",), "updateN": ("or_update_N",), "updateXM": ("or_update_XM",),
    "updateUN": ("or_update_UN",), "updateG": ("or_update_XM", "or_update_Z", "or_update_UN"),
    "updateXOpenMP": ("or_update_X",), "updateMOpenMP": ("or_update_M",), "updateZOpenMP": ("or_update_Z",),
    "updateUOpenMP": ("or_update_U",), "updateNOpenMP": ("or_update_N",),
}


@pytest.fixture(autouse=True)
def generic_layout(monkeypatch):
    monkeypatch.setenv("PARADMM_LAYOUT", "generic")


def _compare_all(o, g, tol):
    for n in ("X", "M", "N", "U", "Z"):
        err = max_rel(g.array(n), o.array(OR_NAME[n]))
        assert err < tol, (n, err)


@pytest.mark.parametrize("problem", ["square", "triangle"])
@pytest.mark.parametrize("call", sorted(LEGACY))
def test_legacy_single_pass(problem, call):
    if problem == "square":
        g, o = square_pair(9, 3)
        randomize(o, g, 11, lo=0.0, hi=1.0)
    else:
        g, o = triangle_lib(6, 3)
        randomize(o, g, 12)
    getattr(g.lib, call)(g.ptr)
    for fn in LEGACY[call]:
        getattr(o.lib, fn)(o.h)
    assert g.layout() == 1
    _compare_all(o, g, PASS_TOL)


def test_legacy_respects_host_writes_between_calls():
    g, o = square_pair(8, 5, rho=2.0)
    for it in range(3):
        g.legacy("updateXM"); o.lib.or_update_XM(o.h)
        # the caller changes rho and U between passes, as host.c writes G.RHOS
        g.array("RHOS")[::3] = 4.0 + it
        o.array("RHO")[::3] = 4.0 + it
        g.array("U")[::5] += 0.01
        o.array("U")[::5] += 0.01
        g.legacy("updateZ"); o.lib.or_update_Z(o.h)
        g.legacy("updateUN"); o.lib.or_update_UN(o.h)
    _compare_all(o, g, 1e-5)


@pytest.mark.parametrize("fixture,iters_tol", [("square20_seed0.npz", 1e-5), ("square20_seed0_rho100.npz", 1e-5)])
def test_legacy_loop_trajectory_vs_reference(fixture, iters_tol):
    """host.c's serial loop (updateXM, updateZ, updateUN) against the reference's Z trajectory."""
    import os
    from conftest import GOLDEN
    f = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    g, _ = square_pair(int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    hist = np.zeros_like(f["z"])
    for it in range(hist.shape[0]):
        g.legacy("updateXM")
        g.legacy("updateZ")
        g.legacy("updateUN")
        hist[it] = g.array("Z")
    err = trajectory_error(hist, f["z"])
    print(f"{fixture}: legacy loop max rel Z error over 200 iterations {err.max():.3e}")
    assert err.max() < iters_tol


@pytest.mark.parametrize("fixture", ["square20_seed0.npz", "square20_seed0_rho100.npz"])
def test_fused_iterate_trajectory_vs_reference(fixture):
    import os
    from conftest import GOLDEN
    f = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    g, _ = square_pair(int(f["n"]), int(f["seed"]), float(f["rho"]), float(f["alpha"]), float(f["a"]))
    hist = np.zeros_like(f["z"])
    for it in range(hist.shape[0]):
        g.iterate(1)
        g.sync_to_host()
        hist[it] = g.array("Z")
    err = trajectory_error(hist, f["z"])
    print(f"{fixture}: fused max rel Z error over 200 iterations {err.max():.3e}")
    assert err.max() < 1e-5


def test_fused_iterate_full_state_vs_oracle():
    """After k device iterations every host array equals the oracle's (M and N included)."""
    g, o = square_pair(12, 9, rho=5.0, alpha=0.5)
    g.iterate(5)
    o.iterate(5)
    g.sync_to_host()
    _compare_all(o, g, 2e-5)
    # N and M read back are those of the reference state after an iteration
    ev = g.edge_var()
    Z = g.array("Z").reshape(-1, 3)
    N, U, X, M = (g.array(n).reshape(-1, 3) for n in ("N", "U", "X", "M"))
    assert np.allclose(N, Z[ev] - U, rtol=0, atol=1e-6)
    assert np.allclose(M, X + U - 0.5 * (X - Z[ev]), rtol=0, atol=1e-6)


def test_triangle_fused_vs_oracle():
    g, o = triangle_lib(7, 4)
    g.legacy("updateGOpenMP", 50, 8)
    hist = o.iterate(50, record=True)
    err = max_rel(g.array("Z"), hist[-1])
    assert err < 1e-4, err
    _compare_all(o, g, 1e-4)


def test_graph_replay_matches_eager():
    """Iterations through the captured hipGraph equal one-by-one iterations bitwise."""
    g1, _ = square_pair(15, 21, rho=3.0)
    g2, _ = square_pair(15, 21, rho=3.0)
    g1.iterate(40)          # 1 eager + 2 graph launches of 16 + 7 eager
    for _ in range(40):
        g2.iterate(1)       # all eager
    g1.sync_to_host()
    g2.sync_to_host()
    assert np.array_equal(g1.array("Z"), g2.array("Z"))
    assert np.array_equal(g1.array("U"), g2.array("U"))


def test_residuals_match_host_recomputation():
    g, o = square_pair(14, 2, rho=4.0, alpha=1.0)
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
    p_ref = np.sum((X - Z[ev]) ** 2)
    d_ref = np.sum(rho ** 2 * np.sum((Z - Zp)[ev] ** 2, axis=1))
    assert p2 == pytest.approx(p_ref, rel=1e-5)
    assert d2 == pytest.approx(d_ref, rel=1e-5)
    # and the oracle, 26 iterations in FP64
    o.iterate(26)
    po2, od2 = o.residuals()
    assert p2 == pytest.approx(po2, rel=1e-2)
    assert d2 == pytest.approx(od2, rel=1e-2)


@pytest.mark.parametrize("maxiter", [0, 1, 4])
def test_warm_start_vs_oracle(maxiter):
    import ctypes as C
    g, o = square_pair(7, 6, rho=3.0)
    rng = np.random.default_rng(maxiter)
    zinit = rng.uniform(0, 1, size=g.nvar * 3)
    uinit = rng.uniform(-0.1, 0.1, size=g.nedge * 3)
    DP = C.POINTER(C.c_double)
    g.lib.initialize_X_N_Z_M_U_optimal(g.ptr, zinit.ctypes.data_as(DP), uinit.ctypes.data_as(DP), maxiter)
    o.lib.or_warm_start(o.h, zinit.ctypes.data_as(DP), uinit.ctypes.data_as(DP), maxiter)
    _compare_all(o, g, 2e-5)


def test_deterministic():
    out = []
    for _ in range(2):
        g, _o = square_pair(25, 8, rho=2.0)
        g.iterate(30)
        g.sync_to_host()
        out.append(g.array("Z").copy())
        g.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("nc", [7, 70])
def test_triangle_pair_rows_equal_per_factor_pass(monkeypatch, nc):
    """The reference demo's triangle packing has its pair edges in rows of
    the generic layout; its pair X pass then runs as coalesced tiles
    (k_xpass_pair2d_rows, 32 x 32 circle pairs, the ragged last block and
    the diagonal tiles included at N = 70).  Same prox on the same inputs:
    30 fused iterations and a legacy X pass agree bit for bit with the
    per-factor pass (PARADMM_PAIR_ROWS=0)."""
    runs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("PARADMM_PAIR_ROWS", mode)
        g, o = triangle_lib(nc, 5)
        g.iterate(30)
        kernels = g.profile_kernels(1)
        g.sync_to_host()
        g.legacy("updateXM")
        g.sync_to_host()
        runs[mode] = {n: g.array(n).copy() for n in ("Z", "X", "U", "M")}
        print(f"N={nc}, rows {mode}: kernels {sorted(kernels)}")
        assert ("disk_collision_2d_rows" in kernels) == (mode == "1")
        o.close()
        g.close()
    for n in ("Z", "X", "U", "M"):
        assert np.array_equal(runs["1"][n], runs["0"][n]), n


def test_triangle_pair_rows_vs_oracle():
    """The tiled pair pass against the FP64 oracle on the triangle demo at
    N = 100 (four circle blocks): 5 iterations within 1e-5."""
    g, o = triangle_lib(100, 2)
    for it in range(5):
        g.iterate(1)
        o.iterate(1)
        err = max_rel(g.read("Z").ravel(), o.array("Z"))
        assert err < 1e-5, (it, err)
    assert "disk_collision_2d_rows" in g.profile_kernels(1)


@pytest.mark.parametrize("problem", ["triangle", "square"])
def test_zpass_workgroup_form_matches_wavefront_form(monkeypatch, problem):
    """The fused Z pass of the generic layout with a workgroup per variable
    and its edges in registers (default where every variable fits) sums the
    same terms in another order than the wavefront-per-variable pass
    (PARADMM_ZPASS_WG=0): 3 iterations agree to float rounding (1e-6
    relative), and both stay within 1e-5 of the FP64 oracle."""
    zs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("PARADMM_ZPASS_WG", mode)
        monkeypatch.setenv("PARADMM_LAYOUT", "generic")
        if problem == "triangle":
            g, o = triangle_lib(70, 5)
        else:
            g, o = square_pair(60, 3, rho=2.0)
        g.iterate(3)
        zs[mode] = g.read("Z").ravel()
        o.iterate(3)
        assert max_rel(zs[mode], o.array("Z")) < 1e-5, mode
        o.close()
        g.close()
    err = max_rel(zs["1"], zs["0"])
    print(f"{problem}: workgroup vs wavefront Z pass after 3 iterations: {err:.1e}")
    assert err < 1e-6
