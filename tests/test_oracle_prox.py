"""Pins the oracle's operators and definitions that the reference does not
cover, by closed forms, brute force and invariants chosen so that a plausible
slip (dropped term, wrong sign or index, swapped weight) fails one of them.
"""
import itertools

import numpy as np
import pytest


def prox(oracle, name, n, rhos, nvars, ndims, params=None):
    lib = oracle.load()
    x = np.zeros_like(n)
    oracle.call_prox(getattr(lib, name), x, n.copy(), np.asarray(rhos, dtype=np.float64).copy(), nvars, ndims,
                     None if params is None else np.asarray(params, dtype=np.float64))
    return x


# ------------------------------------------------------------------ box (dim 3)

BOX_A = np.array([[-1, 0, 1], [1, 0, 1], [0, -1, 1], [0, 1, 1], [0, 0, -1]], dtype=np.float64)  # A y <= b
BOX_B = np.array([0, 1, 0, 1, 0], dtype=np.float64)


def box_bruteforce(n):
    """Projection of n onto {A y <= b} by enumerating active sets: among the
    subsets whose equality-constrained projection is feasible, the nearest."""
    best, bestd = None, np.inf
    for k in range(0, 4):
        for S in itertools.combinations(range(5), k):
            if k == 0:
                y = n.copy()
            else:
                A = BOX_A[list(S)]
                if np.linalg.matrix_rank(A) < k:
                    continue
                lam = np.linalg.solve(A @ A.T, A @ n - BOX_B[list(S)])
                y = n - A.T @ lam
            if np.all(BOX_A @ y <= BOX_B + 1e-12):
                d = np.sum((y - n) ** 2)
                if d < bestd:
                    best, bestd = y, d
    return best


def test_box_matches_bruteforce(oracle):
    rng = np.random.default_rng(7)
    pts = np.concatenate([
        rng.uniform(-0.5, 1.5, size=(300, 3)),
        np.column_stack([rng.uniform(0, 1, 200), rng.uniform(0, 1, 200), rng.uniform(-0.2, 0.8, 200)]),
        np.column_stack([rng.uniform(0.3, 0.7, 100), rng.uniform(-1, 2, 100), rng.uniform(0.3, 1.5, 100)]),
    ])
    for n in pts:
        x = prox(oracle, "or_prox_box3", n, [1.0], 1, 3)
        y = box_bruteforce(n)
        assert np.allclose(x, y, atol=1e-12), (n, x, y)


def test_box_rho_independent_and_feasible_identity(oracle):
    n = np.array([0.3, 0.6, 0.1])
    assert np.array_equal(prox(oracle, "or_prox_box3", n, [1.0], 1, 3), n)  # feasible point is fixed
    m = np.array([1.2, -0.3, 0.4])
    assert np.array_equal(prox(oracle, "or_prox_box3", m, [1.0], 1, 3), prox(oracle, "or_prox_box3", m, [37.0], 1, 3))


# ------------------------------------------------------------------ radius

@pytest.mark.parametrize("rho", [0.5, 1.0, 7.0])
def test_radius_a1_closed_form(oracle, rho):
    """a = 1: argmin -r + rho/2 (r-n)^2 = n + 1/rho, clamped at 0."""
    lib = oracle.load()
    for n in [-3.0, -0.5, 0.01, 0.3, 2.0]:
        r = lib.or_max_alpha_radius(n, rho, 1.0)
        assert r == pytest.approx(max(0.0, n + 1.0 / rho), abs=1e-14)


def test_radius_a2_closed_form(oracle):
    """a = 2: stationary point r = rho n / (rho - 2) (rho > 2, n > 0)."""
    lib = oracle.load()
    for rho in [2.5, 5.0, 10.0]:
        for n in [0.01, 0.2, 1.5]:
            assert lib.or_max_alpha_radius(n, rho, 2.0) == pytest.approx(rho * n / (rho - 2.0), rel=1e-13)


def test_radius3_only_changes_r(oracle):
    n = np.array([0.25, 0.75, 0.1])
    x = prox(oracle, "or_prox_radius3", n, [4.0], 1, 3, [1.0])
    assert x[0] == n[0] and x[1] == n[1] and x[2] == pytest.approx(0.1 + 0.25)


# ------------------------------------------------------------------ pair

def test_pair3_touching_invariant_and_weights(oracle):
    rng = np.random.default_rng(3)
    for _ in range(300):
        c1 = rng.uniform(0, 1, 2)
        c2 = c1 + rng.uniform(-0.2, 0.2, 2)
        r1, r2 = rng.uniform(0.05, 0.3, 2)
        rho = rng.uniform(0.5, 50, 2)
        n = np.array([c1[0], c1[1], r1, c2[0], c2[1], r2])
        x = prox(oracle, "or_prox_pair3", n, rho, 2, 3)
        dist = np.linalg.norm(c1 - c2)
        if dist >= r1 + r2:
            assert np.array_equal(x, n)
            continue
        # after projection the circles touch
        assert np.linalg.norm(x[0:2] - x[3:5]) == pytest.approx(x[2] + x[5], abs=1e-12)
        # centres move along the centre line, away from each other
        u = (c2 - c1) / dist
        d1, d2 = x[0:2] - c1, x[3:5] - c2
        assert abs(d1[0] * u[1] - d1[1] * u[0]) < 1e-12 and np.dot(d1, u) <= 0
        assert abs(d2[0] * u[1] - d2[1] * u[0]) < 1e-12 and np.dot(d2, u) >= 0
        # rho-weighted: moves scale with 1/rho, centre and radius of one circle alike
        m1, m2 = np.linalg.norm(d1), np.linalg.norm(d2)
        assert m1 * rho[0] == pytest.approx(m2 * rho[1], rel=1e-9)
        assert (r1 - x[2]) == pytest.approx(m1, rel=1e-9)
        assert (r2 - x[5]) == pytest.approx(m2, rel=1e-9)


def test_pair3_is_reference_pair2_with_shared_radius(oracle):
    """dim-3 pair == the reference's dim-2 pair with w_x = w_r = rho per circle."""
    rng = np.random.default_rng(9)
    for _ in range(200):
        c1 = rng.uniform(0, 1, 2)
        c2 = c1 + rng.uniform(-0.3, 0.3, 2)
        r1, r2 = rng.uniform(-0.1, 0.3, 2)
        rho = rng.uniform(0.5, 50, 2)
        n3 = np.array([c1[0], c1[1], r1, c2[0], c2[1], r2])
        x3 = prox(oracle, "or_prox_pair3", n3, rho, 2, 3)
        n2 = np.array([c1[0], c1[1], r1, r1, c2[0], c2[1], r2, r2])
        x2 = prox(oracle, "or_prox_pair2", n2, [rho[0], rho[0], rho[1], rho[1]], 4, 2)
        assert np.allclose(x3, [x2[0], x2[1], x2[2], x2[4], x2[5], x2[6]], rtol=1e-13, atol=1e-13)


def test_pair_equal_rho_symmetric(oracle):
    n = np.array([0.4, 0.5, 0.2, 0.6, 0.5, 0.2])
    x = prox(oracle, "or_prox_pair3", n, [3.0, 3.0], 2, 3)
    assert x[2] == pytest.approx(x[5]) and (x[0] + x[3]) / 2 == pytest.approx(0.5)
    # overlap 0.2, four unit weights: each coordinate moves by 0.05
    assert x[0] == pytest.approx(0.35) and x[3] == pytest.approx(0.65) and x[2] == pytest.approx(0.15)


def test_wall2_active_lands_on_plane(oracle):
    rng = np.random.default_rng(5)
    s = np.sin(np.pi / 3)
    for _ in range(200):
        v = np.array([-s, 0.5])
        p = np.array([0.0, 0.0])
        n = rng.uniform(-1, 2, 4)
        rho = rng.uniform(0.5, 20, 2)
        x = prox(oracle, "or_prox_wall2", n, rho, 2, 2, np.concatenate([v, p]))
        inner = v @ (n[0:2] - p)
        if inner + n[2] > 0:
            assert x[2] == x[3] >= 0
            if x[2] > 0:
                assert v @ (x[0:2] - p) + x[2] == pytest.approx(0.0, abs=1e-12)
        else:
            assert np.array_equal(x, n)


# ------------------------------------------------------------------ one iteration by hand

def test_one_iteration_two_circles_by_hand(oracle):
    """A 2-circle square packing, one X/M/Z/U/N iteration computed directly."""
    g = oracle.Graph.square(2, 11, 2.0, 0.5, 1.0)
    Z0 = g.array("Z").reshape(2, 3).copy()
    ev = g.edge_var()
    rho = g.array("RHO").copy()
    g.iterate(1)
    # factor order: radius0, radius1, box0, box1, pair(0,1): edges 0..5
    assert list(ev) == [0, 1, 0, 1, 0, 1]
    N = Z0[ev]  # N = Z broadcast, U = 0
    X = np.zeros((6, 3))
    for e in range(2):  # radius: r = max(0, n + 1/rho)
        X[e] = [N[e, 0], N[e, 1], max(0.0, N[e, 2] + 1.0 / rho[e])]
    for e in range(2, 4):
        X[e] = prox(oracle, "or_prox_box3", N[e].copy(), [rho[e]], 1, 3)
    X[4:6] = prox(oracle, "or_prox_pair3", N[4:6].ravel().copy(), rho[4:6], 2, 3).reshape(2, 3)
    M = X  # U = 0
    Z = np.zeros((2, 3))
    for v in range(2):
        es = np.where(ev == v)[0]
        Z[v] = (rho[es, None] * M[es]).sum(0) / rho[es].sum()
    U = 0.5 * (X - Z[ev])
    assert np.allclose(g.array("Z").reshape(2, 3), Z, atol=1e-15)
    assert np.allclose(g.array("U").reshape(6, 3), U, atol=1e-15)
    assert np.allclose(g.array("N").reshape(6, 3), Z[ev] - U, atol=1e-15)


def test_z_is_permutation_invariant_weighted_mean(oracle):
    g = oracle.Graph.square(5, 2, 1.0, 1.0, 1.0)
    g.array("RHO")[:] = np.random.default_rng(1).uniform(0.5, 3.0, g.nedge)
    g.iterate(3)
    X, U, rho, ev = g.array("X").reshape(-1, 3), g.array("U").reshape(-1, 3), g.array("RHO"), g.edge_var()
    M = g.array("M").reshape(-1, 3)
    Z = g.array("Z").reshape(-1, 3)
    for v in range(g.nvar):
        es = np.where(ev == v)[0][::-1]  # reversed order
        assert np.allclose(Z[v], (rho[es, None] * M[es]).sum(0) / rho[es].sum(), atol=1e-14)
    assert X.shape == U.shape


# ------------------------------------------------------------------ residuals

def test_residuals_match_dual_update(oracle):
    """primal^2 = sum ||X - Z||^2 equals sum ||U_k - U_{k-1}||^2 / alpha^2, and
    dual^2 = sum rho^2 ||Z_k - Z_{k-1}||^2; both independent of the oracle's own sums."""
    alpha = 0.5
    g = oracle.Graph.square(6, 4, 3.0, alpha, 1.0)
    g.iterate(5)
    U0 = g.array("U").copy()
    Z0 = g.array("Z").reshape(-1, 3).copy()
    g.iterate(1)
    p2, d2 = g.residuals()
    dU = g.array("U") - U0
    assert p2 == pytest.approx(np.sum(dU ** 2) / alpha ** 2, rel=1e-10)
    dZ = g.array("Z").reshape(-1, 3) - Z0
    ev = g.edge_var()
    assert d2 == pytest.approx(np.sum(g.array("RHO") ** 2 * np.sum(dZ[ev] ** 2, axis=1)), rel=1e-12)


def test_splitmix64_known_values(oracle):
    """splitmix64 from seed 0: first outputs of the published generator."""
    import ctypes as C
    lib = oracle.load()
    s = C.c_uint64(0)
    out = [lib.or_splitmix64(C.byref(s)) for _ in range(3)]
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_pair_coincident_centres_separate_along_x(oracle):
    """Coincident centres (the reference divides 0/0, packing/host.c:123): the
    pair separates along +x and ends touching, weights as usual."""
    n = np.array([0.3, 0.4, 0.1, 0.3, 0.4, 0.2])
    x = prox(oracle, "or_prox_pair3", n, [2.0, 6.0], 2, 3)
    assert np.all(np.isfinite(x))
    assert x[1] == 0.4 and x[4] == 0.4               # no y motion
    assert x[0] < 0.3 < x[3]                          # circle 1 to -x, circle 2 to +x
    assert x[3] - x[0] == pytest.approx(x[2] + x[5], abs=1e-12)
    assert (0.3 - x[0]) * 2.0 == pytest.approx((x[3] - 0.3) * 6.0, rel=1e-12)
    # listed the other way round, the smaller circle still goes to -x: the
    # same result with the circles swapped
    n2 = np.concatenate([n[3:], n[:3]])
    x2 = prox(oracle, "or_prox_pair3", n2, [6.0, 2.0], 2, 3)
    assert x2 == pytest.approx(np.concatenate([x[3:], x[:3]]), rel=1e-12, abs=1e-15)


def test_radius_a1_at_zero_is_finite(oracle):
    """n = 0 with a = 1: the curvature term is identically zero; the result is 1/rho."""
    lib = oracle.load()
    assert lib.or_max_alpha_radius(0.0, 4.0, 1.0) == pytest.approx(0.25)


def _quality_of(oracle, z):
    """or_square_quality on a square graph whose Z is set to z (rows cx, cy, r)."""
    z = np.asarray(z, dtype=np.float64)
    g = oracle.Graph.square(z.shape[0], 0)
    g.array("Z")[:] = z.ravel()
    return g.square_quality(z.shape[0])


def test_square_quality_two_circles_closed_form(oracle):
    """Two circles 0.3 apart with radii 0.2 and 0.15 overlap by exactly
    0.05; both lie inside the square, so the container violation is 0 and
    the objective is 0.35.  Moving the second circle to distance 0.5 leaves
    no overlap (the measure is clamped at 0, not negative)."""
    ov, bx, sr = _quality_of(oracle, [[0.3, 0.5, 0.2], [0.6, 0.5, 0.15]])
    assert ov == pytest.approx(0.05, abs=1e-15)
    assert bx == 0.0
    assert sr == pytest.approx(0.35, abs=1e-15)
    ov, bx, sr = _quality_of(oracle, [[0.25, 0.5, 0.2], [0.75, 0.5, 0.15]])
    assert ov == 0.0
    # a diagonal pair: distance 5 * 0.06 (a 3-4-5 triangle), radii sum 0.31
    ov, _, _ = _quality_of(oracle, [[0.2, 0.2, 0.16], [0.38, 0.44, 0.15]])
    assert ov == pytest.approx(0.01, abs=1e-15)


@pytest.mark.parametrize("circle,viol", [
    ((0.05, 0.5, 0.1), 0.05),    # left wall: r - cx
    ((0.97, 0.5, 0.1), 0.07),    # right wall: cx + r - 1
    ((0.5, 0.02, 0.1), 0.08),    # bottom wall: r - cy
    ((0.5, 0.95, 0.1), 0.05),    # top wall: cy + r - 1
    ((0.5, 0.5, -0.01), 0.01),   # negative radius: -r
    ((0.5, 0.5, 0.5), 0.0),      # touching all four walls
])
def test_square_quality_container_closed_form(oracle, circle, viol):
    """One circle crossing one wall by a known amount (the other circle well
    inside and apart): the container violation is that amount."""
    ov, bx, sr = _quality_of(oracle, [circle, (0.9, 0.9, 0.0)] if circle[2] != 0.5 else [circle])
    assert bx == pytest.approx(viol, abs=1e-15)


def test_square_quality_maxima_over_many_pairs(oracle):
    """Among many circles the reported overlap is the largest one: a row of
    touching circles (overlap 0) with one pair pushed together by 0.013,
    and the container violation is the largest of several."""
    r = 0.04
    z = [[0.1 + 2 * r * k, 0.5, r] for k in range(10)]
    z[6][0] -= 0.013
    z[0][1] = 0.03   # bottom violation r - cy = 0.01
    z[9][1] = 0.975  # top violation cy + r - 1 = 0.015
    ov, bx, sr = _quality_of(oracle, z)
    assert ov == pytest.approx(0.013, abs=1e-14)
    assert bx == pytest.approx(0.015, abs=1e-15)
    assert sr == pytest.approx(10 * r, abs=1e-15)
