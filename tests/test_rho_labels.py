"""Host logic of the rho-policy training data: the raised-edge labels and
the dataset files (no GPU)."""
import numpy as np

from engine_amd import rho_train as rt


def _features(n, frac, hot_pairs, hot_unary=0):
    """n rows: residual feature log10(1e-6)/5 except hot_pairs pair rows at
    log10(1e-3)/5 (and hot_unary radius rows at the same residual)."""
    f = np.zeros((n, 6), dtype=np.float32)
    f[:, 0] = np.log10(1e-6) / 5
    f[:, 3] = 1.0
    f[:hot_pairs, 0] = np.log10(1e-3) / 5
    f[hot_pairs:hot_pairs + hot_unary, 0] = np.log10(1e-3) / 5
    f[hot_pairs:hot_pairs + hot_unary, 3] = 0.0
    f[hot_pairs:hot_pairs + hot_unary, 4] = 1.0
    f[:, 5] = frac
    return f


def test_late_high_residual_pair_edges_are_raised_and_kept():
    rng = np.random.default_rng(0)
    f = _features(5000, 0.5, hot_pairs=7, hot_unary=3)
    fr, t = rt._label(f, 200.0, rng, 100)
    base = np.log10(200.0)
    hot = (fr[:, 0] > np.log10(rt.RAISE_RESIDUAL) / 5) & (fr[:, 3] == 1.0)
    # every hot pair row of the snapshot is present RAISE_WEIGHT more times
    assert hot.sum() >= 7 * rt.RAISE_WEIGHT
    assert np.allclose(t[hot], base + np.log10(rt.RAISE_FACTOR))
    assert np.allclose(t[~hot], base)
    # a radius edge with the same residual is not raised
    unary_hot = (fr[:, 0] > np.log10(rt.RAISE_RESIDUAL) / 5) & (fr[:, 4] == 1.0)
    assert np.allclose(t[unary_hot], base)


def test_early_snapshots_are_not_raised():
    rng = np.random.default_rng(1)
    f = _features(3000, rt.RAISE_FROM * 0.5, hot_pairs=50)
    fr, t = rt._label(f, 70.0, rng, 3000)
    assert fr.shape[0] == 3000
    assert np.allclose(t, np.log10(70.0))


def test_rows_round_trip(tmp_path):
    f = np.random.default_rng(2).standard_normal((100, 6)).astype(np.float32)
    t = np.arange(100, dtype=np.float32)
    teacher = {"64/0": {"rho_star": 50.0}}
    p = tmp_path / "rows.npz"
    rt.save_rows(p, f, t, teacher)
    f2, t2, te2 = rt.load_rows(p)
    assert np.array_equal(f, f2) and np.array_equal(t, t2) and te2 == teacher
