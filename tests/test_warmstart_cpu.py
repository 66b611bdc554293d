"""CPU checks of the learned-start host logic (normalisation folding, A/B
summary) against the FP64 MLP oracle."""
import numpy as np

import mlp_oracle as M
from rod_dynamics_py import warmstart as WS


def test_fold_equals_normalised_network():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, M.NPARAM)
    X = rng.normal(3.0, 5.0, size=(50, M.IN))
    mx, sx = X.mean(0), X.std(0) + 0.5
    mt, st = np.array([0.1, -2.0, 3.0]), np.array([2.0, 0.5, 4.0])
    y_norm = M.forward(p, (X - mx) / sx) * st + mt
    pf = WS.fold(p.astype(np.float32), mx, sx, mt, st).astype(np.float64)
    y_fold = M.forward(pf, X)
    assert np.allclose(y_fold, y_norm, rtol=1e-4, atol=1e-4 * np.abs(y_norm).max())
    # hidden layer untouched
    assert np.array_equal(M.unpack(pf)[2], M.unpack(p.astype(np.float32).astype(np.float64))[2])


def test_normalisation_floor():
    X = np.zeros((10, M.IN))
    T = np.ones((10, M.OUT))
    mx, sx, mt, st = WS.normalisation(X, T)
    assert np.all(sx == 1e-3) and np.all(st == 1e-3) and np.all(mt == 1.0)


def test_ab_summary_ratios():
    rows = [{"manifolds": 10,
             "cold": {"sweeps": 100, "residual": 1e-9, "manifold_updates": 1000, "ms": 2.0},
             "previous": {"sweeps": 40, "residual": 1e-9, "manifold_updates": 400, "ms": 1.0},
             "learned": {"sweeps": 20, "residual": 5e-9, "manifold_updates": 100, "ms": 0.5},
             "learned_all": {"sweeps": 10, "residual": 6e-9, "manifold_updates": 80, "ms": 0.4},
             "learned_fresh": {"sweeps": 50, "residual": 7e-9, "manifold_updates": 160, "ms": 0.6}}, None]
    s = WS.ab_summary(rows)
    assert s["steps"] == 1
    assert s["sweeps_ratio_previous_over_learned"] == 2.0 and s["sweeps_ratio_cold_over_learned"] == 5.0
    assert s["sweeps_ratio_previous_over_learned_all"] == 4.0 and s["learned_all"]["residual_max"] == 6e-9
    assert s["updates_ratio_previous_over_learned"] == 4.0
    assert s["learned"]["residual_max"] == 5e-9
    assert s["sweeps_ratio_previous_over_learned_fresh"] == 0.8 and s["updates_ratio_cold_over_learned_fresh"] == 6.25
