"""GPU parity of the learned PSOR warm start: FP32 training gradients and
AdamW against the FP64 oracle, BF16 matrix-core inference against the
BF16-emulating oracle forward, the feature kernel against the oracle's
restatement, the warm_mode 2 start, and that a solve from the learned start
still reaches the residual the oracle confirms."""
import numpy as np
import pytest

import mlp_oracle as M
import oracle_py as O
from rod_dynamics_py import _lib as L

from test_gpu_nsc import gpu_load, oracle_load, overlapping_bed

pytestmark = pytest.mark.gpu


def _rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, M.NPARAM).astype(np.float32)


def _blocks(g):
    return dict(zip(["W1", "b1", "W2", "b2", "W3", "b3"], M.unpack(g)))


@pytest.mark.parametrize("n", [5, 1001, 20000])
def test_grad_matches_fp64_oracle(gpu_ctx_factory, n):
    print(f"mlp grad n={n}", flush=True)
    ctx = gpu_ctx_factory()
    p = _rand_params(100 + n)
    ctx.mlp_set_params(p)
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, M.IN)).astype(np.float32)
    T = rng.normal(size=(n, M.OUT)).astype(np.float32)
    g, loss = ctx.mlp_grad(X, T)
    p64 = p.astype(np.float64)
    gr = M.grad(p64, X.astype(np.float64), T.astype(np.float64))
    lr = M.loss(p64, X, T)
    assert abs(loss - lr) <= 1e-5 * lr
    # FP32 sums of n terms: relative error per block well inside 1e-4
    for name, (a, b) in zip(_blocks(g), zip(M.unpack(g), M.unpack(gr))):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < 1e-4, (name, err)


def test_adamw_steps_match_oracle(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    p0 = _rand_params(7)
    ctx.mlp_set_params(p0)
    rng = np.random.default_rng(8)
    n = 512
    X = rng.normal(size=(n, M.IN)).astype(np.float32)
    T = rng.normal(size=(n, M.OUT)).astype(np.float32)
    lr, wd = 1e-3, 0.01
    ctx.mlp_train(X, T, epochs=1, batch=n, lr=lr, weight_decay=wd, seed=1)  # one full-batch step
    p1 = ctx.mlp_get_params().astype(np.float64)
    g0 = M.grad(p0.astype(np.float64), X, T)
    z = np.zeros(M.NPARAM)
    e1, m1, v1 = M.adamw_step(p0.astype(np.float64), g0, z, z, 1, lr, wd)
    big = np.abs(g0) > 1e-3 * np.abs(g0).max()  # sign(g) is unambiguous here
    assert np.max(np.abs(p1 - e1)[big]) < 2e-6
    ctx.mlp_train(X, T, epochs=1, batch=n, lr=lr, weight_decay=wd, seed=2)  # second step, bias-corrected
    p2 = ctx.mlp_get_params().astype(np.float64)
    g1 = M.grad(p1, X, T)
    e2, _, _ = M.adamw_step(p1, g1, m1, v1, 2, lr, wd)
    assert np.max(np.abs(p2 - e2)[big]) < 2e-6


@pytest.mark.parametrize("n", [1, 16, 1007, 70001])
def test_bf16_inference_matches_oracle(gpu_ctx_factory, n):
    print(f"mlp infer n={n}", flush=True)
    ctx = gpu_ctx_factory()
    p = _rand_params(3 + n)
    ctx.mlp_set_params(p)
    rng = np.random.default_rng(n + 1)
    X = rng.normal(size=(n, M.IN)).astype(np.float32)
    Y = ctx.mlp_infer(X)
    yb = M.forward_bf16(p, X)
    y = M.forward(p.astype(np.float64), X)
    # BF16 emulation: differences come only from FP32 vs FP64 sums flipping a
    # BF16 rounding of a hidden activation
    assert np.all(np.abs(Y - yb) <= 1e-2 * (1.0 + np.abs(yb))), np.max(np.abs(Y - yb))
    scale = np.max(np.abs(y)) + 1.0
    assert np.max(np.abs(Y - y)) < 0.05 * scale
    # a fragment-layout mistake would move outputs by O(scale)
    assert np.mean(np.abs(Y - yb)) < 1e-3 * scale


def _settled_pair_of_builds(ctx, B, p):
    """Solve one step (fills the previous-step cache), advance, build again."""
    gpu_load(ctx, B)
    ctx.detect(L.detect_params())
    nm0 = ctx.nsc_build(p)
    ctx.nsc_solve(p, allow_noconv=True)
    rec0, ab0, lam0 = ctx.nsc_manifolds(nm0)
    ctx.integrate(L.RDK_NSC_POS)
    ctx.integrate(L.RDK_NSC_FREE)
    ctx.detect(L.detect_params())
    return ab0, lam0


def test_features_match_oracle(gpu_ctx_factory):
    B = overlapping_bed(n=2000, seed=61, mu=0.4)
    ctx = gpu_ctx_factory()
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                     warm_mode=1, warm_walls=1)
    ab0, lam0 = _settled_pair_of_builds(ctx, B, p)
    nm = ctx.nsc_build(p)
    rec, ab, lam_start = ctx.nsc_manifolds(nm)
    Xg, lsg = ctx.mlp_features(nm)
    s = ctx.download_state()
    prev = {(int(a), int(b)): l for (a, b), l in zip(ab0, lam0[:, 0]) if b >= 0}
    prev_w = {(int(a), -1 - int(b)): l for (a, b), l in zip(ab0, lam0[:, 0]) if b < 0}
    nwall = int((ab[:, 1] < 0).sum())
    assert nm > 500 and nwall > 20
    assert sum((int(a), int(b)) in prev for a, b in ab if b >= 0) > 100  # persisting pairs
    Xo, lso = M.features(rec.T, ab[:, 0], ab[:, 1], s, B["r"], np.full(len(B["r"]), 0.4), 1e-3, prev, prev_w)
    assert np.allclose(lsg, lso, rtol=1e-14, atol=0)
    err = np.abs(Xg - Xo) / (1.0 + np.abs(Xo))
    assert err.max() < 1e-6, err.max(axis=0)
    # the previous-step start of these manifolds is the cached solution
    assert np.allclose(Xo[:, 11] * lso, np.maximum(lam_start[:, 0], 0.0), rtol=1e-12, atol=1e-300)


def test_learned_start_keeps_cached_impulses_and_predicts_the_rest(gpu_ctx_factory):
    B = overlapping_bed(n=2000, seed=62, mu=0.4)
    ctx = gpu_ctx_factory()
    prm = _rand_params(5, 0.1)
    prm[M.OFF_W3:M.OFF_B3] = 0.0
    prm[M.OFF_B3:] = [2.0, 0.5, -0.5]  # output = b3: a load of 2 lambda_s on every uncached manifold
    ctx.mlp_set_params(prm)
    p1 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=1, warm_walls=1)
    _settled_pair_of_builds(ctx, B, p1)
    nm = ctx.nsc_build(p1)
    _, _, lam_prev = ctx.nsc_manifolds(nm)
    X, ls = ctx.mlp_features(nm)
    Y = ctx.mlp_infer(X)
    p2 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=2, warm_walls=1)
    assert ctx.nsc_build(p2) == nm
    _, _, lam_learned = ctx.nsc_manifolds(nm)
    expect = M.apply_prediction(Y, ls, 0.4, lam_prev)
    assert np.allclose(lam_learned, expect, rtol=1e-12, atol=1e-15 * np.abs(expect).max())
    kept = lam_prev[:, 0] > 0
    assert kept.sum() > 100 and (~kept).sum() > 10
    assert np.array_equal(lam_learned[kept], lam_prev[kept])
    assert np.all(lam_prev[~kept, 0] == 0.0)
    assert np.allclose(lam_learned[~kept, 0], 2.0 * ls[~kept], rtol=1e-6)


def test_solve_from_learned_start_reaches_residual(gpu_ctx_factory):
    B = overlapping_bed(n=2000, seed=63, mu=0.4)
    ctx = gpu_ctx_factory()
    ctx.mlp_init(seed=4)
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                     warm_mode=2, warm_walls=1)
    _settled_pair_of_builds(ctx, B, p)
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p)
    assert st["residual"] <= 1e-8
    _, _, lam = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(x=s1["x"], q=s1["q"], v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam)
    W.apply_warm()
    so = W.state()
    assert np.allclose(so["v"], s1["v"], rtol=0, atol=1e-11)
    W.set_state(v=s1["v"], w=s1["w"])
    op = O.nsc_params(mu=0.4, beta=0.0, cfm=0.0, omega=1.3, friction_proj=1)
    assert W.residual(op) <= 1e-8 * (1 + 1e-6)
    assert lam[:, 0].min() >= 0.0
    assert np.all(np.hypot(lam[:, 1], lam[:, 2]) <= 0.4 * lam[:, 0] * (1 + 1e-12) + 1e-15)


def test_training_fits_a_smooth_target(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    ctx.mlp_init(seed=9)
    rng = np.random.default_rng(10)
    n = 8192
    X = rng.normal(size=(n, M.IN)).astype(np.float32)
    T = np.column_stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], np.tanh(X[:, 3]), 0.3 * X[:, 4] ** 2 - 0.3])
    T = T.astype(np.float32)
    p0 = ctx.mlp_get_params()
    loss0 = M.loss(p0.astype(np.float64), X, T)
    loss = ctx.mlp_train(X, T, epochs=30, batch=256, lr=3e-3, weight_decay=1e-4, seed=11)
    p1 = ctx.mlp_get_params().astype(np.float64)
    loss1 = M.loss(p1, X, T)
    assert loss1 < 0.05 * loss0, (loss0, loss1, loss)
    # the trained BF16 inference tracks the FP64 forward of the trained weights
    Y = ctx.mlp_infer(X[:1000])
    assert np.mean(np.abs(Y - M.forward(p1, X[:1000]))) < 0.02


def test_learned_start_for_every_manifold(gpu_ctx_factory):
    """warm_mode 3: the prediction replaces the cached start on every manifold
    (the cached impulse is one of the inputs)."""
    B = overlapping_bed(n=2000, seed=64, mu=0.4)
    ctx = gpu_ctx_factory()
    prm = _rand_params(6, 0.1)
    prm[M.OFF_W3:M.OFF_B3] = 0.0
    prm[M.OFF_B3:] = [2.0, 0.5, -0.5]
    ctx.mlp_set_params(prm)
    p1 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=1, warm_walls=1)
    _settled_pair_of_builds(ctx, B, p1)
    nm = ctx.nsc_build(p1)
    _, _, lam_prev = ctx.nsc_manifolds(nm)
    X, ls = ctx.mlp_features(nm)
    Y = ctx.mlp_infer(X)
    p3 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=3, warm_walls=1)
    assert ctx.nsc_build(p3) == nm
    _, _, lam3 = ctx.nsc_manifolds(nm)
    expect = M.apply_prediction(Y, ls, 0.4, lam_prev, all_manifolds=True)
    assert np.allclose(lam3, expect, rtol=1e-12, atol=1e-15 * np.abs(expect).max())
    assert (lam_prev[:, 0] > 0).sum() > 100
    assert np.allclose(lam3[:, 0], 2.0 * ls, rtol=1e-6)


def test_learned_start_from_the_contact_alone(gpu_ctx_factory):
    """warm_mode 4: the prediction on every manifold, with the previous-step
    impulse input left at 0 (the cache is not read): the network's output for
    the features of the cached build with feature 11 zeroed."""
    B = overlapping_bed(n=2000, seed=65, mu=0.4)
    ctx = gpu_ctx_factory()
    prm = _rand_params(7, 0.1)
    ctx.mlp_set_params(prm)
    p1 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=1, warm_walls=1)
    _settled_pair_of_builds(ctx, B, p1)
    nm = ctx.nsc_build(p1)
    _, _, lam_prev = ctx.nsc_manifolds(nm)
    X, ls = ctx.mlp_features(nm)
    assert (X[:, 11] > 0).sum() > 100  # cached impulses are an input of the cached build
    X0 = X.copy()
    X0[:, 11] = 0.0
    Y0 = ctx.mlp_infer(X0)
    p4 = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=1,
                      warm_mode=4, warm_walls=1)
    assert ctx.nsc_build(p4) == nm
    _, _, lam4 = ctx.nsc_manifolds(nm)
    expect = M.apply_prediction(Y0, ls, 0.4, lam_prev, all_manifolds=True)
    assert np.allclose(lam4, expect, rtol=1e-12, atol=1e-15 * np.abs(expect).max())
    # the mode really ignores the cache: the cached-input prediction differs
    Y = ctx.mlp_infer(X)
    assert not np.allclose(M.apply_prediction(Y, ls, 0.4, lam_prev, all_manifolds=True), lam4)
