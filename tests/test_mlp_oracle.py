"""CPU pins of the FP64 MLP oracle (learned PSOR warm start)."""
import numpy as np

import mlp_oracle as M


def _params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, M.NPARAM)


def test_layout_size():
    assert M.NPARAM == 12 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3 == 5187
    p = _params(0)
    assert np.array_equal(M.pack(*M.unpack(p)), p)


def test_dsilu_matches_difference_quotient():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.dsilu(z))) < 1e-9
    assert M.silu(0.0) == 0.0 and abs(M.dsilu(0.0) - 0.5) < 1e-15


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2, 0.3)
    X = rng.normal(size=(7, M.IN))
    T = rng.normal(size=(7, M.OUT))
    g = M.grad(p, X, T)
    # every parameter of the small output layers and a sample of the rest
    idx = np.concatenate([np.arange(M.OFF_W3, M.NPARAM), rng.choice(M.OFF_W3, 200, replace=False)])
    h = 1e-6
    for k in idx:
        e = np.zeros(M.NPARAM)
        e[k] = h
        fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_gradient_closed_form_with_zero_output_layer():
    rng = np.random.default_rng(3)
    W1, b1, W2, b2, W3, b3 = M.unpack(_params(4))
    W3 = np.zeros_like(W3)
    p = M.pack(W1, b1, W2, b2, W3, b3)
    X = rng.normal(size=(9, M.IN))
    T = rng.normal(size=(9, M.OUT))
    g = M.grad(p, X, T)
    gW1, gb1, gW2, gb2, gW3, gb3 = M.unpack(g)
    assert np.all(gW1 == 0) and np.all(gb1 == 0) and np.all(gW2 == 0) and np.all(gb2 == 0)
    Y = np.broadcast_to(b3, T.shape)
    assert np.allclose(gb3, 2.0 * (Y - T).sum(0) / T.size, rtol=1e-14, atol=1e-15)


def test_planted_error_is_caught():
    """A gradient with a dropped SiLU' factor fails the finite-difference check."""
    rng = np.random.default_rng(5)
    p = _params(6, 0.3)
    X = rng.normal(size=(5, M.IN))
    T = rng.normal(size=(5, M.OUT))
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    z1 = X @ W1 + b1
    h1 = M.silu(z1)
    z2 = h1 @ W2 + b2
    h2 = M.silu(z2)
    dY = 2.0 * (h2 @ W3 + b3 - T) / T.size
    wrong_gb2 = (dY @ W3.T).sum(0)  # missing dsilu(z2)
    h = 1e-6
    k = M.OFF_B2 + 3
    e = np.zeros(M.NPARAM)
    e[k] = h
    fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / (2 * h)
    assert abs(fd - M.grad(p, X, T)[k]) < 1e-8
    assert abs(fd - wrong_gb2[3]) > 1e-4


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 0.0, np.inf], np.float32)
    r = M.bf16(a)
    # ties to even: 1+2^-8 -> 1, 1+3*2^-8 -> 1+2^-6
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6]
    assert r[4] == -2.5 and r[5] == 0.0 and np.isinf(r[6])


def test_forward_bf16_equals_forward_on_representable_case():
    # weights and inputs exactly representable; ReLU-free network with all
    # hidden pre-activations 0 (SiLU(0) = 0 is exact) -> identical outputs
    p = np.zeros(M.NPARAM)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    b3 = np.array([0.5, -0.25, 1.0])
    p = M.pack(W1, b1, W2, b2, W3, b3)
    X = np.ones((4, M.IN))
    assert np.array_equal(M.forward_bf16(p, X), M.forward(p, X))
    # generic weights: within the BF16 rounding bound of the FP64 forward
    rng = np.random.default_rng(7)
    p = _params(8, 0.2)
    X = rng.normal(size=(64, M.IN))
    y, yb = M.forward(p, X), M.forward_bf16(p, X)
    scale = np.max(np.abs(y)) + 1.0
    assert np.max(np.abs(y - yb)) < 0.05 * scale
    assert np.max(np.abs(y - yb)) > 0.0


def test_features_single_contacts():
    # one rod resting on the floor: identity orientation, contact normal +z
    # (from the rod towards the wall is -z; the manifold normal points A->B)
    mf = np.zeros((42, 2))
    mf[0:3, 0] = [0, 0, -1]
    mf[3:6, 0] = [1, 0, 0]
    mf[6:9, 0] = [0, -1, 0]
    mf[9:12, 0] = [2.0, 0, -0.5]
    mf[35, 0] = 4.0
    mf[38, 0] = -0.05
    mf[39, 0] = -0.3
    # rod-rod contact between bodies 0 and 1, body 1 rotated 90 deg about z
    mf[0:3, 1] = [1, 0, 0]
    mf[3:6, 1] = [0, 1, 0]
    mf[6:9, 1] = [0, 0, 1]
    mf[35, 1] = 2.0
    mf[38, 1] = 0.01
    c, s = np.cos(np.pi / 4), np.sin(np.pi / 4)
    state = {
        "q": np.array([[1.0, 0, 0, 0], [c, 0, 0, s]]),
        "v": np.array([[0.0, 0, -0.3], [1.0, 0, 0]]),
        "w": np.array([[0.0, 0.1, 0.0], [0.0, 0, 0]]),
    }
    r = np.array([0.5, 0.25])
    vs = 0.01
    X, ls = M.features(mf, [0, 0], [-1, 1], state, r, np.array([0.5, 0.5]), vs,
                       prev={(0, 1): 0.004}, prev_wall={(0, 0): -1.0})
    # wall contact: v_a at the contact = v + w x r_a = (0,0,-0.3) + (0,0.1,0)x(2,0,-0.5) = (-0.05,0,-0.5)
    assert np.allclose(X[0], [-0.1, 0, 0, -1, 0, 0, 0, -30.0, 5.0, 0.0, 0.5, 0.0])
    assert np.isclose(ls[0], vs / 4.0)
    # rod-rod: n in B's frame = R_z(-90) (1,0,0) = (0,-1,0); r_a = 0 so
    # v_rel = v_b - v_a = (1,0,0) - (0,0,-0.3) = (1,0,0.3): t1 part 0, t2 part 0.3
    assert np.allclose(X[1], [0.01 / 0.75, 1, 0, 0, 0, -1, 0, 0, 0, 30.0, 0.5, 0.004 / (vs / 2.0)])


def test_apply_prediction_cone():
    Y = np.array([[2.0, 3.0, 4.0], [-1.0, 0.1, 0.1], [1.0, 0.1, 0.0]])
    L = M.apply_prediction(Y, np.array([1.0, 1.0, 2.0]), 0.5)
    assert np.allclose(L[0], [2.0, 0.6, 0.8])   # |t| = 5 scaled to mu*ln = 1
    assert np.allclose(L[1], [0.0, 0.0, 0.0])
    assert np.allclose(L[2], [2.0, 0.2, 0.0])   # inside the disc, kept


def test_apply_prediction_keeps_previous_impulses():
    Y = np.array([[0.5, 0.0, 0.0], [3.0, 1.0, 0.0], [-3.0, 0.0, 0.0]])
    prev = np.array([[1.0, 0.2, -0.1], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    L = M.apply_prediction(Y, np.array([2.0, 1.0, 1.0]), 0.5, prev)
    assert np.array_equal(L[0], prev[0])           # a previous impulse is kept as is
    assert np.allclose(L[1], [3.0, 1.0, 0.0])      # new contact: prediction, inside the disc (|t| = 1 <= 1.5)
    assert np.allclose(L[2], [0.0, 0.0, 0.0])      # predicted separation


def test_apply_prediction_to_every_manifold():
    Y = np.array([[0.5, 0.0, 0.0], [3.0, 1.0, 0.0]])
    prev = np.array([[1.0, 0.2, -0.1], [0.0, 0.0, 0.0]])
    L = M.apply_prediction(Y, np.array([2.0, 1.0]), 0.5, prev, all_manifolds=True)
    assert np.allclose(L[0], [1.0, 0.0, 0.0])      # the prediction replaces the cached impulse
    assert np.allclose(L[1], [3.0, 1.0, 0.0])


def test_adamw_first_step_closed_form():
    # first step: m_hat = g, v_hat = g^2, so the step is lr * g / (|g| + eps)
    rng = np.random.default_rng(9)
    p = rng.normal(size=M.NPARAM)
    g = rng.normal(size=M.NPARAM)
    z = np.zeros(M.NPARAM)
    p1, m1, v1 = M.adamw_step(p, g, z, z, 1, lr=1e-2, wd=0.1)
    decay = np.full(M.NPARAM, 1.0)
    decay[M.OFF_W1:M.OFF_B1] = decay[M.OFF_W2:M.OFF_B2] = decay[M.OFF_W3:M.OFF_B3] = 1.0 - 1e-3
    assert np.allclose(p1, p * decay - 1e-2 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 1e-3 * g * g)
    # biases are not decayed
    assert p1[M.OFF_B1] == p[M.OFF_B1] - 1e-2 * g[M.OFF_B1] / (abs(g[M.OFF_B1]) + 1e-8)
