"""The reference's randomInit stream (App::resetScene, main.cpp:2213-2294).

The library's host function (rdk_random_init, compiled by hipcc's clang) is
compared bit for bit with the oracle's restatement compiled by g++ (the
reference's compiler), which writes the reference's own expressions and so
lets g++ choose the evaluation order of `dw(gen) * e1 + dw(gen) * e2`.  Both
are pinned independently of libstdc++ by a numpy restatement of the published
algorithms underneath: std::mt19937's raw outputs (pinned by its 10000th
output elsewhere), the canonical float u = float(raw) / 2^32, and Marsaglia's
polar method (pairs x = 2u - 1, y = 2u' - 1, rejected unless 0 < r2 <= 1,
mult = sqrt(-2 log r2 / r2), first draw y mult, second x mult from the same
distribution object).  No GPU is used.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

f = np.float32


def _quats(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    return L.glm_normalize_quat_f32(q)


def _axis(q):  # glm: v + ((qv x v) w + qv x (qv x v)) 2 with v = +Y, FP32
    qv = q[:, 1:]
    v = np.zeros_like(qv)
    v[:, 1] = 1.0
    uv = np.stack([qv[:, 1] * v[:, 2] - v[:, 1] * qv[:, 2], qv[:, 2] * v[:, 0] - v[:, 2] * qv[:, 0],
                   qv[:, 0] * v[:, 1] - v[:, 0] * qv[:, 1]], 1)
    uuv = np.stack([qv[:, 1] * uv[:, 2] - uv[:, 1] * qv[:, 2], qv[:, 2] * uv[:, 0] - uv[:, 2] * qv[:, 0],
                    qv[:, 0] * uv[:, 1] - uv[:, 0] * qv[:, 1]], 1)
    return v + ((uv * q[:, :1]) + uuv) * f(2.0)


class PolarStream:
    """numpy restatement of normal_distribution<float> on mt19937 raw outputs."""

    def __init__(self, seed, k=4096):
        self.raw = [O.mt_raw(seed, i + 1) for i in range(k)]
        self.i = 0

    def uniform(self):
        u = f(self.raw[self.i]) / f(4294967296.0)
        self.i += 1
        return u if u < f(1.0) else np.nextafter(f(1.0), f(0.0))

    def pair(self):
        while True:
            x = f(2.0) * self.uniform() - f(1.0)
            y = f(2.0) * self.uniform() - f(1.0)
            r2 = x * x + y * y
            if not (r2 > f(1.0) or r2 == f(0.0)):
                break
        mult = np.sqrt(f(-2.0) * np.log(r2) / r2)
        return y * mult, x * mult  # first draw, the saved second draw


def _ulps(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("project", [False, True])
def test_library_stream_equals_gpp_oracle_bitwise(mode, project):
    n = 3000
    q = _quats(n, 11 + mode)
    rng = np.random.default_rng(5)
    mass = rng.uniform(0.1, 3.0, n).astype(np.float32)
    ip = rng.uniform(0.01, 1.0, n).astype(np.float32)
    skip = (rng.random(n) < 0.1).astype(np.uint8)
    args = dict(v_sigma=0.3, w_sigma=0.7, w_speed=1.5, kbt_trans=2.0, kbt_rot=0.5, seed=4242, project=project)
    v, w = L.random_init(q, mode, mass32=mass, i_perp32=ip, skip=skip, **args)
    vo, wo = O.random_init(q, mode, mass=mass, i_perp=ip, skip=skip, **args)
    assert np.array_equal(v, vo) and np.array_equal(w, wo)
    assert not np.any(v[skip == 1]) and not np.any(w[skip == 1])
    assert np.any(v[skip == 0])


def test_gaussian_draws_follow_the_polar_method_per_rod():
    """v = three draws of a fresh N(0, vSigma) object, w = three of a fresh
    N(0, wSigma) object, rod after rod: the second value of the third pair of
    each object is discarded with it.  A distribution shared across rods, or
    the pair's values taken in the other order, fails here."""
    n, seed = 200, 777
    q = _quats(n, 3)
    v, w = L.random_init(q, 0, v_sigma=1.0, w_sigma=2.0, seed=seed, project=False)
    P = PolarStream(seed)
    ev, ew = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    for i in range(n):
        a, b = P.pair()
        c, _ = P.pair()
        ev[i] = [a, b, c]
        a, b = P.pair()
        c, _ = P.pair()
        ew[i] = [a * f(2.0), b * f(2.0), c * f(2.0)]
    # numpy's float32 log may differ from the C library's logf in the last place
    assert _ulps(v, ev).max() <= 4 and _ulps(w, ew).max() <= 4
    # a stream that kept the discarded values (one distribution for all rods) differs at once
    P2 = PolarStream(seed)
    a, b = P2.pair()
    c, d = P2.pair()
    e, _ = P2.pair()
    wrong_w0 = np.array([d, e, 0], np.float32)
    assert _ulps(w[0, :2], wrong_w0[:2] * f(2.0)).max() > 1000


def test_thermal_draw_order_is_gpp_right_operand_first():
    """The reference's w = dw(gen) * e1 + dw(gen) * e2: g++ evaluates the
    right operand first, so e2 carries the first draw.  Checked against the
    polar stream: w . e2 = first draw, w . e1 = second draw (to rounding)."""
    n, seed = 50, 99
    q = _quats(n, 7)
    mass = np.full(n, 2.0, np.float32)
    ip = np.full(n, 0.5, np.float32)
    v, w = L.random_init(q, 1, kbt_trans=2.0, kbt_rot=0.5, seed=seed, mass32=mass, i_perp32=ip)
    P = PolarStream(seed)
    u = _axis(q)
    for i in range(n):
        a, b = P.pair()
        c, _ = P.pair()
        assert _ulps(v[i], np.array([a, b, c], np.float32)).max() <= 4  # sigma_v = sqrt(2 / 2) = 1
        first, second = P.pair()
        ui = u[i].astype(np.float64)
        e1 = np.cross(ui, [1.0, 0.0, 0.0] if abs(ui[0]) < 0.9 else [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(ui, e1)
        wi = w[i].astype(np.float64)
        assert abs(wi @ e2 - float(first)) < 1e-5 and abs(wi @ e1 - float(second)) < 1e-5  # sigma_w = 1
        assert abs(wi @ ui) < 1e-5


def test_projection_removes_the_spin_about_the_axis_and_statistics():
    n = 20000
    q = _quats(n, 13)
    v, w = L.random_init(q, 0, v_sigma=0.1, w_sigma=0.2, seed=42, project=True)
    u = _axis(q).astype(np.float64)
    wd = w.astype(np.float64)
    assert np.abs(np.sum(wd * u, axis=1)).max() < 1e-6
    assert abs(v.mean()) < 3e-3 and abs(v.std() - 0.1) < 2e-3
    # two of three components survive the projection: E|w|^2 = 2 wSigma^2
    assert abs(np.mean(np.sum(wd * wd, axis=1)) - 2 * 0.04) < 2e-3


def test_uniform_mode_speed_and_range():
    n = 5000
    q = _quats(n, 17)
    v, w = L.random_init(q, 2, v_sigma=0.5, w_speed=3.0, seed=8, project=False)
    assert v.min() >= -0.5 and v.max() < 0.5
    assert np.allclose(np.linalg.norm(w.astype(np.float64), axis=1), 3.0, rtol=1e-6)


def test_glm_quaternion_normalize_matches_float64():
    q = _quats(1000, 21)
    n = np.linalg.norm(q.astype(np.float64), axis=1)
    assert np.abs(n - 1.0).max() < 3e-7
    z = L.glm_normalize_quat_f32(np.zeros((2, 4), np.float32))
    assert np.array_equal(z, np.array([[1, 0, 0, 0]] * 2, np.float32))
    assert math.isclose(float(L.glm_normalize_quat_f32(np.array([[0, 0, 3, 4]], np.float32))[0, 3]), 0.8,
                        rel_tol=1e-7)
