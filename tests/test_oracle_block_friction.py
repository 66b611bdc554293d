"""Pins of the oracle's block tangential PSOR updates (friction_proj 2 and 3,
CPU only).

friction_proj 1 steps the tangential impulse by one scalar step omega/g_tb
and scales it radially onto the disc |lambda_t| <= mu lambda_n.  Both block
updates take the step lambda_t - omega G^-1 w_t with the 2x2 tangential
Delassus block G; when that lands outside the disc, friction_proj 2 projects
it in G's metric and friction_proj 3 takes friction_proj 1's step instead.
All three have the same fixed points (lambda_t in the disc, -w_t in the
disc's normal cone at lambda_t: maximal dissipation).

* the metric projection against a brute-force search over the circle, its
  KKT conditions, and the isotropic case (a radial scale);
* one update of an isolated sticking contact zeroes both tangential slip
  components at once (a dropped or mis-signed off-diagonal G12 would not);
* friction_proj 3 reduces to friction_proj 2 on a sticking contact and to
  friction_proj 1 on a sliding one, bit for bit;
* a converged bed satisfies the friction-cone conditions, checked directly
  from the body velocities.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from test_oracle_nsc import random_bed


def brute_projection(G, r, y, n=20001):
    """argmin over |x| = r of (x - y)^T G (x - y) (|y| > r: the minimiser is on
    the circle): dense angle grid, then golden-section refinement."""
    def f(th):
        d = r * np.array([np.cos(th), np.sin(th)]) - y
        return d @ G @ d
    th = np.linspace(-math.pi, math.pi, n)
    X = r * np.stack([np.cos(th), np.sin(th)], 1) - y
    vals = np.einsum("ij,jk,ik->i", X, G, X)
    k = int(np.argmin(vals))
    a, b = th[k] - 2 * math.pi / n, th[k] + 2 * math.pi / n
    g = (math.sqrt(5) - 1) / 2
    for _ in range(200):
        c, d = b - g * (b - a), a + g * (b - a)
        if f(c) < f(d):
            b = d
        else:
            a = c
    t = 0.5 * (a + b)
    return r * np.array([math.cos(t), math.sin(t)])


def spd_cases():
    rng = np.random.default_rng(4)
    out = [(np.diag([3.0, 0.2]), 1.0), (np.array([[2.0, 1.5], [1.5, 2.0]]), 0.3),
           (np.array([[2.0, -1.5], [-1.5, 2.0]]), 0.3), (np.array([[1.0, 0.999], [0.999, 1.0]]), 2.0)]
    for _ in range(12):
        A = rng.normal(size=(2, 2))
        out.append((A @ A.T + rng.uniform(0.01, 1.0) * np.eye(2), rng.uniform(0.05, 3.0)))
    return out


@pytest.mark.parametrize("case", range(16))
def test_metric_projection_matches_brute_force(case):
    G, r = spd_cases()[case]
    rng = np.random.default_rng(100 + case)
    for _ in range(6):
        y = rng.normal(size=2)
        y *= r * rng.uniform(1.01, 30.0) / np.linalg.norm(y)
        x = O.disc_project_metric(G[0, 0], G[0, 1], G[1, 1], r, y)
        xb = brute_projection(G, r, y)
        assert np.linalg.norm(x) == pytest.approx(r, rel=1e-14)
        assert np.allclose(x, xb, rtol=0, atol=1e-7 * r), (x, xb)
        # KKT: G (y - x) = nu x with nu >= 0
        g = G @ (y - x)
        nu = (g @ x) / (x @ x)
        assert nu >= 0 and np.allclose(g, nu * x, rtol=0, atol=1e-10 * np.linalg.norm(g))
        # the check has the power to reject plausible mistakes: the Euclidean
        # radial scale, and G12 taken with the wrong sign
        if abs(G[0, 0] - G[1, 1]) + 2 * abs(G[0, 1]) > 0.2 * np.trace(G):
            assert not np.allclose(r * y / np.linalg.norm(y), xb, rtol=0, atol=1e-7 * r)
        if abs(G[0, 1]) > 0.1 * np.trace(G):
            xs = O.disc_project_metric(G[0, 0], -G[0, 1], G[1, 1], r, y)
            assert not np.allclose(xs, xb, rtol=0, atol=1e-7 * r)


def test_metric_projection_isotropic_is_radial_scale():
    for a in (1e-3, 1.0, 7.5):
        y = np.array([3.0, -4.0])
        x = O.disc_project_metric(a, 0.0, a, 2.0, y)
        assert np.allclose(x, [1.2, -1.6], rtol=1e-15, atol=0)


def _slip(st, rec, ab):
    """relative tangential velocity (t1, t2) of each manifold from body velocities"""
    out = []
    for i in range(len(ab)):
        a, b = ab[i]
        t1, t2, ra, rb = rec[i, 3:6], rec[i, 6:9], rec[i, 9:12], rec[i, 12:15]
        va = st["v"][a] + np.cross(st["w"][a], ra)
        vb = (st["v"][b] + np.cross(st["w"][b], rb)) if b >= 0 else np.zeros(3)
        out.append([t1 @ (vb - va), t2 @ (vb - va)])
    return np.array(out)


def _g12(rec, ab, i):
    ra, rb, t1 = rec[i, 9:12], rec[i, 12:15], rec[i, 3:6]
    g = np.cross(ra, t1) @ rec[i, 21:24]
    if ab[i, 1] >= 0:
        g += np.cross(rb, t1) @ rec[i, 30:33]
    return g


def _one_contact(wall):
    """the bed with its approaching contact (rod-rod or rod-wall) whose
    tangential block is furthest from isotropic"""
    W, B = random_bed(seed=9)
    W.detect(0, 0.0)
    W.nsc_build(O.nsc_params(mu=0.5, cfm=0.0, warm_start=0))
    rec, ab, lam = W.manifolds()
    cand = [i for i in range(len(ab)) if (ab[i, 1] < 0) == wall and rec[i, 39] < -0.1]

    def aniso(i):
        g1, g2, g12 = rec[i, 36], rec[i, 37], _g12(rec, ab, i)
        return (abs(g1 - g2) + 2 * abs(g12)) / (g1 + g2)
    i = max(cand, key=aniso)
    assert aniso(i) > 0.05
    return W, rec[i:i + 1].copy(), ab[i:i + 1].copy(), W.state()


def _update_once(W, rec, ab, st0, proj, mu):
    W.set_state(v=st0["v"], w=st0["w"])
    W.set_manifolds(rec, ab, np.zeros((1, 3)))
    W.sweeps(O.nsc_params(mu=mu, beta=0.0, cfm=0.0, omega=1.0, friction_proj=proj), 1)
    _, _, l1 = W.manifolds()
    assert l1[0, 0] > 0
    return l1, W.state()


@pytest.mark.parametrize("wall", [False, True])
def test_block_step_zeroes_slip_of_one_sticking_contact(wall):
    W, rec, ab, st0 = _one_contact(wall)
    res = {}
    for proj in (1, 2, 3):
        _, st = _update_once(W, rec, ab, st0, proj, 1e6)
        res[proj] = np.abs(_slip(st, rec, ab)[0])
    scale = np.abs(_slip(st0, rec, ab)[0]).max() + 1e-300
    assert res[2].max() <= 1e-12 * scale and res[3].max() <= 1e-12 * scale, (res, scale)
    # the scalar step leaves slip behind
    assert res[1].max() > 1e-3 * scale, (res, scale)


@pytest.mark.parametrize("wall", [False, True])
def test_hybrid_update_reduces_to_block_or_scalar_step(wall):
    W, rec, ab, st0 = _one_contact(wall)
    # sticking (the block step stays in the disc): friction_proj 3 == 2
    l3, s3 = _update_once(W, rec, ab, st0, 3, 1e6)
    l2, s2 = _update_once(W, rec, ab, st0, 2, 1e6)
    assert np.array_equal(l3, l2) and np.array_equal(s3["v"], s2["v"]) and np.array_equal(s3["w"], s2["w"])
    # sliding (mu small enough that the block step leaves the disc): 3 == 1
    l3, s3 = _update_once(W, rec, ab, st0, 3, 1e-3)
    l1, s1 = _update_once(W, rec, ab, st0, 1, 1e-3)
    l2, _ = _update_once(W, rec, ab, st0, 2, 1e-3)
    assert np.hypot(l3[0, 1], l3[0, 2]) == pytest.approx(1e-3 * l3[0, 0], rel=1e-12)
    assert np.array_equal(l3, l1) and np.array_equal(s3["v"], s1["v"]) and np.array_equal(s3["w"], s1["w"])
    assert not np.allclose(l2[0, 1:], l1[0, 1:], rtol=1e-6, atol=0)  # the metric projection differs


@pytest.mark.parametrize("proj", [2, 3])
def test_block_step_converges_to_friction_cone_conditions(proj):
    W, B = random_bed(seed=9)
    W.detect(0, 0.0)
    mu = 0.5
    p = O.nsc_params(mu=mu, beta=0.0, cfm=0.0, omega=1.3, tol=1e-10, friction_proj=proj, warm_start=0,
                     max_sweeps=100000)
    W.nsc_build(p)
    r, s = W.solve(p)
    assert r <= 1e-10 and s > 1
    rec, ab, lam = W.manifolds()
    st = W.state()
    slip = _slip(st, rec, ab)
    # normal relative velocity plus the restitution bias (beta = 0: no penetration term)
    wn = []
    for i in range(len(ab)):
        a, b = ab[i]
        n, ra, rb = rec[i, 0:3], rec[i, 9:12], rec[i, 12:15]
        va = st["v"][a] + np.cross(st["w"][a], ra)
        vb = (st["v"][b] + np.cross(st["w"][b], rb)) if b >= 0 else np.zeros(3)
        e, vpre = rec[i, 40], rec[i, 39]  # restitution bias e v_n,pre on approach
        wn.append(n @ (vb - va) + (e * vpre if vpre < 0 else 0.0))
    wn = np.array(wn)
    eps = 1e-8
    ln, lt = lam[:, 0], lam[:, 1:3]
    assert ln.min() >= 0 and wn.min() >= -eps
    assert np.all(np.minimum(ln * rec[:, 35], np.maximum(wn, 0)) <= eps)
    nt = np.linalg.norm(lt, axis=1)
    assert np.all(nt <= mu * ln * (1 + 1e-12) + 1e-15)
    stick = nt < mu * ln * (1 - 1e-6)
    assert stick.sum() > 10 and (~stick & (ln > 0)).sum() > 10
    # sticking: no slip; sliding: slip opposes the tangential impulse on B
    assert np.abs(slip[stick]).max() <= eps
    sl = ~stick & (ln > 0)
    u = lt[sl] / nt[sl, None]
    along = np.einsum("ij,ij->i", slip[sl], u)
    perp = slip[sl] - along[:, None] * u
    assert along.max() <= eps and np.abs(perp).max() <= eps
