"""Pins of the oracle's force laws and integrators (CPU only).

Closed forms: the soft potential's force is minus its distance derivative;
Hertz scaling delta^2 (reference law, hertz_mindlin.cpp:261-264) and
delta^1.5 (textbook Hertz) with the (4/3) E* sqrt(R*) prefactor; Mindlin
spring xi += v_t dt and Coulomb cap; history pruning after 4 frames
(hertz_mindlin.hpp:143); momentum balance per contact; exact free flight;
constant-omega rotation; kinetic energy (main.cpp:3357-3368).
"""
import math

import numpy as np
import pytest

import oracle_py as O


def two_crossed(gap_dist, rA=0.5, rB=0.5, vb=(0, 0, 0), wb=(0, 0, 0)):
    """Rod A along y at the origin, rod B along z at x = gap_dist."""
    c = math.sqrt(0.5)
    x = np.array([[0.0, 0, 0], [gap_dist, 0, 0]])
    q = np.array([[1.0, 0, 0, 0], [c, c, 0, 0]])
    v = np.array([[0.0, 0, 0], list(vb)])
    w = np.array([[0.0, 0, 0], list(wb)])
    W = O.World(x, q, v, w, np.array([rA, rB]), np.array([2.0, 2.0]), np.array([1.0, 2.0]), np.array([0.5, 0.5]))
    W.set_box([-10] * 3, [10] * 3, [0, 0, 0])
    W.set_physics((0, 0, 0), 1e-4)
    return W


def soft_force_and_pe(d, p):
    W = two_crossed(d)
    W.detect(0, p.delta)
    pe = W.soft_forces(p)
    st = W.state()
    return st["f"][1, 0], pe


@pytest.mark.parametrize("d", [0.9, 0.99, 0.994, 0.997, 1.001, 1.004])
def test_soft_force_against_pe_gradient(d):
    # Quadratic regime (d <= h - delta): F = -dU/dd.  Barrier regime: the
    # reference's potentialGradient (soft_contact.cpp:222-229) is
    # -2/K1^2 log(1+e) e/(1+e), i.e. the true dU/dd divided by K1 = 15/delta;
    # the restatement keeps that, so F = -(dU/dd)/K1 there.
    delta = 0.005
    p = O.soft_params(delta=delta, k_scaler=1000.0, enable_friction=0)
    f, pe = soft_force_and_pe(d, p)
    eps = 1e-7
    _, pp = soft_force_and_pe(d + eps, p)
    _, pm = soft_force_and_pe(d - eps, p)
    dU = (pp - pm) / (2 * eps)
    K1 = 15.0 / delta
    expect = -dU / K1 if d > 1.0 - delta else -dU
    assert f == pytest.approx(expect, rel=1e-5, abs=1e-12)
    assert f > 0  # B pushed along +x, away from A


def test_soft_and_hm_momentum_balance():
    rng = np.random.default_rng(4)
    n, side = 250, 9.0
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.5, n)
    h = rng.uniform(0.5, 2, n)
    m = rng.uniform(0.5, 2, n)
    v = rng.normal(size=(n, 3))
    w = rng.normal(size=(n, 3))
    for model in ("soft", "hm"):
        W = O.World(x, q, v, w, r, h, m, np.full(n, 0.5))
        W.set_box([0] * 3, [side] * 3, [1, 1, 1])
        W.set_physics((0, 0, 0), 1e-4)
        if model == "soft":
            nc = W.detect(0, 0.005)
            W.soft_forces(O.soft_params(delta=0.005, k_scaler=1e3, damping=0.3, mu=0.5, nu=1e-3))
        else:
            nc = W.detect(0, 0.0)
            W.hm_forces(O.hm_params(E=1e5, restitution=0.5, mu_k=0.4, mu_s=0.6, v_c=1e-2))
        c = W.contacts(nc)
        st = W.state()
        assert nc > 50
        fsum = st["f"].sum(axis=0)
        assert np.abs(fsum).max() <= 1e-9 * np.abs(st["f"]).max()
    # angular momentum: friction acts at the two closest points, which are
    # dist apart along n, so only the normal (+damping) forces balance torques
    for model in ("soft", "hm"):
        W = O.World(x, q, v, w, r, h, m, np.full(n, 0.5))
        W.set_box([0] * 3, [side] * 3, [0, 0, 0])
        W.set_physics((0, 0, 0), 1e-4)
        if model == "soft":
            W.detect(0, 0.005)
            W.soft_forces(O.soft_params(delta=0.005, k_scaler=1e3, damping=0.3, enable_friction=0))
        else:
            W.detect(0, 0.0)
            W.hm_forces(O.hm_params(E=1e5, restitution=0.5, enable_tangential=0))
        st = W.state()
        L = st["tau"].sum(axis=0) + np.cross(x, st["f"]).sum(axis=0)
        assert np.abs(L).max() <= 1e-9 * (np.abs(st["tau"]).max() + 1)


@pytest.mark.parametrize("law,power", [(0, 2.0), (1, 1.5)])
def test_hertz_normal_force_closed_form(law, power):
    E, nu = 1e6, 0.25
    Es = E / (2 * (1 - nu * nu))
    Rs = 0.25  # r = 0.5 for both rods
    for delta in (1e-3, 4e-3, 1.6e-2):
        W = two_crossed(1.0 - delta)
        assert W.detect(0, 0.0) == 1
        W.hm_forces(O.hm_params(E=E, nu_poisson=nu, restitution=0.99999999, enable_tangential=0, law=law))
        fb = W.state()["f"][1, 0]
        assert fb == pytest.approx(4.0 / 3.0 * Es * math.sqrt(Rs) * delta ** power, rel=1e-9)


def test_mindlin_spring_and_coulomb_cap():
    E, nu = 1e6, 0.25
    Gs = E / (2 * (1 + nu))
    delta, dt, vt = 4e-3, 1e-4, 0.01
    W = two_crossed(1.0 - delta, vb=(0, vt, 0))
    W.set_physics((0, 0, 0), dt)
    p = O.hm_params(E=E, nu_poisson=nu, restitution=0.99999999, mu_k=10.0, mu_s=10.0, v_c=1.0, law=1,
                    enable_tangential=1)
    W.detect(0, 0.0)
    W.hm_forces(p)
    kt = 8 * Gs * math.sqrt(0.25) * math.sqrt(delta)
    fy = W.state()["f"][1, 1]
    assert fy == pytest.approx(-kt * vt * dt, rel=1e-6)  # elastic spring after one call (damping ~ 0)
    # second call without clearing: xi doubles
    W.set_forces(np.zeros((2, 3)), np.zeros((2, 3)))
    W.detect(0, 0.0)
    W.hm_forces(p)
    assert W.state()["f"][1, 1] == pytest.approx(-2 * kt * vt * dt, rel=1e-6)
    # Coulomb cap: tiny mu -> |F_t| = mu |F_n|
    W = two_crossed(1.0 - delta, vb=(0, 5.0, 0))
    p2 = O.hm_params(E=E, nu_poisson=nu, restitution=0.99999999, mu_k=0.1, mu_s=0.1, v_c=1.0, law=1)
    W.detect(0, 0.0)
    W.hm_forces(p2)
    f = W.state()["f"][1]
    assert abs(f[1]) == pytest.approx(0.1 * f[0], rel=1e-9)


def test_hm_history_pruned_after_four_frames():
    W = two_crossed(0.99)
    p = O.hm_params(E=1e5, restitution=0.5)
    W.detect(0, 0.0)
    W.hm_forces(p)
    assert W.hm_history_size() == 1
    W.set_state(x=np.array([[0.0, 0, 0], [5.0, 0, 0]]))  # separated
    for k in range(4):
        W.detect(0, 0.0)
        W.hm_forces(p)
        assert W.hm_history_size() == 1, k
    W.detect(0, 0.0)
    W.hm_forces(p)
    assert W.hm_history_size() == 0


def single_rod(v, w, q=(1.0, 0, 0, 0)):
    W = O.World(np.array([[1.0, 2.0, 3.0]]), np.array([q]), np.array([v]), np.array([w]), np.array([0.5]),
                np.array([5.0]), np.array([2.0]), np.array([0.0]))
    W.set_box([-100] * 3, [100] * 3, [0, 0, 0])
    return W


def test_free_flight_exact_and_quaternion_unit():
    dt = 1e-3
    W = single_rod([0.3, -0.2, 0.1], [0.0, 0.0, 0.7])
    W.set_physics((0, 0, -1.0), dt)
    for _ in range(1000):
        W.integrate(0)
        W.integrate(1)
    st = W.state()
    t = 1000 * dt
    # symplectic Euler: x_n = x0 + v0 t - g dt^2 n(n+1)/2
    assert st["x"][0, 2] == pytest.approx(3.0 + 0.1 * t - dt * dt * 1000 * 1001 / 2, abs=1e-12)
    assert st["x"][0, 0] == pytest.approx(1.0 + 0.3 * t, abs=1e-12)
    assert np.linalg.norm(st["q"][0]) == pytest.approx(1.0, abs=1e-15)
    # rotation about z (perpendicular to the rod) by |w| t: first-order update, O(dt) angle error
    u = O.axis(st["q"][0])
    ang = math.atan2(-u[0], u[1])
    assert ang == pytest.approx(0.7 * t, abs=1e-3)


def test_verlet_exp_map_rotation_exact():
    dt = 1e-3
    W = single_rod([0, 0, 0], [0.0, 0.0, 0.7])
    W.set_physics((0, 0, 0), dt)
    for _ in range(2000):
        W.integrate(2)
        W.integrate(3)
    u = O.axis(W.state()["q"][0])
    assert math.atan2(-u[0], u[1]) == pytest.approx(0.7 * 2000 * dt, abs=1e-12)


def test_torque_free_rod_conserves_energy_and_angular_momentum():
    # spin with an axial component: the rod axis precesses about L
    m, r, Lr = 2.0, 0.5, 10.0
    Ix, Iy = m * (3 * r * r + Lr * Lr) / 12, m * r * r / 2

    def ang_mom(st):
        u = O.axis(st["q"][0])
        w = st["w"][0]
        return Ix * w + (Iy - Ix) * u * (u @ w)

    drift = []
    for dt, nsteps in ((2e-4, 2500), (1e-4, 5000)):
        W = single_rod([0, 0, 0], [0.3, 0.9, 0.2])
        W.set_physics((0, 0, 0), dt)
        ke0, L0 = W.kinetic_energy(), ang_mom(W.state())
        for _ in range(nsteps):
            W.integrate(2)
            W.integrate(3)
        ke1, L1 = W.kinetic_energy(), ang_mom(W.state())
        assert ke1 == pytest.approx(ke0, rel=1e-4)
        assert np.allclose(L1, L0, rtol=0, atol=1e-3 * np.linalg.norm(L0))
        drift.append(abs(ke1 - ke0))
    # the explicit gyroscopic half kicks drift at first order in dt
    assert drift[1] < 0.7 * drift[0]


def test_kinetic_energy_closed_form():
    q = np.array([math.cos(0.3), 0, 0, math.sin(0.3)])
    W = single_rod([1.0, 2.0, 0.0], [0.5, -0.1, 0.25], q)
    m, r, L = 2.0, 0.5, 10.0
    Ix, Iy = m * (3 * r * r + L * L) / 12, m * r * r / 2
    u = O.axis(q)
    w = np.array([0.5, -0.1, 0.25])
    wpar = w @ u
    ke = 0.5 * m * 5.0 + 0.5 * (Iy * wpar ** 2 + Ix * (w @ w - wpar ** 2))
    assert W.kinetic_energy() == pytest.approx(ke, rel=1e-13)


def test_soft_collision_energy_error_vanishes_with_dt():
    # head-on side collision of crossed rods, no damping or friction.  The
    # reference force jumps at d = h - delta (the barrier gradient is dU/dd / K1),
    # so Verlet's energy error comes from that jump and shrinks with dt.
    errs = []
    for dt in (1e-4, 2.5e-5):
        x = np.array([[0.0, 0, 0], [1.2, 0.0, 0]])
        q = np.array([[1.0, 0, 0, 0], [math.sqrt(0.5), math.sqrt(0.5), 0, 0]])
        v = np.array([[0.1, 0, 0], [-0.1, 0, 0]])
        W = O.World(x, q, v, np.zeros((2, 3)), np.full(2, 0.5), np.full(2, 2.0), np.ones(2), np.zeros(2))
        W.set_box([-50] * 3, [50] * 3, [0, 0, 0])
        W.set_physics((0, 0, 0), dt)
        p = O.soft_params(delta=0.05, k_scaler=100.0, damping=0.0, enable_friction=0)
        ke0 = W.kinetic_energy()
        W.vv_step(0, p, int(4.0 / dt), method=0)
        st = W.state()
        assert st["v"][0, 0] < 0 < st["v"][1, 0]  # bounced
        assert st["v"][0, 0] == pytest.approx(-st["v"][1, 0], rel=1e-12)  # momentum
        errs.append(abs(W.kinetic_energy() / ke0 - 1))
    assert errs[1] < 2e-3 and errs[1] < errs[0] / 4
