"""Closed-form pins of the oracle's soft friction laws and Hertz-Mindlin
rolling torque (CPU only).

Fixture: rod A along y at the origin, rod B crossing it with its centre on
the common normal, so both closest points are the rod centres (no lever arm:
the body forces are the contact forces and the body torques are the contact
torques).  The soft normal force acts along n = A -> B; friction is tangent.

* Smooth model (soft_contact.cpp:458-485): gamma = 2/(1+exp(-K2 v)) - 1 =
  tanh(K2 v / 2) for v <= nu, 1 above; Stribeck mu_eff = mu + (mu_s - mu)
  exp(-v/nu); |F_t| = mu_eff gamma |F_n|, opposing the relative tangential
  velocity of A.
* Cundall-Strack (soft_contact.cpp:317-367): F_t <- P_n(F_t - kt dt v_t),
  capped at mu |F_n|, carried between calls per pair.
* Karnopp (soft_contact.cpp:369-456): below the dead band the hold force
  F = -m_eff g_t - m_eff v_t / dt, scaled to mu_s |F_n| when larger; above it
  mu |F_n| against v_t.
* Rolling (hertz_mindlin.cpp:356-383): tau = -mu_r R* |F_n| w_rel / |w_rel|,
  w_rel = w_a - w_b, half on A and minus half on B.

Each closed form is also checked against a plausible mistake (wrong split,
sign, radius, blend or accumulation), which must NOT match.
"""
import math

import numpy as np
import pytest

import oracle_py as O

C45 = math.sqrt(0.5)


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def crossed(gap, phi=0.0, va=(0, 0, 0), vb=(0, 0, 0), wa=(0, 0, 0), wb=(0, 0, 0), rA=0.5, rB=0.5,
            mA=1.0, mB=2.0, g=(0, 0, 0), dt=1e-3):
    """A along y at the origin; B's axis is z turned by phi about y, i.e.
    (sin phi, 0, cos phi), centred at gap * n with n = (cos phi, 0, -sin phi)
    (the common normal of the two axes)."""
    n = np.array([math.cos(phi), 0.0, -math.sin(phi)])
    qb = qmul([math.cos(phi / 2), 0, math.sin(phi / 2), 0], [C45, C45, 0, 0])
    x = np.array([[0.0, 0, 0], gap * n])
    q = np.array([[1.0, 0, 0, 0], qb])
    W = O.World(x, q, np.array([va, vb], float), np.array([wa, wb], float), np.array([rA, rB]),
                np.array([2.0, 2.0]), np.array([mA, mB]), np.array([0.5, 0.5]))
    W.set_box([-10] * 3, [10] * 3, [0, 0, 0])
    W.set_physics(g, dt)
    return W, n


def soft_eval(W, p):
    W.set_forces(np.zeros((2, 3)), np.zeros((2, 3)))
    assert W.detect(0, p.delta) == 1
    W.soft_forces(p)
    st = W.state()
    c = W.contacts(1)
    return st, c


def split(fA, n):
    """normal magnitude and tangential vector of the force on A"""
    fn = -float(fA @ n)  # A is pushed against n
    return fn, fA + fn * n


# ---------------------------------------------------------------- smooth / Stribeck
@pytest.mark.parametrize("frac", [1e-6, 1e-3, 0.1, 0.5, 0.999, 1.0, 1.001, 3.0, 50.0])
def test_smooth_gamma_is_tanh_and_saturates(frac):
    nu, mu = 1e-3, 0.4
    v = frac * nu
    W, n = crossed(0.99, vb=(0, v, 0))
    p = O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, mu_static=0.0, nu=nu, friction_model=0)
    st, c = soft_eval(W, p)
    assert np.allclose(c["n"][0], n, atol=1e-15)
    fn, ft = split(st["f"][0], n)
    assert fn > 0
    gamma = math.tanh(15.0 / nu * v / 2) if v <= nu else 1.0
    assert np.linalg.norm(ft) == pytest.approx(mu * gamma * fn, rel=1e-12)
    # A's relative velocity at the contact is -v along y: friction on A is +y
    assert ft[1] > 0 and abs(ft[0]) + abs(ft[2]) <= 1e-15 * fn
    # B gets the opposite force
    assert np.allclose(st["f"][1] + st["f"][0], 0, atol=1e-12 * fn)
    if frac <= 1e-3:
        # gamma -> 0 linearly: |F_t| ~ mu (K2 v / 2) |F_n|, not mu |F_n|
        assert np.linalg.norm(ft) == pytest.approx(mu * 7.5 * frac * fn, rel=1e-4)
        assert np.linalg.norm(ft) < 1e-2 * mu * fn


def test_smooth_gamma_continuous_at_nu():
    nu, mu = 1e-3, 0.4
    out = []
    for v in (nu * (1 - 1e-9), nu * (1 + 1e-9)):
        W, n = crossed(0.99, vb=(0, v, 0))
        st, _ = soft_eval(W, O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, mu_static=0.0, nu=nu))
        fn, ft = split(st["f"][0], n)
        out.append(np.linalg.norm(ft) / (mu * fn))
    # below nu: 1 - 2 e^-15 / (1 + e^-15) = 1 - 6.1e-7; above: exactly 1
    assert out[1] == 1.0
    assert out[0] == pytest.approx(1 - 2 * math.exp(-15) / (1 + math.exp(-15)), rel=1e-12)
    assert abs(out[1] - out[0]) < 1e-6


@pytest.mark.parametrize("frac", [1e-7, 0.3, 1.0, 40.0])
def test_stribeck_blend(frac):
    nu, mu, mus = 1e-3, 0.3, 0.8
    v = frac * nu
    W, n = crossed(0.99, vb=(0, v, 0))
    st, _ = soft_eval(W, O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, mu_static=mus, nu=nu))
    fn, ft = split(st["f"][0], n)
    gamma = math.tanh(7.5 * frac) if v <= nu else 1.0
    mu_eff = mu + (mus - mu) * math.exp(-frac)
    ratio = np.linalg.norm(ft) / (gamma * fn)
    assert ratio == pytest.approx(mu_eff, rel=1e-9)
    if frac < 1e-6:
        assert ratio == pytest.approx(mus, rel=1e-6)  # static coefficient at v -> 0
    if frac > 30:
        assert ratio == pytest.approx(mu, rel=1e-12)  # kinetic coefficient for v >> nu
    # a blend that decays the wrong way (mu_s at high speed) does not match
    wrong = mus + (mu - mus) * math.exp(-frac)
    assert abs(ratio - wrong) > 1e-3 or abs(mu_eff - wrong) < 1e-12


def test_friction_within_coulomb_bound_and_opposes_sliding():
    rng = np.random.default_rng(11)
    nu, mu, mus = 1e-2, 0.35, 0.6
    checked = 0
    for trial in range(200):
        phi = rng.uniform(-1.2, 1.2)
        gap = rng.uniform(0.96, 1.004)
        va, vb = rng.normal(size=3) * 10 ** rng.uniform(-4, 0), rng.normal(size=3) * 10 ** rng.uniform(-4, 0)
        wa, wb = rng.normal(size=3) * 0.01, rng.normal(size=3) * 0.01
        W, n = crossed(gap, phi, va=va, vb=vb, wa=wa, wb=wb)
        p = O.soft_params(delta=0.005, k_scaler=1e3, damping=0.2, mu=mu, mu_static=mus, nu=nu)
        W.set_forces(np.zeros((2, 3)), np.zeros((2, 3)))
        if W.detect(0, p.delta) != 1:
            continue
        W.soft_forces(p)
        st = W.state()
        c = W.contacts(1)
        nn = c["n"][0]
        fA = st["f"][0]
        ft = fA - (fA @ nn) * nn
        fn = abs(fA @ nn)
        # relative velocity of A at its contact point (both contact points are centres here)
        ra = c["pa"][0] - 0.0
        rb = c["pb"][0] - (W.state()["x"][1] + c["shift"][0])
        vrel = (va + np.cross(wa, ra)) - (vb + np.cross(wb, rb))
        vt = vrel - (vrel @ nn) * nn
        assert np.linalg.norm(ft) <= max(mu, mus) * fn * (1 + 1e-12) + 1e-300
        if np.linalg.norm(vt) > 1e-9 and np.linalg.norm(ft) > 0:
            cosang = (ft @ vt) / (np.linalg.norm(ft) * np.linalg.norm(vt))
            assert cosang == pytest.approx(-1.0, abs=1e-9)
        checked += 1
    assert checked > 100


# ---------------------------------------------------------------- Cundall-Strack
def test_cundall_accumulates_projects_and_caps():
    kt, dt, mu = 50.0, 1e-3, 0.5
    vb = np.array([0.0, 0.3, 0.2])
    p = O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, kt=kt, friction_model=1)
    # three calls on one pair (the history is keyed by the pair and kept in the
    # world), the normal turning between calls: F <- P_n(F - kt dt v_t)
    W, n0 = crossed(0.99, 0.0, vb=vb, dt=dt)
    expect = np.zeros(3)
    naive = np.zeros(3)
    for k, phi in enumerate((0.0, 0.25, 0.55)):
        if k > 0:
            _, n_ = crossed(0.99, phi)
            qb = qmul([math.cos(phi / 2), 0, math.sin(phi / 2), 0], [C45, C45, 0, 0])
            W.set_state(x=np.array([[0.0, 0, 0], 0.99 * n_]), q=np.array([[1.0, 0, 0, 0], qb]))
        st, c = soft_eval(W, p)
        n = c["n"][0]
        fn = -float(st["f"][0] @ n)
        vrel = -vb  # A at rest, no spin, contact points at the centres
        vt = vrel - (vrel @ n) * n
        expect = expect - kt * dt * vt
        expect = expect - (expect @ n) * n
        naive = naive - kt * dt * vt
        assert np.linalg.norm(expect) < mu * fn  # below the cap on these calls
        ftA = st["f"][0] - (st["f"][0] @ n) * n
        assert np.allclose(ftA, expect, rtol=0, atol=1e-12 * np.linalg.norm(expect))
        if k == 2:
            # without the tangent-plane projection the spring would not match
            assert np.linalg.norm(naive - expect) > 1e-3 * np.linalg.norm(expect)
    # cap: a stiff spring saturates at mu |F_n| against v_t on the first call
    W, n = crossed(0.99, 0.0, vb=vb, dt=dt)
    st, _ = soft_eval(W, O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, kt=1e6, friction_model=1))
    fn, ft = split(st["f"][0], n)
    vt = -vb - (-vb @ n) * n
    assert np.allclose(ft, -mu * fn * vt / np.linalg.norm(vt), atol=1e-12 * fn)
    # and the capped value is what the next call starts from: with v_t reversed
    # the spring unloads by kt dt |v_t| from the cap
    W.set_state(v=np.array([[0.0, 0, 0], -vb]))
    st, _ = soft_eval(W, O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, kt=1e6, friction_model=1))
    fn2, ft2 = split(st["f"][0], n)
    unloaded = -mu * fn * vt / np.linalg.norm(vt) + 1e6 * dt * vt
    want = unloaded if np.linalg.norm(unloaded) <= mu * fn2 else unloaded * (mu * fn2 / np.linalg.norm(unloaded))
    assert np.allclose(ft2, want, atol=1e-10 * fn2)


def test_cundall_three_calls_linear_accumulation():
    kt, dt = 20.0, 1e-3
    vb = np.array([0.0, 0.1, 0.0])
    W, n = crossed(0.99, vb=vb, dt=dt)
    p = O.soft_params(delta=0.005, k_scaler=1e3, mu=0.5, kt=kt, friction_model=1)
    for k in range(1, 4):
        st, _ = soft_eval(W, p)
        fn, ft = split(st["f"][0], n)
        # v_t of A is -vb: spring force on A after k calls is +k kt dt |vb| along y
        assert ft[1] == pytest.approx(k * kt * dt * 0.1, rel=1e-12)
        assert k * kt * dt * 0.1 < 0.5 * fn
        # the rate-independent mistake (kt v_t without dt) would be 1000x larger
        assert abs(ft[1] - k * kt * 0.1) > 1.0


# ---------------------------------------------------------------- Karnopp
def test_karnopp_stick_hold_force():
    dt, mus, mu = 1e-3, 0.9, 0.4
    g = np.array([0.3, -0.8, -1.0])  # tangential part (0, g_y, g_z) for n = x
    vb = np.array([0.0, 2e-4, -1e-4])
    W, n = crossed(0.99, vb=vb, g=g, dt=dt)
    p = O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, mu_static=mus, vel_deadband=1e-3, friction_model=2)
    st, _ = soft_eval(W, p)
    fn, ft = split(st["f"][0], n)
    me = 1.0 * 2.0 / 3.0
    gt = g - (g @ n) * n
    vt = -vb - (-vb @ n) * n
    hold = -me * gt - me * vt / dt
    assert np.linalg.norm(hold) <= mus * fn
    assert np.allclose(ft, hold, atol=1e-12 * np.linalg.norm(hold))
    # mistakes that must not match: +F_drive, or the sum of the masses
    assert np.linalg.norm(ft - (me * gt - me * vt / dt)) > 1e-3
    assert np.linalg.norm(ft - (-3.0 * gt - 3.0 * vt / dt)) > 1e-3


def test_karnopp_breakaway_and_slip():
    dt, mus, mu = 1e-3, 0.05, 0.04
    g = np.array([0.0, -50.0, 0.0])
    vb = np.array([0.0, 2e-4, 0.0])
    W, n = crossed(0.99, vb=vb, g=g, dt=dt)
    p = O.soft_params(delta=0.005, k_scaler=1e3, mu=mu, mu_static=mus, vel_deadband=1e-3, friction_model=2)
    st, _ = soft_eval(W, p)
    fn, ft = split(st["f"][0], n)
    me = 2.0 / 3.0
    hold = -me * (g - (g @ n) * n) - me * (-vb) / dt
    assert np.linalg.norm(hold) > mus * fn
    assert np.allclose(ft, hold * (mus * fn / np.linalg.norm(hold)), atol=1e-12 * fn)
    # above the dead band: kinetic mu against v_t
    vb2 = np.array([0.0, 0.05, 0.0])
    W, n = crossed(0.99, vb=vb2, g=g, dt=dt)
    st, _ = soft_eval(W, p)
    fn, ft = split(st["f"][0], n)
    assert np.allclose(ft, [0.0, mu * fn, 0.0], atol=1e-13 * fn)


# ---------------------------------------------------------------- HM rolling
@pytest.mark.parametrize("wa,wb", [((0.3, -0.2, 0.7), (0, 0, 0)), ((0.1, 0.4, 0.0), (-0.5, 0.2, 0.3))])
def test_rolling_torque_closed_form(wa, wb):
    E, nu, mur, rA, rB = 1e6, 0.25, 0.07, 0.5, 0.3
    delta = 4e-3
    W, n = crossed(rA + rB - delta, wa=wa, wb=wb, rA=rA, rB=rB)
    assert W.detect(0, 0.0) == 1
    p = O.hm_params(E=E, nu_poisson=nu, restitution=0.99999999, enable_tangential=0, enable_rolling=1,
                    mu_roll=mur, law=0)
    W.hm_forces(p)
    st = W.state()
    fn = np.linalg.norm(st["f"][1])
    Es = E / (2 * (1 - nu * nu))
    Rs = rA * rB / (rA + rB)
    assert fn == pytest.approx(4.0 / 3.0 * Es * math.sqrt(Rs) * delta ** 2, rel=1e-6)
    wr = np.array(wa, float) - np.array(wb, float)
    tau = -mur * Rs * fn * wr / np.linalg.norm(wr)
    # contact points are the centres: the body torques are the rolling torques
    assert np.allclose(st["tau"][0], 0.5 * tau, rtol=0, atol=1e-12 * np.linalg.norm(tau))
    assert np.allclose(st["tau"][1], -0.5 * tau, rtol=0, atol=1e-12 * np.linalg.norm(tau))
    # plausible mistakes: the whole torque on each body, the opposite sense,
    # the radius of one rod instead of R*
    for wrong in (tau, -0.5 * tau, 0.5 * tau * rA / Rs):
        assert np.linalg.norm(st["tau"][0] - wrong) > 1e-3 * np.linalg.norm(tau)


def test_rolling_zero_for_equal_spins():
    W, n = crossed(0.996, wa=(0.2, 0.1, 0.0), wb=(0.2, 0.1, 0.0))
    W.detect(0, 0.0)
    W.hm_forces(O.hm_params(E=1e6, restitution=0.99999999, enable_tangential=0, enable_rolling=1, mu_roll=0.1))
    assert np.all(W.state()["tau"] == 0.0)
