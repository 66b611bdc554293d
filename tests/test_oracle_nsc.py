"""Pins of the oracle's NSC path (CPU only).

Reference fixtures: smoke_test.py:165-204 (exact elastic exchange of crossed
rods, and the Table S1 cfm=0.05 bound).  Closed forms: Delassus diagonal
from an independent inertia computation, frictionless LCP solved by brute
force over active sets.
"""
import itertools
import math

import numpy as np
import pytest

import oracle_py as O


def crossed_pair(vx=0.1):
    c = math.sqrt(0.5)
    r, L, rho = 0.01, 1.0, 1000.0
    m = rho * math.pi * r * r * L  # cylinder-only mass (rigid_body.cpp:19-24)
    x = np.array([[-0.06, 0, 0], [0.06, 0, 0]])
    q = np.array([[1.0, 0, 0, 0], [c, c, 0, 0]])
    v = np.array([[vx, 0, 0], [0, 0, 0]])
    W = O.World(x, q, v, np.zeros((2, 3)), np.full(2, r), np.full(2, L / 2), np.full(2, m), np.ones(2))
    W.set_box([-10] * 3, [10] * 3, [0, 0, 0])
    W.set_physics((0, 0, 0), 1e-3, 0.0, 0.0, 80.0)
    return W


def test_elastic_exchange_reference_smoke():
    # smoke_test.py:165-195: e=1, mu=0, beta=0, cfm=0, 200 iters, 3000 steps
    W = crossed_pair()
    p = O.nsc_params(mu=0.0, beta=0.0, cfm=0.0, omega=1.0, max_sweeps=200, stop_rule=0)
    W.nsc_step(p, 3000, method=0)
    v = W.state()["v"]
    assert abs(v[0, 0]) < 1e-3 and abs(v[1, 0] - 0.1) < 1e-3
    # with the build-defined residual stop rule the exchange is exact to 1e-12
    W = crossed_pair()
    p = O.nsc_params(mu=0.0, beta=0.0, cfm=0.0, omega=1.0, tol=1e-12, stop_rule=1)
    W.nsc_step(p, 3000, method=1)
    v = W.state()["v"]
    assert abs(v[0, 0]) < 1e-12 and abs(v[1, 0] - 0.1) < 1e-12
    ke = 0.5 * (v ** 2).sum()
    assert ke == pytest.approx(0.5 * 0.01, rel=1e-10)


def test_table_s1_cfm_bound_reference_smoke():
    # smoke_test.py:197-204: cfm=0.05 stays near-elastic for thin rods
    W = crossed_pair()
    p = O.nsc_params(mu=0.0, beta=0.0, cfm=0.05, omega=1.0, max_sweeps=200, stop_rule=0)
    W.nsc_step(p, 3000, method=0)
    v = W.state()["v"]
    assert v[1, 0] > 0.09 and (v[0, 0] ** 2 + v[1, 0] ** 2) / 0.01 > 0.95


def random_bed(n=300, side=9.0, seed=3, mu=0.5):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.5, n)
    h = rng.uniform(0.5, 2.0, n)
    m = rng.uniform(0.5, 2.0, n)
    v = rng.normal(size=(n, 3))
    w = rng.normal(size=(n, 3))
    W = O.World(x, q, v, w, r, h, m, np.full(n, 0.2))
    W.set_box([0] * 3, [side] * 3, [1, 1, 0], 1, 1)
    W.set_physics((0, 0, -1), 1e-3)
    return W, dict(x=x, q=q, v=v, w=w, r=r, h=h, m=m)


def rot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_manifold_fields_against_independent_inertia():
    W, B = random_bed()
    W.detect(0, 0.0)
    p = O.nsc_params(mu=0.5, cfm=0.01, warm_start=0)
    m = W.nsc_build(p)
    rec, ab, lam = W.manifolds()
    assert m > 50 and (ab[:, 1] < 0).any()
    for i in range(m):
        a, b = ab[i]
        n, t1, t2, ra, rb = (rec[i, 3 * k:3 * k + 3] for k in range(5))
        # orthonormal contact frame
        F = np.stack([n, t1, t2])
        assert np.allclose(F @ F.T, np.eye(3), atol=1e-12)
        # body-frame inertia of a solid cylinder about local +Y, rotated to the world
        def Iinv(j):
            L = 2 * B["h"][j]
            Ix = B["m"][j] * (3 * B["r"][j] ** 2 + L * L) / 12
            Iy = B["m"][j] * B["r"][j] ** 2 / 2
            R = rot(B["q"][j])
            return R @ np.diag([1 / Ix, 1 / Iy, 1 / Ix]) @ R.T
        IA = Iinv(a)
        for k, d in enumerate((n, t1, t2)):
            assert np.allclose(rec[i, 15 + 3 * k:18 + 3 * k], IA @ np.cross(ra, d), rtol=1e-12, atol=1e-13)
            g = 1 / B["m"][a] + np.cross(ra, d) @ IA @ np.cross(ra, d) + 0.01
            if b >= 0:
                IB = Iinv(b)
                assert np.allclose(rec[i, 24 + 3 * k:27 + 3 * k], IB @ np.cross(rb, d), rtol=1e-12, atol=1e-13)
                g += 1 / B["m"][b] + np.cross(rb, d) @ IB @ np.cross(rb, d)
            assert rec[i, 35 + k] == pytest.approx(g, rel=1e-12)
        # pre-solve normal velocity
        va = B["v"][a] + np.cross(B["w"][a], ra)
        vb = (B["v"][b] + np.cross(B["w"][b], rb)) if b >= 0 else np.zeros(3)
        assert rec[i, 39] == pytest.approx(n @ (vb - va), rel=1e-12, abs=1e-13)


def delassus(rec, ab, nbody, masses):
    """Normal-normal block of J M^-1 J^T from the manifold records."""
    m = len(ab)
    A = np.zeros((m, m))
    for i in range(m):
        for j in range(m):
            ai, bi = ab[i]
            aj, bj = ab[j]
            ni, nj = rec[i, 0:3], rec[j, 0:3]
            rai, rbi = rec[i, 9:12], rec[i, 12:15]
            s = 0.0
            # impulse +lambda_j on B_j, -lambda_j on A_j; measure w_i
            for body_j, sign_j, Iv in ((aj, -1.0, rec[j, 15:18]), (bj, 1.0, rec[j, 24:27])):
                if body_j < 0:
                    continue
                dv = sign_j * nj / masses[body_j]
                dw = sign_j * Iv
                for body_i, sign_i, r in ((ai, -1.0, rai), (bi, 1.0, rbi)):
                    if body_i == body_j:
                        s += sign_i * ni @ (dv + np.cross(dw, r))
            A[i, j] = s
    return A


def test_frictionless_lcp_equals_brute_force_solution():
    # a small cluster: mu = 0, so only the normal LCP remains
    rng = np.random.default_rng(12)
    for trial in range(5):
        n = 7
        x = rng.uniform(0, 2.2, size=(n, 3))
        q = rng.normal(size=(n, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        r = np.full(n, 0.45)
        h = np.full(n, 0.6)
        mass = rng.uniform(0.5, 2, n)
        v = rng.normal(size=(n, 3))
        w = rng.normal(size=(n, 3))
        W = O.World(x, q, v, w, r, h, mass, np.full(n, 0.0))
        W.set_box([-5] * 3, [5] * 3, [0, 0, 0])
        W.set_physics((0, 0, 0), 1e-3)
        nc = W.detect(0, 0.0)
        if nc < 2 or nc > 10:
            continue
        p = O.nsc_params(mu=0.0, beta=0.0, cfm=1e-6, omega=1.0, tol=1e-13, max_sweeps=200000, warm_start=0)
        W.nsc_build(p)
        rec, ab, _ = W.manifolds()
        A = delassus(rec, ab, n, mass) + 1e-6 * np.eye(len(ab))
        bvec = rec[:, 39].copy()  # e = 0 and beta = 0: w = A lambda + v_n_pre
        # the diagonal stored in the record is the Delassus diagonal + cfm
        assert np.allclose(np.diag(A), rec[:, 35], rtol=1e-12)
        sol = None
        for act in itertools.product([0, 1], repeat=len(ab)):
            S = [i for i in range(len(ab)) if act[i]]
            lam = np.zeros(len(ab))
            if S:
                lam[S] = np.linalg.solve(A[np.ix_(S, S)], -bvec[S])
            wv = A @ lam + bvec
            if lam.min() >= -1e-12 and wv.min() >= -1e-10:
                sol = lam
                break
        assert sol is not None
        r_, s_ = W.solve(p)
        assert r_ <= 1e-13
        _, _, lam = W.manifolds()
        assert np.allclose(lam[:, 0], sol, rtol=1e-8, atol=1e-12), (lam[:, 0], sol)


@pytest.mark.parametrize("proj", [0, 1])
def test_solve_converges_and_satisfies_friction_bounds(proj):
    W, B = random_bed(seed=9)
    W.detect(0, 0.0)
    if proj == 1:
        # Euclidean disc: the projected Gauss-Seidel map converges to the fixed point
        p = O.nsc_params(mu=0.5, beta=0.0, cfm=0.0, omega=1.3, tol=1e-10, friction_proj=1, warm_start=0)
        W.nsc_build(p)
        r, s = W.solve(p)
        assert r <= 1e-10 and s > 1
    else:
        # the reference's sequential clamp (nsc_contact.cpp:503-533) is not a
        # contraction on dense sliding beds; run it with the reference stop rule
        p = O.nsc_params(mu=0.5, beta=0.1, cfm=0.0, omega=1.3, max_sweeps=200, friction_proj=0, stop_rule=0,
                         warm_start=0)
        W.nsc_build(p)
        r, s = W.solve(p)
        assert s == 200 and np.isfinite(r)
    _, _, lam = W.manifolds()
    assert lam[:, 0].min() >= 0
    if proj == 0:
        assert np.all(np.abs(lam[:, 1]) <= 0.5 * lam[:, 0] * (1 + 1e-12) + 1e-15)
        assert np.all(np.abs(lam[:, 2]) <= 0.5 * lam[:, 0] * (1 + 1e-12) + 1e-15)
    else:
        assert np.all(np.hypot(lam[:, 1], lam[:, 2]) <= 0.5 * lam[:, 0] * (1 + 1e-12) + 1e-15)
    if proj == 0:
        return
    # one more sweep from a converged state changes lambda by at most ~tol/g
    _, _, l0 = W.manifolds()
    W.sweeps(p, 1)
    _, _, l1 = W.manifolds()
    rec, _, _ = W.manifolds()
    assert np.max(np.abs(l1 - l0) * rec[:, 35:38].min(axis=1, keepdims=True)) < 1e-9


def test_residual_detects_planted_error():
    # a wrong sign in one impulse must show up in the residual
    W, B = random_bed(seed=17)
    W.detect(0, 0.0)
    p = O.nsc_params(mu=0.5, beta=0.0, omega=1.0, tol=1e-12, friction_proj=1, warm_start=0, max_sweeps=50000)
    W.nsc_build(p)
    r, _ = W.solve(p)
    assert r <= 1e-12
    rec, ab, lam = W.manifolds()
    i = int(np.argmax(lam[:, 0]))
    lam2 = lam.copy()
    lam2[i, 0] *= 1.01
    # apply the impulse change so that velocities stay consistent with lambda
    W.set_manifolds(rec, ab, lam2 - lam)
    W.apply_warm()
    W.set_manifolds(rec, ab, lam2)
    assert W.residual(p) > 1e-6
    # lambda alone changed (velocities not): invisible to the fixed-point map
    # when cfm = 0, which is why the GPU tests also check v = v0 + M^-1 J^T lambda
    W.set_manifolds(rec, ab, -(lam2 - lam))
    W.apply_warm()
    W.set_manifolds(rec, ab, lam)
    assert W.residual(p) <= 1e-11


def test_warm_start_reduces_sweeps_on_a_resting_bed():
    W, B = random_bed(n=200, side=8.0, seed=23, mu=0.5)
    W.set_state(v=np.zeros((200, 3)), w=np.zeros((200, 3)))
    cold = O.nsc_params(mu=0.5, omega=1.0, tol=1e-8, friction_proj=1, warm_start=0)
    warm = O.nsc_params(mu=0.5, omega=1.0, tol=1e-8, friction_proj=1, warm_start=1)
    s_cold, _ = W.nsc_step(cold, 3, method=0)
    W2, _ = random_bed(n=200, side=8.0, seed=23, mu=0.5)
    W2.set_state(v=np.zeros((200, 3)), w=np.zeros((200, 3)))
    s_warm, _ = W2.nsc_step(warm, 3, method=0)
    assert s_warm < s_cold


def test_overrelaxed_psor_limit_cycle_fixture():
    """Captured case: with mu = 0.977 and restitution 0.5, PSOR at omega = 1.3
    settles on a limit cycle (residual 0.242) while omega = 1 converges; the
    stall rule's omega fallback exists for this."""
    import os

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "c4_limit_cycle.npz"))
    W = O.World(d["x"], d["q"], d["v"], d["w"], d["r"], d["h"], d["mass"], d["rest"])
    W.set_box(list(d["lo"]), list(d["hi"]), [1, 1, 1], 0, 0)
    W.set_physics((0, 0, 0), 1e-3)
    assert W.detect(0, 0.0) == 11
    assert W.nsc_build(O.nsc_params(mu=float(d["mu"]), warm_start=0)) == 11
    rec, ab, _ = W.manifolds()
    assert np.array_equal(ab, d["ab"])
    res = {}
    for om in (1.3, 1.0):
        op = O.nsc_params(mu=float(d["mu"]), beta=0.0, cfm=0.0, omega=om, friction_proj=1)
        W.set_state(v=d["v"], w=d["w"])
        W.set_manifolds(rec, ab, d["lam0"])
        W.apply_warm()
        W.sweeps(op, 2000, None)
        res[om] = W.residual(op)
    assert res[1.3] > 0.1 and res[1.0] < 1e-12, res


def _crossing_pair(gap):
    """Rod A along x at the origin, rod B along y above it; the contact is at
    both rod centres, so the correction is a pure translation."""
    import math

    c = math.cos(math.pi / 4)
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0 + gap]])
    q = np.array([[c, 0.0, 0.0, c], [1.0, 0.0, 0.0, 0.0]])  # A's axis (local +Y) rotated onto x
    W = O.World(x, q, np.zeros((2, 3)), np.zeros((2, 3)), np.full(2, 0.5), np.full(2, 2.0), np.array([1.0, 3.0]))
    W.set_box([-50] * 3, [50] * 3, [0, 0, 0])
    return W


def test_position_projection_single_contact_closed_form():
    # phi = -0.05: one pass, omega = 1, cfm = 0 moves the pair apart to phi = -slop
    W = _crossing_pair(-0.05)
    assert W.detect(0, 0.0) == 1
    p = O.nsc_params(omega=1.0, cfm=0.0, position_stabilization=1, position_iters=1, position_psor_iters=50, slop=1e-4)
    passes, npos = W.project(p)
    assert passes == 1 and npos == 1
    st = W.state()
    gap = st["x"][1, 2] - st["x"][0, 2] - 1.0
    assert abs(gap - (-1e-4)) < 1e-14
    # the correction splits by inverse mass: dx_A = -lambda/m_A, dx_B = +lambda/m_B, lambda = 0.0499/(1 + 1/3)
    lam = (0.05 - 1e-4) / (1.0 + 1.0 / 3.0)
    assert np.allclose(st["x"][0], [0, 0, -lam], atol=1e-15) and np.allclose(st["x"][1], [0, 0, 0.95 + lam / 3], atol=1e-15)
    # within slop: nothing to do
    W2 = _crossing_pair(-5e-5)
    W2.detect(0, 0.0)
    assert W2.project(p) == (0, 0)


def test_position_projection_conserves_mass_weighted_position_and_reduces_penetration():
    rng = np.random.default_rng(17)
    n, side = 400, 14.0
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.5, n)
    h = rng.uniform(0.5, 2.0, n)
    m = rng.uniform(0.5, 2.0, n)
    W = O.World(x, q, np.zeros((n, 3)), np.zeros((n, 3)), r, h, m)
    W.set_box([-100] * 3, [100] * 3, [0, 0, 0])

    def pen():
        nc = W.detect(0, 0.0)
        c = W.contacts(nc)
        d = (r[c["a"]] + r[c["b"]]) - c["dist"]
        return nc, float(d.sum()), float(np.median(d))

    nc0, pen0, med0 = pen()
    assert nc0 > 100 and pen0 > 100.0 and med0 > 0.1
    x0 = W.state()["x"]
    p = O.nsc_params(omega=1.0, cfm=0.0, position_stabilization=1, position_iters=5, position_psor_iters=50, slop=1e-4)
    passes, npos = W.project(p)
    assert passes >= 1 and npos > 50
    x1 = W.state()["x"]
    # corrections are equal and opposite along each normal, weighted by inverse mass
    assert np.allclose((m[:, None] * (x1 - x0)).sum(0), 0.0, atol=1e-11)
    nc1, pen1, med1 = pen()
    # total depth drops by > 5x; the typical pair ends exactly at the allowed slop
    assert pen1 < 0.2 * pen0
    assert abs(med1 - 1e-4) < 1e-6
