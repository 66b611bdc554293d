"""A settling-bed island whose solve stalls: impulses drift along an internal
squeeze of two rods pressed together on the floor (CPU only).

Fixture tests/golden/wedged_island.npz (make_wedged_island.py): the smallest
island (28 manifolds, 19 rods) that the 1e5-rod settling workload's island
solve (block tangential update, stall window 1000) leaves above the
tolerance, at step 4189, with the previous-step impulses it starts from and
the library's recorded result.

What the oracle shows: the solve stops at the stall rule after 2000 sweeps
with the same residual the library recorded; the impulses move by the same
amount every block of sweeps (a linear drift, not a geometric decay that
extrapolation could shortcut) while the residual stays flat; the drift sits in
two rods that press on each other and slide on the floor (both wall contacts
on the friction limit); run on without the stall rule, the drift reaches an
active-set change after about 30 000 sweeps and the residual then falls to
round-off.
"""
import os

import numpy as np

import oracle_py as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wedged_island.npz")


def world(d):
    W = O.World(d["x"], d["q"], d["v_free"], d["w_free"], d["r"], d["h"], d["mass"])
    W.set_box(d["lo"], d["hi"], [1, 1, 0], 1, 1)
    W.set_physics(tuple(d["gravity"]), float(d["dt"]))
    W.set_manifolds(d["rec"], d["ab"], d["lam0"])
    W.apply_warm()
    return W


def params(d, max_sweeps=10000):
    return O.nsc_params(mu=float(d["mu"]), beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=max_sweeps,
                        check_every=8, friction_proj=int(d["friction_proj"]), stop_rule=1, warm_start=1)


def test_fixture_is_one_small_island():
    d = np.load(FIX)
    ab = d["ab"]
    assert len(ab) == 28 and len(d["r"]) == 19 and (ab[:, 1] < 0).sum() == 10
    assert sorted(d["order"].tolist()) == list(range(28))
    par = np.arange(19)

    def find(i):
        while par[i] != i:
            i = par[i]
        return i

    for a, b in ab:
        if b >= 0:
            par[find(a)] = find(b)
    assert len({find(a) for a in range(19)}) == 1


def test_stall_rule_stops_where_the_library_did():
    d = np.load(FIX)
    W = world(d)
    sw, res, upd, hs = W.solve_island(params(d), d["order"], stall_sweeps=int(d["stall_sweeps"]), stall_omega=0.0,
                                      hot=0)
    assert sw == 2000 and int(d["gpu_sweeps"]) == 2000
    # the island held the bed's largest residual: the library's recorded value
    assert res == float(d["gpu_residual"]) and 1e-8 < res < 2e-8


def test_impulses_drift_linearly_at_a_flat_residual():
    d = np.load(FIX)
    W = world(d)
    p = params(d)
    order = d["order"]
    W.sweeps(p, 1024, order)
    lams, res = [], []
    for _ in range(3):
        lams.append(W.manifolds()[2].copy())
        res.append(W.residual(p))
        W.sweeps(p, 1024, order)
    d1, d2 = lams[1] - lams[0], lams[2] - lams[1]
    n1, n2 = np.linalg.norm(d1), np.linalg.norm(d2)
    assert 0.99 < n2 / n1 < 1.01 and float((d1 * d2).sum()) / (n1 * n2) > 0.999
    assert abs(res[2] - res[0]) < 0.01 * res[0] and res[0] > 1e-8
    # where: the three manifolds that move most are two sliding floor contacts
    # and the contact between their two rods
    ab, lam = d["ab"], lams[2]
    top = np.argsort(-np.abs(d2).sum(1))[:3]
    walls = [i for i in top if ab[i, 1] < 0]
    rods = [i for i in top if ab[i, 1] >= 0]
    assert len(walls) == 2 and len(rods) == 1
    for i in walls:
        assert np.hypot(lam[i, 1], lam[i, 2]) >= float(d["mu"]) * lam[i, 0] * (1 - 1e-9)
    assert {ab[i, 0] for i in walls} == set(ab[rods[0]].tolist())


def test_without_the_stall_rule_the_drift_ends_in_an_active_set_change():
    d = np.load(FIX)
    W = world(d)
    p = params(d, max_sweeps=40000)
    sw, res, _, _ = W.solve_island(p, d["order"], stall_sweeps=0, stall_omega=0.0, hot=0)
    assert 20000 < sw < 40000 and res <= 1e-8
    # past the change the iteration converges fast: 64 more sweeps to round-off
    W.sweeps(p, 64, d["order"])
    assert W.residual(p) <= 1e-13
