"""The settled-bed island that the full-sweep solve gave up on (CPU only).

Fixture tests/golden/settled_island.npz: the largest contact island (2546
manifolds, 1838 rods, 595 floor contacts) of the 1e5-rod settling workload
at step 4626, captured from the library's GPU run with the manifolds it
built and the previous-step impulses they start from (make_settled_island.py).
The library's full-sweep solve stopped there at the stall rule with residual
5.8e-7 (recorded in the fixture).

The oracle shows what happens: with full sweeps the residual sits on a
plateau (about 5.9e-7 from a few hundred to 2000 sweeps), so the stall rule
(window 1000, stop at the first stall) ends the solve unconverged; without
the stall rule the same iteration does reach 1e-8 (a slow stick/slip mode of
a few wall contacts, not a fixed floor); the hot-set phase of the LDS island
solve (Gauss-Seidel on the manifolds around the few contacts above tol)
reaches 1e-8 with the stall rule on, with a fraction of the manifold
updates, to the same solution.
"""
import os

import numpy as np

import oracle_py as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "settled_island.npz")


def island_world(d):
    W = O.World(d["x"], d["q"], d["v_free"], d["w_free"], d["r"], d["h"], d["mass"])
    W.set_box(d["lo"], d["hi"], [1, 1, 0], 1, 1)
    W.set_physics(tuple(d["gravity"]), float(d["dt"]))
    W.set_manifolds(d["rec"], d["ab"], d["lam0"])
    W.apply_warm()
    return W


def params(d):
    return O.nsc_params(mu=float(d["mu"]), beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=8,
                        friction_proj=1, stop_rule=1, warm_start=1)


def test_fixture_is_the_stalled_island():
    d = np.load(FIX)
    m = len(d["ab"])
    assert m == 2546 and len(d["r"]) == 1838
    assert int(d["gpu_nonconverged"]) == 1 and 1e-7 < float(d["gpu_residual"]) < 1e-6
    # one island: every rod-rod manifold connects the same component
    par = np.arange(len(d["r"]))

    def find(i):
        while par[i] != i:
            par[i] = par[par[i]]
            i = par[i]
        return i

    for a, b in d["ab"]:
        if b >= 0:
            par[find(a)] = find(b)
    assert len({find(a) for a in range(len(d["r"]))}) == 1


def test_full_sweeps_stall_on_a_plateau_and_hot_sets_converge():
    d = np.load(FIX)
    p = params(d)
    ix = np.arange(len(d["ab"]))
    # full sweeps, stall rule on: stops unconverged at the 2000-sweep milestone
    W = island_world(d)
    sw, res, upd, hs = W.solve_island(p, ix, stall_sweeps=1000, stall_omega=0.0, hot=0)
    assert sw == 2000 and 5e-7 < res < 7e-7 and hs == 0
    # the same iteration without the stall rule: slow, but it converges
    W2 = island_world(d)
    sw2, res2, upd2, _ = W2.solve_island(p, ix, stall_sweeps=0, hot=0)
    assert res2 <= 1e-8 and 2000 < sw2 < 4000
    assert W2.residual(p) == res2
    # hot sets, stall rule on: converged, 4x fewer manifold updates
    W3 = island_world(d)
    sw3, res3, upd3, hs3 = W3.solve_island(p, ix, stall_sweeps=1000, stall_omega=0.0, hot=1)
    assert res3 <= 1e-8 and W3.residual(p) == res3
    assert hs3 > 0 and upd3 * 4 < upd2
    # same solution within the tolerance: impulses and velocities of the two
    # converged solves agree to the residual scale
    _, _, l2 = W2.manifolds()
    _, _, l3 = W3.manifolds()
    s2, s3 = W2.state(), W3.state()
    assert np.max(np.abs(s2["v"] - s3["v"])) < 1e-6
    assert abs(l2[:, 0].sum() - l3[:, 0].sum()) < 1e-4 * l2[:, 0].sum()


def test_hot_set_off_is_the_plain_solve():
    d = np.load(FIX)
    p = params(d)
    ix = np.arange(len(d["ab"]))
    W = island_world(d)
    sw, res, _, _ = W.solve_island(p, ix, stall_sweeps=0, hot=0)
    V = island_world(d)
    res_v, sw_v = V.solve(p, ix.astype(np.int32))
    assert (sw, res) == (sw_v, res_v)
    _, _, la = W.manifolds()
    _, _, lb = V.manifolds()
    assert np.array_equal(la, lb)
