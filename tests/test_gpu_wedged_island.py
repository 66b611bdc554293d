"""The wedged settling-bed island (tests/golden/wedged_island.npz, 28
manifolds, 19 rods: the register island class) solved by the library from
the fixture's previous-step impulses, equal bit for bit to the oracle's
island solve: stopped by the stall rule as in the settling run, and run on
without it until the drift ends and the island converges."""
import os

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wedged_island.npz")


@pytest.mark.parametrize("stall,max_sweeps", [(1000, 10000), (0, 40000)])
def test_wedged_island_solve_bitwise_equal_oracle(gpu_ctx_factory, stall, max_sweeps):
    print(f"wedged island solve stall_sweeps={stall} max_sweeps={max_sweeps}", flush=True)
    d = np.load(FIX)
    n = len(d["r"])
    mu = float(d["mu"])
    ctx = gpu_ctx_factory()
    ctx.set_bodies(d["x"], d["q"], d["v_free"], d["w_free"], d["r"], d["h"], d["mass"], np.zeros(n))
    ctx.set_systems([0, n], [L.system_desc(d["lo"], d["hi"], (1, 1, 0), True, True, mu)])
    ctx.set_physics(L.physics(tuple(d["gravity"]), float(d["dt"]), 0.08, 0.12, 80.0))
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=max_sweeps, check_every=8,
                     friction_proj=int(d["friction_proj"]), warm_mode=1, stop_rule=1, stall_sweeps=stall,
                     stall_omega=0.0, hot_set=int(d["hot_set"]))
    ctx.detect(L.detect_params(0.0, 16.0, 1.0))
    nm = ctx.nsc_build(p)
    assert nm == len(d["ab"])
    rec, ab, lam_cold = ctx.nsc_manifolds(nm)
    # the fixture's manifolds, possibly in another order: match them by (a, b, normal)
    key = lambda r_, a_: [tuple(a_[i].tolist()) + tuple(np.round(r_[i, 0:3], 12)) for i in range(len(a_))]
    kf = {k: i for i, k in enumerate(key(d["rec"], d["ab"]))}
    perm = np.array([kf[k] for k in key(rec, ab)])
    assert np.array_equal(rec, d["rec"][perm]) and np.array_equal(ab, d["ab"][perm])
    ctx.nsc_set_lambda(d["lam0"][perm])  # the previous-step impulses of the settling run
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p, allow_noconv=True)
    nisl, counts = ctx.island_info()
    assert nisl == 1 and counts == [0, 1, 0, 0], counts  # one island, register class of up to 64 manifolds
    order = ctx.solve_order(nm)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()

    # the oracle holds the manifolds in the library's solve order, so that its
    # warm-start pass adds each body's impulses in the same (colour) order
    W = O.World(d["x"], d["q"], s0["v"], s0["w"], d["r"], d["h"], d["mass"])
    W.set_box(d["lo"], d["hi"], [1, 1, 0], 1, 1)
    W.set_physics(tuple(d["gravity"]), float(d["dt"]))
    W.set_manifolds(rec[order], ab[order], d["lam0"][perm][order])
    W.apply_warm()
    op = O.nsc_params(mu=mu, beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=max_sweeps, check_every=8,
                      friction_proj=int(d["friction_proj"]), stop_rule=1, warm_start=1)
    sw, res, _, _ = W.solve_island(op, np.arange(nm, dtype=np.int32), stall_sweeps=stall, stall_omega=0.0, hot=0)
    print(f"  GPU {st['sweeps']} sweeps, residual {st['residual']:.3e}; oracle {sw}, {res:.3e}", flush=True)
    assert st["sweeps"] == sw and st["residual"] == res
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g[order], lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])
    if stall:
        assert sw == 2000 and res == float(d["gpu_residual"])  # as in the settling run
    else:
        assert 20000 < sw < 40000 and res <= 1e-8
