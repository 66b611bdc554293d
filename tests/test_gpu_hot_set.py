"""LDS island solve to the tolerance, with and without its hot-set phase and
selective rounds, equal bit for bit to the oracle's restatement
(orc_nsc_solve_island: stall rule, hot sets, selective rounds; sweeps, hot
sweeps and manifold updates counted alike), on the settled-bed island of tests/golden/settled_island.npz
(2546 manifolds, 1838 rods: one island of the LDS class), solved from a
cold start, with the scalar (friction_proj 1), the block (friction_proj 2)
and the hybrid (friction_proj 3) tangential update."""
import os

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "settled_island.npz")


@pytest.mark.parametrize("proj", [1, 2, 3])
@pytest.mark.parametrize("hot", [0, 1, 2])
def test_lds_island_solve_bitwise_equal_oracle(gpu_ctx_factory, hot, proj):
    print(f"settled island solve hot_set={hot} friction_proj={proj}", flush=True)
    d = np.load(FIX)
    n = len(d["r"])
    mu = float(d["mu"])
    ctx = gpu_ctx_factory()
    ctx.set_bodies(d["x"], d["q"], d["v_free"], d["w_free"], d["r"], d["h"], d["mass"], np.zeros(n))
    ctx.set_systems([0, n], [L.system_desc(d["lo"], d["hi"], (1, 1, 0), True, True, mu)])
    ctx.set_physics(L.physics(tuple(d["gravity"]), float(d["dt"]), 0.08, 0.12, 80.0))
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=8, friction_proj=proj,
                     warm_mode=0, stop_rule=1, stall_sweeps=1000, stall_omega=0.0, hot_set=hot)
    ctx.detect(L.detect_params(0.0, 16.0, 1.0))
    nm = ctx.nsc_build(p)
    assert nm == len(d["ab"])
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    assert not lam0.any()
    s0 = ctx.download_state()
    wc0 = ctx.work_counters()
    st = ctx.nsc_solve(p, allow_noconv=True)
    nisl, counts = ctx.island_info()
    assert nisl == 1 and counts == [0, 0, 0, 1], counts  # one island, LDS class
    order = ctx.solve_order(nm)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    wc1 = ctx.work_counters()
    hot_g = wc1["hot_sweeps"] - wc0["hot_sweeps"]
    upd_g = wc1["lds_manifold_sweeps"] - wc0["lds_manifold_sweeps"]

    W = O.World(d["x"], d["q"], s0["v"], s0["w"], d["r"], d["h"], d["mass"])
    W.set_box(d["lo"], d["hi"], [1, 1, 0], 1, 1)
    W.set_physics(tuple(d["gravity"]), float(d["dt"]))
    W.set_manifolds(rec, ab, lam0)
    op = O.nsc_params(mu=mu, beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=8,
                      friction_proj=proj, stop_rule=1, warm_start=0)
    sw, res, upd, hs = W.solve_island(op, order, stall_sweeps=1000, stall_omega=0.0, hot=hot)
    print(f"  GPU {st['sweeps']} sweeps (hot {hot_g}), residual {st['residual']:.3e}; oracle {sw} (hot {hs}), "
          f"{res:.3e}, {upd} updates", flush=True)
    assert st["sweeps"] == sw and hot_g == hs and upd_g == upd
    assert st["residual"] == res
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g, lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])
    if hot:
        assert hs > 0 and res <= 1e-8  # the phase ran and the island converged
    if hot == 2:
        # selective rounds ran: fewer updates than the same solve with hot sets only
        W.set_state(v=s0["v"], w=s0["w"])
        W.set_manifolds(rec, ab, lam0)
        _, _, upd1, _ = W.solve_island(op, order, stall_sweeps=1000, stall_omega=0.0, hot=1)
        assert upd < upd1, (upd, upd1)


def _solve_from(d, mode, weights=None):
    n = len(d["r"])
    ctx = L.Context(0)
    try:
        ctx.set_bodies(d["x"], d["q"], d["v_free"], d["w_free"], d["r"], d["h"], d["mass"], np.zeros(n))
        ctx.set_systems([0, n], [L.system_desc(d["lo"], d["hi"], (1, 1, 0), True, True, float(d["mu"]))])
        ctx.set_physics(L.physics(tuple(d["gravity"]), float(d["dt"]), 0.08, 0.12, 80.0))
        wm = {"cold": 0, "previous": 1, "learned": 4}[mode]
        p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=8, friction_proj=1,
                         warm_mode=wm, stop_rule=1, stall_sweeps=1000, stall_omega=0.0, hot_set=1, warm_walls=1)
        if weights is not None:
            ctx.mlp_set_params(weights)
        ctx.detect(L.detect_params(0.0, 16.0, 1.0))
        nm = ctx.nsc_build(p)
        assert nm == len(d["ab"])
        if mode == "previous":
            ctx.nsc_set_lambda(d["lam0"])  # the impulses the protocol's warm cache held
        st = ctx.nsc_solve(p, allow_noconv=True)
        rec, ab, lam = ctx.nsc_manifolds(nm)
        s1 = ctx.download_state()
    finally:
        ctx.close()
    # the oracle's residual of the GPU's final impulses and velocities
    W = O.World(d["x"], d["q"], s1["v"], s1["w"], d["r"], d["h"], d["mass"])
    W.set_box(d["lo"], d["hi"], [1, 1, 0], 1, 1)
    W.set_physics(tuple(d["gravity"]), float(d["dt"]))
    W.set_manifolds(rec, ab, lam)
    op = O.nsc_params(mu=float(d["mu"]), beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, friction_proj=1)
    return st["sweeps"], W.residual(op)


def test_learned_start_on_the_settled_island():
    """The shipped network (weights/warmstart_c2.npy, trained on the settling
    workload's own PSOR solutions) predicting every impulse of the captured
    island from the contact alone (warm_mode 4), against the cold start and
    the previous-step impulses, each solved to 1e-8 (hot sets on) and checked
    by the oracle's residual.  On this island the prediction needs 1176 sweeps
    against 6000 cold and 4360 from the previous-step impulses.  Over the whole
    settling protocol it does not beat the previous-step start (1752 against
    1045 sweeps per step on the settled bed), so the workloads keep the
    previous-step start; this test pins what the network gains where it
    does."""
    import os as _os

    d = np.load(FIX)
    w = np.load(_os.path.join(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))), "rod-dynamics-3d_amd",
                              "weights", "warmstart_c2.npy"))
    out = {m: _solve_from(d, m, w if m == "learned" else None) for m in ("cold", "previous", "learned")}
    print(f"settled island sweeps: {out}", flush=True)
    for m, (sw, res) in out.items():
        assert res <= 1e-8 * (1 + 1e-6), (m, res)
    assert out["learned"][0] * 2 <= out["cold"][0]
    assert out["learned"][0] * 2 <= out["previous"][0]
