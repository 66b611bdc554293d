"""GPU parity of position stabilisation (nsc_contact.cpp:659-880): one pass
against the oracle swept in the same order, several passes against the
oracle's penetration statistics, and inside the NSC step."""
import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu


def bed(seed=17, n=400, side=14.0, periodic=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return dict(x=x, q=q, r=rng.uniform(0.3, 0.5, n), h=rng.uniform(0.5, 2.0, n), m=rng.uniform(0.5, 2.0, n),
                side=side, periodic=(1, 1, 1) if periodic else (0, 0, 0))


def load(ctx, B):
    n = len(B["r"])
    ctx.set_bodies(B["x"], B["q"], np.zeros((n, 3)), np.zeros((n, 3)), B["r"], B["h"], B["m"], np.zeros(n))
    lo, hi = ([0.0] * 3, [B["side"]] * 3) if B["periodic"][0] else ([-100.0] * 3, [100.0] * 3)
    ctx.set_systems([0, n], [L.system_desc(lo, hi, B["periodic"], 0, 0, 0.5)])
    ctx.set_physics(L.physics((0, 0, 0), 1e-3, 0, 0, 1e30))
    W = O.World(B["x"], B["q"], np.zeros((n, 3)), np.zeros((n, 3)), B["r"], B["h"], B["m"])
    W.set_box(lo, hi, B["periodic"])
    return W


@pytest.mark.parametrize("periodic", [False, True])
def test_one_pass_matches_oracle_in_the_same_order(gpu_ctx_factory, periodic):
    B = bed(periodic=periodic)
    ctx = gpu_ctx_factory()
    W = load(ctx, B)
    p = L.nsc_params(omega=1.0, cfm=1e-6, position_stabilization=1, position_iters=1, position_psor_iters=50,
                     slop=1e-4)
    passes, nfirst = ctx.nsc_project(L.detect_params(0.0, 1.0, 1.0), p)
    assert passes == 1 and nfirst > 50
    order = ctx.solve_order(nfirst)
    s = ctx.download_state()
    W.detect(0, 0.0)
    op = O.nsc_params(omega=1.0, cfm=1e-6, position_stabilization=1, position_iters=1, position_psor_iters=50,
                      slop=1e-4)
    po, no = W.project(op, order)
    assert (po, no) == (passes, nfirst)
    so = W.state()
    # closed-form vs matrix inertia products differ in the last bits only
    assert np.max(np.abs(s["x"] - so["x"])) < 1e-11
    assert np.max(np.abs(s["q"] - so["q"])) < 1e-11


def test_several_passes_reduce_penetration_like_the_oracle(gpu_ctx_factory):
    B = bed(seed=23)
    ctx = gpu_ctx_factory()
    W = load(ctx, B)
    r = B["r"]

    def pen_gpu():
        nc = ctx.detect(L.detect_params(0.0, 1.0, 1.0))
        c = ctx.contacts(nc)
        d = (r[c["a"]] + r[c["b"]]) - c["dist"]
        return float(d.sum()), float(np.median(d))

    def pen_oracle():
        nc = W.detect(0, 0.0)
        c = W.contacts(nc)
        d = (r[c["a"]] + r[c["b"]]) - c["dist"]
        return float(d.sum()), float(np.median(d))

    s0, _ = pen_gpu()
    p = L.nsc_params(omega=1.0, cfm=0.0, position_stabilization=1, position_iters=5, position_psor_iters=50, slop=1e-4)
    ctx.nsc_project(L.detect_params(0.0, 1.0, 1.0), p)
    W.detect(0, 0.0)  # the oracle projects over the current contact list; the GPU entry point detects itself
    W.project(O.nsc_params(omega=1.0, cfm=0.0, position_stabilization=1, position_iters=5, position_psor_iters=50,
                           slop=1e-4))
    sg, mg = pen_gpu()
    so, mo = pen_oracle()
    assert s0 > 100.0
    assert sg < 0.2 * s0 and so < 0.2 * s0
    assert abs(mg - 1e-4) < 1e-6 and abs(mo - 1e-4) < 1e-6


def test_nsc_step_with_position_stabilisation(gpu_ctx_factory):
    B = bed(seed=29, periodic=True)
    ctx = gpu_ctx_factory()
    load(ctx, B)
    r = B["r"]
    dp = L.detect_params(0.0, 1.0, 1.0)
    nc = ctx.detect(dp)
    c = ctx.contacts(nc)
    pen0 = float(((r[c["a"]] + r[c["b"]]) - c["dist"]).sum())
    p = L.nsc_params(beta=0.2, cfm=1e-6, omega=1.0, tol=1e-8, max_sweeps=2000, check_every=4, friction_proj=1,
                     position_stabilization=1, position_iters=5, position_psor_iters=50, slop=1e-4, stall_sweeps=250)
    st = ctx.step_nsc(dp, p, 3)
    assert st["steps"] == 3
    nc = ctx.detect(dp)
    c = ctx.contacts(nc)
    pen = float(((r[c["a"]] + r[c["b"]]) - c["dist"]).sum())
    assert pen < 0.2 * pen0
    s = ctx.download_state()
    assert np.all(np.isfinite(s["x"])) and np.allclose(np.linalg.norm(s["q"], axis=1), 1.0, atol=1e-12)
