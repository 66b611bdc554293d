"""GPU parity of the force laws, the integrators and the fused steps.

Forces and torques: within 1e-12 relative to the largest force (FP64;
exp/log/pow come from the device math library, not glibc).  Integrators:
the NSC stages are bit-identical except for the closed-form inertia, the
Verlet stages use device sin/cos; 1e-13 relative.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L
from rod_dynamics_py import scenes

pytestmark = pytest.mark.gpu


def packing(n=2500, side=20.0, seed=3, periodic=(1, 1, 1)):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return dict(x=x, q=q, v=rng.normal(size=(n, 3)), w=rng.normal(size=(n, 3)), r=rng.uniform(0.3, 0.5, n),
                h=rng.uniform(0.5, 2.0, n), m=rng.uniform(0.5, 2.0, n), rest=np.full(n, 0.5), side=side,
                periodic=periodic, f=rng.normal(size=(n, 3)), tau=rng.normal(size=(n, 3)))


def both(ctx, B, dt=1e-4, g=(0, 0, -1.0), damp=(0.0, 0.0, 1e30)):
    n = len(B["r"])
    ctx.set_bodies(B["x"], B["q"], B["v"], B["w"], B["r"], B["h"], B["m"], B["rest"])
    ctx.set_systems([0, n], [L.system_desc([0] * 3, [B["side"]] * 3, B["periodic"], mu=0.5)])
    ctx.set_physics(L.physics(g, dt, *damp))
    W = O.World(B["x"], B["q"], B["v"], B["w"], B["r"], B["h"], B["m"], B["rest"])
    W.set_box([0] * 3, [B["side"]] * 3, B["periodic"])
    W.set_physics(g, dt, *damp)
    return W


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_soft_forces_match_oracle(gpu_ctx_factory, model):
    B = packing()
    ctx = gpu_ctx_factory()
    W = both(ctx, B)
    gp = L.soft_params(delta=0.005, k_scaler=1e3, damping=0.2, mu=0.5, mu_static=0.8, nu=1e-2, kt=500.0,
                       vel_deadband=0.5, friction_model=model)
    op = O.soft_params(delta=0.005, k_scaler=1e3, damping=0.2, mu=0.5, mu_static=0.8, nu=1e-2, kt=500.0,
                       vel_deadband=0.5, friction_model=model)
    for it in range(3):  # repeated calls exercise the Cundall-Strack history
        nc = ctx.detect(L.detect_params(0.005, 0.0, 1.0))
        assert W.detect(0, 0.005) == nc and nc > 200
        pe_g = ctx.forces_soft(gp)
        pe_o = W.soft_forces(op)
        sg, so = ctx.download_state(), W.state()
        assert rel(sg["f"], so["f"]) < 1e-12 and rel(sg["tau"], so["tau"]) < 1e-12
        assert pe_g == pytest.approx(pe_o, rel=1e-11)


@pytest.mark.parametrize("law", [0, 1])
def test_hm_forces_and_history_match_oracle(gpu_ctx_factory, law):
    B = packing(seed=7)
    ctx = gpu_ctx_factory()
    W = both(ctx, B)
    gp = L.hm_params(E=1e5, nu_poisson=0.25, restitution=0.5, mu_k=0.4, mu_s=0.6, v_c=1e-2, mu_roll=0.05,
                     enable_tangential=1, enable_rolling=1, law=law)
    op = O.hm_params(E=1e5, nu_poisson=0.25, restitution=0.5, mu_k=0.4, mu_s=0.6, v_c=1e-2, mu_roll=0.05,
                     enable_tangential=1, enable_rolling=1, law=law)
    rng = np.random.default_rng(1)
    x = B["x"].copy()
    for it in range(8):
        nc = ctx.detect(L.detect_params(0.0, 0.0, 1.0))
        assert W.detect(0, 0.0) == nc
        ctx.forces_hm(gp)
        W.hm_forces(op)
        sg, so = ctx.download_state(), W.state()
        assert rel(sg["f"], so["f"]) < 1e-12, it
        assert rel(sg["tau"], so["tau"]) < 1e-12, it
        assert ctx.history_size() == W.hm_history_size(), it
        # move the rods a little so contacts appear and disappear
        x = x + rng.normal(scale=0.05, size=x.shape)
        zero = np.zeros_like(x)
        ctx.upload_state(x=x)
        ctx.upload_forces(zero, zero)
        W.set_state(x=x)
        W.set_forces(zero, zero)


@pytest.mark.parametrize("stage", [0, 1, 2, 3, 4])
def test_integrator_stages_match_oracle(gpu_ctx_factory, stage):
    B = packing(n=5000, periodic=(1, 1, 0))
    B["x"][:100] += 25.0  # wrap on periodic axes
    ctx = gpu_ctx_factory()
    W = both(ctx, B, dt=1e-3, damp=(0.08, 0.12, 2.5))
    ctx.upload_forces(B["f"], B["tau"])
    W.set_forces(B["f"], B["tau"])
    ctx.integrate(stage)
    W.integrate(stage)
    sg, so = ctx.download_state(), W.state()
    for k in ("x", "v", "w", "f", "tau"):
        assert rel(sg[k], so[k]) < 1e-13, k
    assert rel(sg["q"], so["q"]) < 1e-13
    if stage == 1:
        assert np.array_equal(sg["x"], so["x"])  # position update and wrap: same operations


@pytest.mark.parametrize("warm_walls", [0, 1])
def test_nsc_steps_match_oracle_short_horizon(gpu_ctx_factory, warm_walls):
    sc = scenes.c2_settling(n=1500)
    v0 = np.tile([0.0, 0.0, -5.0], (sc.n, 1))
    ctx = gpu_ctx_factory()
    sc.load(ctx)
    ctx.upload_state(v=v0)
    p = sc.params
    p.tol = 1e-12
    p.check_every = 1
    p.warm_walls = warm_walls
    st = ctx.step_nsc(sc.detect, p, 30)
    assert st["nonconverged_steps"] == 0 and st["manifolds_total"] > 0
    W = O.world_from_scene(sc)
    W.set_state(v=v0)
    op = O.nsc_params(mu=0.5, beta=0.0, cfm=0.0, omega=1.3, tol=1e-12, friction_proj=1, warm_start=1,
                      warm_walls=warm_walls)
    W.nsc_step(op, 30, method=1)
    sg, so = ctx.download_state(), W.state()
    # colour-order vs sorted-order Gauss-Seidel: equal at convergence up to the
    # residual tolerance; positions agree to well below 1e-9
    assert np.abs(sg["x"] - so["x"]).max() < 1e-9
    assert np.abs(sg["v"] - so["v"]).max() < 1e-7
    ke_g, phi_g = ctx.observables()
    assert ke_g == pytest.approx(W.kinetic_energy(), rel=1e-9)
    assert phi_g == pytest.approx(W.packing_fraction(), rel=1e-12)


@pytest.mark.parametrize("model", ["soft", "hm"])
def test_verlet_steps_match_oracle_short_horizon(gpu_ctx_factory, model):
    B = packing(n=1500, side=18.0, seed=19)
    B["v"] *= 0.1
    B["w"] *= 0.1
    ctx = gpu_ctx_factory()
    W = both(ctx, B, dt=1e-4, g=(0, 0, 0))
    if model == "soft":
        gp = L.soft_params(delta=0.005, k_scaler=1e3, damping=0.1, mu=0.5, nu=1e-2)
        op = O.soft_params(delta=0.005, k_scaler=1e3, damping=0.1, mu=0.5, nu=1e-2)
        st = ctx.step_vv(L.detect_params(0.005, 0.0, 1.0), 0, gp, 20)
        W.vv_step(0, op, 20, method=1)
    else:
        gp = L.hm_params(E=1e4, restitution=0.5, mu_k=0.4, mu_s=0.4, v_c=1e-2)
        op = O.hm_params(E=1e4, restitution=0.5, mu_k=0.4, mu_s=0.4, v_c=1e-2)
        st = ctx.step_vv(L.detect_params(0.0, 0.0, 1.0), 1, gp, 20)
        W.vv_step(1, op, 20, method=1)
    sg, so = ctx.download_state(), W.state()
    assert st["contacts_total"] > 1000
    assert np.abs(sg["x"] - so["x"]).max() < 1e-11
    assert np.abs(sg["v"] - so["v"]).max() < 1e-9
    assert np.abs(sg["w"] - so["w"]).max() < 1e-9


@pytest.mark.parametrize("model", [0, 1, 2])
def test_kept_soft_contact_forces(gpu_ctx_factory, model):
    """The per-contact soft forces kept for the network log: scattered onto
    the bodies (b gets force_b + friction_b, a their negatives) they give the
    body forces, which match the oracle; force_b lies along the contact
    normal and the friction force across it."""
    B = packing(seed=5)
    B["f"] = np.zeros_like(B["f"])
    B["tau"] = np.zeros_like(B["tau"])
    ctx = gpu_ctx_factory()
    W = both(ctx, B, g=(0, 0, 0))
    ctx.keep_soft_forces(True)
    gp = L.soft_params(delta=0.005, k_scaler=1e3, damping=0.2, mu=0.5, mu_static=0.8, nu=1e-2, kt=500.0,
                       vel_deadband=0.5, friction_model=model)
    op = O.soft_params(delta=0.005, k_scaler=1e3, damping=0.2, mu=0.5, mu_static=0.8, nu=1e-2, kt=500.0,
                       vel_deadband=0.5, friction_model=model)
    for it in range(2):  # the second call reads the Cundall-Strack history of the first
        nc = ctx.detect(L.detect_params(0.005, 0.0, 1.0))
        assert W.detect(0, 0.005) == nc and nc > 200
        f0 = ctx.download_state()["f"]  # forces accumulate over the calls
        ctx.forces_soft(gp)
        W.soft_forces(op)
    raw, kf = ctx.contacts_raw(), ctx.soft_forces()
    assert kf.shape == (len(raw), 6)
    fb, fra = kf[:, 0:3], kf[:, 3:6]
    tot = np.zeros((len(B["r"]), 3))
    np.add.at(tot, raw["b"], fb - fra)
    np.add.at(tot, raw["a"], fra - fb)
    sg, so = ctx.download_state(), W.state()
    assert rel(sg["f"], so["f"]) < 1e-12
    assert rel(tot, sg["f"] - f0) < 1e-12
    n = np.stack([raw["nx"], raw["ny"], raw["nz"]], 1).astype(np.float64)
    assert np.abs(np.cross(fb, n)).max() <= 1e-6 * np.abs(fb).max()
    assert np.abs(np.sum(fra * n, 1)).max() <= 1e-6 * max(np.abs(fra).max(), 1e-300)
    assert np.abs(fra).max() > 0
    ctx.keep_soft_forces(False)
    assert len(ctx.soft_forces()) == 0
