"""Reduced versions of the dense Hertz-Mindlin workload (C3) and the batched
hard-contact workload (C4) on the GPU against the oracle.

C3: polydisperse rods (L lognormal, median 16), generated at phi = 0.05 and
compressed on the GPU to phi = 0.30, thermal velocities; then 20
velocity-Verlet Hertz-Mindlin steps with tangential spring history, each
preceded by the affine strain x <- s x, box <- s box, s = 1 - 1e-3 dt, on
both sides from the same compressed state (hertz_mindlin.cpp:183-247 with
the strain protocol of the workload).

C4: several independent periodic packings of different size, aspect ratio and
friction coefficient in one context, 40 hard-contact steps (thermal gas at
T = 25 so that every system collides), against one
oracle world per system with that system's mu (nsc_contact.cpp:239-655).
"""
import math

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L
from rod_dynamics_py import protocols, scenes

pytestmark = pytest.mark.gpu


def test_c3_dense_strained_hm_matches_oracle(gpu_ctx_factory):
    print("c3 reduced: 5000 rods, compress then 20 strained HM steps", flush=True)
    sc = scenes.c3_dense(n=5000)
    sc.detect = L.detect_params(0.0, 4.0, 1.0)
    ctx = gpu_ctx_factory()
    sc.load(ctx)
    protocols.compress_nsc(ctx, sc, sc.extra["side1"])
    side = float(sc.extra["side_now"])
    assert protocols.packing_fraction_from_side(sc, side) == pytest.approx(0.30, rel=1e-6)
    st0 = ctx.download_state()
    v, w = L.gen_thermal(st0["q"], sc.r, sc.h, sc.mass, 1.0, sc.extra["seed"] + 29)
    ctx.upload_state(v=v, w=w)
    hp = sc.params
    W = O.World(st0["x"], st0["q"], v, w, sc.r, sc.h, sc.mass, sc.rest)
    W.set_box([0.0] * 3, [side] * 3, [1, 1, 1])
    W.set_physics((0.0, 0.0, 0.0), sc.phys.dt)
    op = O.hm_params(E=hp.E, nu_poisson=hp.nu_poisson, restitution=hp.restitution, mu_k=hp.mu_k, mu_s=hp.mu_s,
                     v_c=hp.v_c, mu_roll=hp.mu_roll, enable_tangential=1, enable_rolling=0, law=hp.law)
    s = 1.0 - 1e-3 * sc.phys.dt
    span = side
    contacts = 0
    for k in range(20):
        ctx.affine_scale(s)
        st = ctx.step_vv(sc.detect, 1, hp, 1)
        contacts += st["contacts_total"]
        x = W.state()["x"]
        W.set_state(x=0.0 + s * (x - 0.0))
        span *= s
        W.set_box([0.0] * 3, [span] * 3, [1, 1, 1])
        W.vv_step(1, op, 1, method=1)
    sg, so = ctx.download_state(), W.state()
    assert np.array_equal(sg["x"].max(axis=0) <= span, np.array([True] * 3))
    assert contacts > 20 * 2 * 5000  # dense: several contacts per rod in every evaluation
    assert np.abs(sg["x"] - so["x"]).max() < 1e-11
    assert np.abs(sg["v"] - so["v"]).max() < 1e-8 * max(1.0, np.abs(so["v"]).max())
    assert np.abs(sg["w"] - so["w"]).max() < 1e-8 * max(1.0, np.abs(so["w"]).max())
    hg = C_hist(ctx)
    assert hg == W.hm_history_size() and hg > 5000


def C_hist(ctx):
    import ctypes

    n = ctypes.c_int()
    ctx._check(ctx._L.rdk_history_size(ctx._h, ctypes.byref(n)))
    return n.value


def test_c4_batched_systems_match_per_system_oracle(gpu_ctx_factory):
    print("c4 reduced: 8 systems, 40 hard-contact steps", flush=True)
    sc = scenes.c4_batched(nsys=8, seed=scenes.SEED0 + 44, kT=25.0, threads=8, n_range=(300, 1500),
                           aspects=(4, 8, 16, 32))
    mus = [d.mu for d in sc.descs]
    assert len(set(round(m, 6) for m in mus)) == 8 and max(mus) - min(mus) > 0.3
    p = sc.params
    p.tol, p.omega, p.check_every, p.stall_sweeps = 1e-12, 1.0, 1, 0
    ctx = gpu_ctx_factory()
    sc.load(ctx)
    det = L.detect_params(0.0, 4.0, 1.0)
    st = ctx.step_nsc(det, p, 40)
    print(f"  {st['manifolds_total']} manifolds, {st['sweeps_total']} sweeps over 40 steps", flush=True)
    assert st["nonconverged_steps"] == 0 and st["manifolds_total"] > 1000
    sg = ctx.download_state()
    for k, d in enumerate(sc.descs):
        a0, a1 = int(sc.offsets[k]), int(sc.offsets[k + 1])
        W = O.World(sc.x[a0:a1], sc.q[a0:a1], sc.v[a0:a1], sc.w[a0:a1], sc.r[a0:a1], sc.h[a0:a1], sc.mass[a0:a1],
                    sc.rest[a0:a1])
        W.set_box(list(d.lo), list(d.hi), [int(v) for v in d.periodic], 0, 0)
        W.set_physics(tuple(sc.phys.gravity[:]), sc.phys.dt)
        op = O.nsc_params(mu=d.mu, beta=0.0, cfm=0.0, omega=1.0, tol=1e-12, max_sweeps=10000, check_every=1,
                          friction_proj=p.friction_proj, stop_rule=1, warm_start=1)
        W.nsc_step(op, 40, method=1)
        so = W.state()
        # colour-order vs sorted-order Gauss-Seidel, both to 1e-12: equal up to the tolerance
        assert np.abs(sg["x"][a0:a1] - so["x"]).max() < 1e-9, k
        assert np.abs(sg["v"][a0:a1] - so["v"]).max() < 1e-7, k
        assert np.abs(sg["w"][a0:a1] - so["w"]).max() < 1e-7, k
