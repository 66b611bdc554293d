"""Preparation protocols built on the library's own steps.

compress_nsc: the dense Hertz-Mindlin workload is specified at phi = 0.30,
which random sequential addition cannot reach for rods of L/d ~ 16, and the
reference has no compression routine.  The packing is generated at phi = 0.05
and compressed here: the boxes are scaled affinely in small increments and
after each increment overdamped hard-contact steps (Baumgarte beta, velocities
removed by the damping) push the rods out of the overlaps the increment made.
Contact mechanics of the timed run are untouched by this preparation.
"""
import math

from . import _lib as L


def compress_nsc(ctx, sc, side_target, factor=0.998, steps_per=2, sweeps=30, log=None):
    """Compress scene ``sc`` (already loaded in ``ctx``) from its current box side
    to ``side_target``; restores the scene's physics afterwards.  Returns the
    number of hard-contact steps taken."""
    side = float(sc.extra.get("side_now", sc.extra["side"]))
    dt = 1e-2
    ctx.set_physics(L.physics((0.0, 0.0, 0.0), dt, 1.0 / dt, 1.0 / dt, 1e30))  # v, w -> 0 after each step
    p = L.nsc_params(beta=0.8, cfm=1e-4, omega=1.0, tol=0.0, max_sweeps=sweeps, check_every=sweeps, friction_proj=1,
                     warm_mode=1, stop_rule=1)
    det = L.detect_params(0.0, sc.detect.chunk_len if sc.detect.chunk_len > 0 else 2.0, sc.detect.cell_scale)
    steps = 0
    k = 0
    while side > side_target * (1.0 + 1e-12):
        s = max(factor, side_target / side)
        ctx.affine_scale(s)
        side *= s
        ctx.step_nsc(det, p, steps_per)
        steps += steps_per
        k += 1
        if log is not None and k % 25 == 0:
            print(f"[compress] side {side:.3f} -> {side_target:.3f} ({steps} hard-contact steps)", file=log,
                  flush=True)
    ctx.step_nsc(det, p, 4 * steps_per)  # relax the last increment
    steps += 4 * steps_per
    ctx.set_physics(sc.phys)
    sc.extra["side_now"] = side
    return steps


def strained_vv(ctx, sc, nsteps, strain_rate, model=1):
    """Velocity-Verlet steps of the scene's force model with an isotropic
    affine strain rate applied once per step (x <- x (1 - rate dt), box
    likewise).  Returns the accumulated step statistics."""
    s = 1.0 - strain_rate * sc.phys.dt
    tot = None
    for _ in range(nsteps):
        if strain_rate != 0.0:
            ctx.affine_scale(s)
        st = ctx.step_vv(sc.detect, model, sc.params, 1)
        if tot is None:
            tot = dict(st)
        else:
            for key in tot:
                tot[key] = max(tot[key], st[key]) if key == "residual_max" else tot[key] + st[key]
    if strain_rate != 0.0:
        sc.extra["side_now"] = float(sc.extra.get("side_now", sc.extra["side"])) * s ** nsteps
    return tot


def packing_fraction_from_side(sc, side):
    vol = float(sum(math.pi * r * r * 2.0 * h + 4.0 / 3.0 * math.pi * r ** 3 for r, h in zip(sc.r, sc.h)))
    return vol / side ** 3
