"""Writes wedged_island.npz: the smallest contact island of the 1e5-rod
settling workload (hard contact, L/d 20, mu 0.5, omega 1.3, the workload's
tangential update and hot-set settings) that the island solve with the stall
rule leaves above the tolerance, at the first step from `--step` on where
one does, reduced to that island's rods.  Runs on a GPU (the settling
protocol up to the step, seconds on one MI355X); the island is chosen with
the FP64 oracle's restatement of the island solve.

Contents (island rods renumbered 0..nb-1 in ascending original index):
  step            the protocol step
  x, q, v, w      state at the end of step `step` - 1 (before the free-flight
                  update of step `step`)
  v_free, w_free  velocities after that free-flight update (the solve's start)
  r, h, mass      rod constants; box lo/hi, periodic x,y, floor and ceiling
  ab, rec, lam0   the step's manifolds of the island as the library built
                  them (natural order; b < 0 for walls), with the
                  previous-step impulses they start from
  order           the island's manifolds in the library's solve order
  friction_proj, hot_set, stall_sweeps
                  the solve settings of the workload
  gpu_sweeps, gpu_residual
                  what the library's solve did on the whole bed at this step

Usage: python tests/golden/make_wedged_island.py [--step 4189] [--out PATH]
"""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "rod-dynamics-3d_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def islands(ab, n):
    par = np.arange(n)

    def find(i):
        while par[i] != i:
            par[i] = par[par[i]]
            i = par[i]
        return i

    for a, b in ab:
        if b >= 0:
            ra, rb = find(a), find(b)
            if ra != rb:
                par[max(ra, rb)] = min(ra, rb)
    return np.array([find(a) for a in ab[:, 0]])


def main():
    import oracle_py as O
    from rod_dynamics_py import _lib as L
    from rod_dynamics_py import scenes

    ap = argparse.ArgumentParser()
    ap.add_argument("--step", type=int, default=4189)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "wedged_island.npz"))
    a = ap.parse_args()
    sc = scenes.c2_settling(n=100_000)
    p = sc.params
    p.check_every = 8
    det = L.detect_params(0.0, 16.0, 1.0)
    ctx = L.Context(0)
    sc.load(ctx)
    done = 0
    while done < a.step - 1:
        k = min(500, a.step - 1 - done)
        ctx.step_nsc(det, p, k)
        done += k
        print(f"step {done}", flush=True)
    d = sc.descs[0]
    op = O.nsc_params(mu=sc.extra["mu"], beta=0.0, cfm=0.0, omega=p.omega, tol=p.tol, max_sweeps=p.max_sweeps,
                      check_every=p.check_every, friction_proj=p.friction_proj, stop_rule=1, warm_start=1)
    # from `step` on, the first step whose solve leaves an island above the tolerance
    for step in range(a.step, a.step + 400):
        s_prev = ctx.download_state()
        ctx.integrate(L.RDK_NSC_FREE)
        s_free = ctx.download_state()
        ctx.detect(det)
        nm = ctx.nsc_build(p)
        rec, ab, lam0 = ctx.nsc_manifolds(nm)
        st = ctx.nsc_solve(p, allow_noconv=True)
        order = ctx.solve_order(nm)
        ctx.integrate(L.RDK_NSC_POS)
        if st["residual"] <= p.tol:
            continue
        # the oracle's island solve on every island, from the same start
        W = O.World(s_prev["x"], s_prev["q"], s_free["v"], s_free["w"], sc.r, sc.h, sc.mass)
        W.set_box(list(d.lo), list(d.hi), [1, 1, 0], 1, 1)
        W.set_physics(tuple(sc.phys.gravity[:]), sc.phys.dt)
        W.set_manifolds(rec, ab, lam0)
        W.apply_warm()
        lab = islands(ab, sc.n)
        pos = np.empty(nm, np.int64)
        pos[order] = np.arange(nm)
        best = None
        for isl in np.unique(lab):
            ix = np.where(lab == isl)[0]
            if len(ix) < 2 or (best is not None and len(ix) >= len(best)):
                continue
            ixo = np.ascontiguousarray(ix[np.argsort(pos[ix])], np.int32)
            _, res, _, _ = W.solve_island(op, ixo, stall_sweeps=p.stall_sweeps, stall_omega=p.stall_omega,
                                          hot=p.hot_set if len(ix) > 256 else 0)
            if not res <= p.tol:
                best = ix
        if best is not None:
            break
        print(f"step {step}: not converged on the GPU, every island converged in the oracle", flush=True)
    else:
        raise SystemExit(f"no island above the tolerance in steps {a.step}..{a.step + 399}")
    ctx.close()
    sel = best
    bodies = np.unique(np.concatenate([ab[sel, 0], ab[sel, 1][ab[sel, 1] >= 0]]))
    remap = -np.ones(sc.n, np.int64)
    remap[bodies] = np.arange(len(bodies))
    ab_l = ab[sel].copy()
    ab_l[:, 0] = remap[ab_l[:, 0]]
    rod = ab_l[:, 1] >= 0
    ab_l[rod, 1] = remap[ab_l[rod, 1]]
    local = -np.ones(nm, np.int64)
    local[sel] = np.arange(len(sel))
    order_l = local[sel[np.argsort(pos[sel])]]
    np.savez_compressed(a.out, step=step, x=s_prev["x"][bodies], q=s_prev["q"][bodies], v=s_prev["v"][bodies],
                        w=s_prev["w"][bodies], v_free=s_free["v"][bodies], w_free=s_free["w"][bodies],
                        r=sc.r[bodies], h=sc.h[bodies], mass=sc.mass[bodies], lo=np.array(d.lo[:]),
                        hi=np.array(d.hi[:]), ab=ab_l.astype(np.int32), rec=rec[sel], lam0=lam0[sel],
                        order=order_l.astype(np.int32), friction_proj=p.friction_proj, hot_set=p.hot_set,
                        stall_sweeps=p.stall_sweeps, gpu_sweeps=st["sweeps"], gpu_residual=st["residual"],
                        mu=sc.extra["mu"], gravity=np.array(sc.phys.gravity[:]), dt=sc.phys.dt)
    print(f"island of {len(sel)} manifolds, {len(bodies)} rods at step {step}; whole-bed solve: {st['sweeps']} "
          f"sweeps, residual {st['residual']:.3e} -> {a.out}", flush=True)


if __name__ == "__main__":
    main()
