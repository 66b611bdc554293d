#!/usr/bin/env python3
"""Benchmark: hard-contact PSOR settling (N=1e5 capsules, L/d=20) on one GPU.

A "step" is one full hard-contact time step of the packing: free flight,
cell-list broad phase + FP64 narrow phase, manifold build (rod-rod and
floor/ceiling walls), contact-graph colouring, coloured PSOR to the
fixed-point residual 1e-8, position/quaternion update.  The workload is the
settling protocol of the reference's hard-contact path: 1e5 rods, d=1, L=20,
random non-overlapping start at phi=0.05 in a 319^3 box periodic in x,y with
floor and ceiling walls, gravity -z, mu=0.5, e=0, omega=1.3, dt=1e-3,
5000 steps (warmup + timed).  State lives in HBM for the whole run.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rod-dynamics-3d_amd"))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per unit (ideal DRAM model, FP64 = 8 B, index = 4 B)
BYTES = {
    "broad_narrow": ("bodies", 72),      # x, q, r, h of every body once per pass
    "psor_sweep": ("sweep_contacts", 568),  # manifold 328 + lambda r/w 48 + 2x(v,w) gather 96 + scatter 96
    "residual": ("resid_contacts", 472),
    "manifold": ("manifolds", 696),
    "nsc_free": ("bodies", 280),
    "nsc_pos": ("bodies", 280),
    "contact_record": ("contacts", 112 + 144),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4980)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk-len", type=float, default=4.0)
    ap.add_argument("--check-every", type=int, default=4)
    ap.add_argument("--block", type=int, default=250)
    ap.add_argument("--ab-steps", type=int, default=20,
                    help="steps after the timed protocol solved from cold, previous-step and learned starts")
    ap.add_argument("--weights", default=os.path.join(ROOT, "rod-dynamics-3d_amd", "weights", "warmstart_c2.npy"))
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of this run
    give the scale factor between the counters and bytes moved."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": 4 * n, "launches": 5}), flush=True)


def cpu_baseline(sc, nsteps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    W = O.world_from_scene(sc)
    p = O.nsc_params(mu=sc.extra["mu"], beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=1,
                     friction_proj=sc.params.friction_proj, stop_rule=1, warm_start=1,
                     warm_walls=sc.params.warm_walls)
    W.nsc_step(p, 1, method=1)  # warm-up step (allocations)
    t0 = time.perf_counter()
    sweeps, rmax = W.nsc_step(p, nsteps, method=1)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sc.n * nsteps / dt, "unit": "rod-steps/s", "cores": cores, "kind": "port",
            "sample": f"steps 2..{nsteps + 1} of the same N={sc.n} settling run (same seed), FP64 oracle: reference "
                      f"spatial hash + serial sorted-order PSOR to residual 1e-8, OpenMP {cores} threads",
            "seconds": dt, "sweeps_per_step": sweeps / max(nsteps, 1)}


def learned_ab(ctx, sc, a):
    """Sweeps/step from the cold, previous-step and learned starts on the same
    manifolds and pre-solve velocities, over the steps that follow the timed
    protocol (the settled bed)."""
    import numpy as np
    from rod_dynamics_py import warmstart as WS

    if not os.path.exists(a.weights):
        return {"status": f"no trained weights at {os.path.relpath(a.weights, ROOT)}"}
    ctx.set_profiling(False)
    ctx.mlp_set_params(np.load(a.weights))
    rows = []
    for k in range(a.ab_steps):
        rows.append(WS.ab_step(ctx, sc.detect, sc.params))
        if (k + 1) % 5 == 0:
            print(f"[bench] learned-start A/B step {k + 1}/{a.ab_steps}", file=sys.stderr, flush=True)
    res = WS.ab_summary(rows)
    res["weights"] = os.path.relpath(a.weights, ROOT)
    res["bed"] = f"steps {a.warmup + a.steps + 1}..{a.warmup + a.steps + a.ab_steps} of the same run"
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (one MI355X per box); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if a.copy_calib > 0:
        copy_calibration(a.copy_calib)
        return
    from rod_dynamics_py import _lib as L
    from rod_dynamics_py import scenes

    t0 = time.perf_counter()
    sc = scenes.c2_settling(n=a.n)
    print(f"[bench] scene {sc.name} built in {time.perf_counter() - t0:.1f}s (box {sc.extra['side']:.2f})",
          file=sys.stderr, flush=True)
    p = sc.params
    p.check_every = a.check_every
    det = L.detect_params(0.0, a.chunk_len, 1.0)
    ctx = L.Context(0)
    sc.load(ctx)
    # warmup with every kernel timed, to rank the kernels
    ctx.set_profiling(True)
    ctx.step_nsc(det, p, a.warmup)
    kt = ctx.kernel_times()
    ranked = sorted(((ms, name) for name, (n, ms) in kt.items() if name in BYTES), reverse=True)
    dom = ranked[0][1] if ranked else "broad_narrow"
    names = [ctx._L.rdk_kernel_name(k).decode() for k in range(ctx._L.rdk_num_kernels())]
    ctx.set_profiling(True)  # resets counters
    ctx.set_profile_mask(1 << names.index(dom))
    ctx.sync()
    print(f"[bench] warmup {a.warmup} steps done; dominant kernel {dom}; timing {a.steps} steps", file=sys.stderr,
          flush=True)
    t1 = time.perf_counter()
    st = None
    done = 0
    while done < a.steps:  # blocks of steps, one progress line each (stderr)
        k = min(a.block, a.steps - done)
        s_ = ctx.step_nsc(det, p, k)
        done += k
        if st is None:
            st = dict(s_)
        else:
            for key in st:
                st[key] = max(st[key], s_[key]) if key == "residual_max" else st[key] + s_[key]
        print(f"[bench] step {a.warmup + done}: {time.perf_counter() - t1:.2f}s, manifolds/step "
              f"{s_['manifolds_total'] / k:.0f}, sweeps/step {s_['sweeps_total'] / k:.1f}", file=sys.stderr, flush=True)
    ctx.sync()
    el = time.perf_counter() - t1
    kt = ctx.kernel_times()
    wc = ctx.work_counters()
    launches, kms = kt[dom]
    unit, bpu = BYTES[dom]
    units = {"bodies": wc["body_passes"] * sc.n if dom == "broad_narrow" else launches * sc.n,
             "sweep_contacts": wc["sweep_contacts"], "resid_contacts": wc["resid_contacts"],
             "manifolds": wc["manifolds"], "contacts": wc["contacts"]}[unit]
    achieved = (bpu * units / 1e9) / (kms / 1e3) if kms > 0 else 0.0
    value = sc.n * a.steps / el
    out = {
        "metric": "rod-steps/s at PSOR residual 1e-8",
        "value": value,
        "unit": "rod-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "hard-contact PSOR settling (C2)", "rods": sc.n, "L_over_d": 20, "phi0": 0.05,
                   "box": round(sc.extra["side"], 3), "periodic": "x,y", "walls": "floor+ceiling", "mu": 0.5,
                   "restitution": 0.0, "omega": 1.3, "residual_tol": 1e-8, "dt": 1e-3,
                   "friction_projection": "euclidean disc" if p.friction_proj == 1 else "reference sequential",
                   "steps_total": a.warmup + a.steps, "seed": sc.extra["seed"], "check_every": a.check_every,
                   "chunk_len": a.chunk_len},
        "psor": {"sweeps_per_step": st["sweeps_total"] / max(st["steps"], 1),
                 "manifolds_per_step": st["manifolds_total"] / max(st["steps"], 1),
                 "rod_contacts_per_step": st["contacts_total"] / max(st["steps"], 1),
                 "colors_per_step": st["colors_total"] / max(st["steps"], 1),
                 "residual_max": st["residual_max"], "nonconverged_steps": st["nonconverged_steps"],
                 "warm_start": "previous step (rod-rod and plane-wall contacts)"},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launches": launches, "avg_launch_ms": kms / max(launches, 1),
                     "bytes_per_unit": bpu, "unit_name": unit, "units": units},
    }
    if a.ab_steps > 0:
        out["learned_warm_start"] = learned_ab(ctx, sc, a)
    ctx.close()
    if not a.no_cpu:
        print(f"[bench] cpu baseline: {a.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
