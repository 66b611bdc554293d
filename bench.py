#!/usr/bin/env python3
"""Benchmark of the rod-dynamics step on one GPU.  Prints one JSON line.

Default workload (--workload c2): hard-contact PSOR settling, the protocol the
headline metric is quoted on.  A "step" is one full hard-contact time step:
free flight, cell-list broad phase + FP64 narrow phase, manifold build (rod-rod
and floor/ceiling walls), contact islands and colouring, PSOR to the
fixed-point residual 1e-8, position/quaternion update.  1e5 rods, d=1, L=20,
random non-overlapping start at phi=0.05 in a 319^3 box periodic in x,y with
floor and ceiling walls, gravity -z, mu=0.5, e=0, omega=1.3, dt=1e-3, 5000
steps (20 warm-up + 4980 timed).  State lives in HBM for the whole run.  After
the timed protocol the settled bed is solved from the cold, previous-step and
learned starts (same manifolds and velocities) to report sweeps/step.

--workload c1: Hertz-Mindlin thermal gas, 1000 rods L/d=10, phi=0.10,
periodic, dt=1e-4, 10 000 steps (velocity Verlet, two force evaluations per
step as the reference integrator does).
--workload c3: dense polydisperse Hertz-Mindlin, 1e6 rods (L lognormal,
median 16), generated at phi=0.05, compressed to phi=0.30 (untimed
preparation), then 2000 timed steps at dt=5e-5 with an affine strain rate 1e-3.
--workload c4: 256 independent periodic hard-contact packings of uneven size
(N 500..20000, L/d 4..64, mu 0..1) stepped together, one launch per stage for
all systems, chunk length autotuned in the warm-up, 1000 steps.
--workload c5: the settling protocol at 2e5 rods from an evaluation seed, then
the settled bed solved from the cold, previous-step and learned starts (the
learned start trained on other seeds).
--workload populate: the reference's non-overlapping populate (serial random
sequential addition, FP32, std::mt19937 stream) reproduced on the GPU: 1e6
rods, L=20, d=1, periodic box at phi=0.05; rods placed per second, with the
serial oracle timed on a 1e5-rod sample of the same density (and the GPU
result on that sample compared with it bit for bit).
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
    "psor_sweep": ("psor_timed_updates", 568),  # manifold 328 + lambda r/w 48 + 2x(v,w) gather 96 + scatter 96
    "island_large": ("lds_manifold_sweeps", 376),  # manifold 328 + lambda r/w 48 (body velocities stay in LDS)
    "residual": ("resid_contacts", 472),
    "manifold": ("manifolds", 696),
    "nsc_free": ("bodies", 280),
    "nsc_pos": ("bodies", 280),
    "contact_record": ("contacts", 112 + 144),
    "force": ("force_contacts", 320),    # record 112 + x,q,v,w of two bodies 208
    "body_reduce": ("force_contacts", 192),
    "history": ("force_contacts", 64),
    "vv_half1": ("bodies", 280),
    "vv_half2": ("bodies", 280),
    "populate": ("candidates", 28),      # candidate centre + direction 24, decision 4
}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0, from a calibrated 1 GiB copy)
# for (workload, kernel); recorded with the command in profiles/
TRAFFIC = {("c2", "broad_narrow"): 11760.8 * 1024 * 2.0 + 153.9 * 1024,
           ("c4", "broad_narrow"): 127276.0 * 1024 * 2.0 + 3022.6 * 1024,
           ("c3", "broad_narrow"): 697756.0 * 1024 * 2.0 + 48165.0 * 1024,
           ("c2", "island_large"): 848.6 * 1024 * 2.0 + 383.1 * 1024}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["c2", "c1", "c3", "c4", "c5", "populate"], default="c2")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None, help="CPU baseline sample size (steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk-len", type=float, default=0.0, help="0: autotune over a few steps of the warm-up")
    ap.add_argument("--check-every", type=int, default=8,
                    help="sweeps between residual checks on the settling workloads (4: 14.4M, 8: 15.1M, 16: 15.4M "
                         "rod-steps/s at 1e5 rods, 0.7 / 1.4 %% more sweeps)")
    ap.add_argument("--warm-mode", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="override the settling workloads' PSOR start (1 previous step, 4 learned for every "
                         "manifold from the contact alone; learned modes load --weights)")
    ap.add_argument("--hot-set", type=int, default=None, choices=[0, 1, 2],
                    help="override the workload's hot-set phase of the LDS island solve (hard-contact workloads)")
    ap.add_argument("--friction-proj", type=int, default=None, choices=[0, 1, 2, 3],
                    help="override the hard-contact workloads' tangential PSOR update (1 scalar step on the "
                         "Euclidean disc, 2 block step projected in the tangential block's metric)")
    ap.add_argument("--stall-sweeps", type=int, default=None,
                    help="override the workload's stall-rule window (hard-contact workloads)")
    ap.add_argument("--step-graph", type=int, default=None, choices=[0, 1, 2],
                    help="hard-contact step: 0 eager (counts read back to the host), 1 fixed-capacity launches, "
                         "2 the fixed step captured once in a HIP graph and replayed (default 2; the settling "
                         "workloads' run 1, which times the kernels, is always eager)")
    ap.add_argument("--roof-steps", type=int, default=50,
                    help="batched workload with a captured step: eager steps after the timed run in which the "
                         "dominant kernel is timed with HIP events")
    ap.add_argument("--block", type=int, default=250)
    ap.add_argument("--protocol-steps", type=int, default=PROTOCOL_STEPS,
                    help="settling workloads: length of the protocol the timed blocks are spread over")
    ap.add_argument("--ab-steps", type=int, default=20,
                    help="steps after the timed protocol solved from cold, previous-step and learned starts")
    ap.add_argument("--weights", default=None,
                    help="learned-start weights (default: weights/warmstart_c2.npy, warmstart_c5.npy for --workload c5)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (HBM counter calibration)")
    a = ap.parse_args()
    defaults = {"c2": (4980, 100_000, 5), "c5": (4980, 200_000, 5), "c1": (9980, 1000, 2000), "c3": (2000, 1_000_000, 2),
                "c4": (980, 256, 8), "populate": (1, 1_000_000, 100_000)}[a.workload]
    if a.weights is None:
        a.weights = os.path.join(ROOT, "rod-dynamics-3d_amd", "weights",
                                 "warmstart_c5.npy" if a.workload == "c5" else "warmstart_c2.npy")
    if a.steps is None:
        a.steps = defaults[0]
    if a.n is None:
        a.n = defaults[1]
    if a.cpu_steps is None:
        a.cpu_steps = defaults[2]
    return a


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    return O


def cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def autotune(ctx, run_steps, cands, reps=3, blocks=(64, 128, 256, 512)):
    """Time `reps` steps (after one untimed step) for each chunk length in
    `cands`, then, at the best one, for each broad-phase workgroup size in
    `blocks`, then with the stencil rows staged in LDS at the best size; the
    contact set depends on none of them, so these are ordinary steps of the
    run.  Leaves the best workgroup size and staging set.  Returns (best chunk
    length, {"chunk_len_ms": ms per step by chunk length, "broad_block": best
    workgroup size, "broad_block_ms": ms per step by workgroup size,
    "broad_stage": 0/1, "broad_stage_ms": ms per step staged}, steps taken)."""
    from rod_dynamics_py import _lib as L

    def timed(cl):
        run_steps(cl, 1)
        ctx.sync()
        t = time.perf_counter()
        run_steps(cl, reps)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t) / reps

    tune = {cl: timed(cl) for cl in cands}
    best = min(tune, key=tune.get)
    btune = {}
    for b in blocks:
        ctx.set_tuning(L.TUNE_BROAD_BLOCK, b)
        btune[b] = timed(best)
    bbest = min(btune, key=btune.get)
    ctx.set_tuning(L.TUNE_BROAD_BLOCK, bbest)
    # stencil rows staged in LDS at the best workgroup size
    ctx.set_tuning(L.TUNE_BROAD_STAGE, 1)
    staged_ms = timed(best)
    stage = 1 if staged_ms < btune[bbest] else 0
    ctx.set_tuning(L.TUNE_BROAD_STAGE, stage)
    log(f"autotune chunk_len {best} ({', '.join(f'{k}: {v:.2f} ms' for k, v in tune.items())}); broad-phase "
        f"workgroup {bbest} ({', '.join(f'{k}: {v:.2f} ms' for k, v in btune.items())}); LDS-staged rows "
        f"{staged_ms:.2f} ms -> stage {stage}")
    return best, {"chunk_len_ms": {str(k): v for k, v in tune.items()}, "broad_block": bbest,
                  "broad_block_ms": {str(k): v for k, v in btune.items()}, "broad_stage": stage,
                  "broad_stage_ms": staged_ms}, (len(cands) + len(blocks) + 1) * (reps + 1)


def rank_kernels(ctx, run_warmup, fixed=None):
    """Warm-up with every kernel timed; returns the dominant kernel among those
    with a byte model (or `fixed`) and leaves only that kernel's timer on."""
    ctx.set_profiling(True)
    run_warmup()
    kt = ctx.kernel_times()
    ranked = sorted(((ms, name) for name, (n, ms) in kt.items() if name in BYTES), reverse=True)
    dom = fixed or (ranked[0][1] if ranked else "broad_narrow")
    names = [ctx._L.rdk_kernel_name(k).decode() for k in range(ctx._L.rdk_num_kernels())]
    ctx.set_profiling(True)  # resets counters
    ctx.set_profile_mask(1 << names.index(dom))
    ctx.sync()
    return dom


def roofline_from(dom, kt, wc, n, workload):
    """kt = (timed launches, summed event ms) of kernel `dom`; wc = work
    counters over the same launches."""
    launches, kms = kt
    unit, bpu = BYTES[dom]
    units = {"bodies": wc.get("body_passes", 0) * n if dom == "broad_narrow" else launches * n,
             "sweep_contacts": wc.get("sweep_contacts", 0), "resid_contacts": wc.get("resid_contacts", 0),
             "psor_timed_updates": wc.get("psor_timed_updates", 0),
             "lds_manifold_sweeps": wc.get("lds_manifold_sweeps", 0),
             "manifolds": wc.get("manifolds", 0), "contacts": wc.get("contacts", 0),
             "force_contacts": wc.get("force_contacts", 0)}[unit]
    achieved = (bpu * units / 1e9) / (kms / 1e3) if kms > 0 else 0.0
    return {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC.get((workload, dom)), "launches": launches,
            "avg_launch_ms": kms / max(launches, 1), "bytes_per_unit": bpu, "unit_name": unit, "units": units}


def roofline(ctx, dom, n, workload):
    return roofline_from(dom, ctx.kernel_times()[dom], ctx.work_counters(), n, workload)


def base_line(metric, value, a, el, dtype, config):
    return {"metric": metric, "value": value, "unit": "rod-steps/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


# ---------------------------------------------------------------------------- C2
PROTOCOL_STEPS = 5000  # the settling protocol: 5000 steps at dt = 1e-3


def phase_blocks(total, warm, k, max_blocks=20):
    """The timed steps of a protocol of `total` steps whose first `warm` are
    warm-up.  When warm + k covers the protocol, one block of every step after
    the warm-up.  Otherwise the k steps are split into B equal blocks (B the
    largest divisor of k not above max_blocks), block b centred in the b-th of
    B equal strata of steps warm..total, so that the mean step time of the
    blocks estimates the mean over the protocol (stratified sample)."""
    span = total - warm
    if k >= span:
        return [(warm, span)]
    nb = max(d for d in range(1, min(k, max_blocks) + 1) if k % d == 0)
    size = k // nb
    out = []
    for b in range(nb):
        s = int(warm + (b + 0.5) * span / nb - size / 2)
        s = max(s, out[-1][0] + out[-1][1] if out else warm)
        out.append((s, size))
    return out


def advance(ctx, det, p, cur, target, t_ref, block=250, acc=None):
    """Steps cur..target, one progress line per `block` steps; the step
    statistics are added into `acc` when given."""
    first = cur
    while cur < target:
        k = min(block, target - cur)
        st = ctx.step_nsc(det, p, k)
        if acc is not None:
            for key, val in st.items():
                acc[key] = max(acc.get(key, 0.0), val) if key == "residual_max" else acc.get(key, 0) + val
        cur += k
        if k == block or (cur == target and target - first >= 10):
            log(f"  step {cur} ({time.perf_counter() - t_ref:.1f}s)")
    return cur


def cpu_baseline_c2(sc, snaps, total_steps, O):
    """The FP64 oracle (reference spatial hash + serial sorted-order PSOR to
    1e-8, OpenMP elsewhere) from GPU-state snapshots at the start of timed
    blocks (one per stratum of the protocol): at each, one unmeasured step
    (fills the warm cache), then one timed step; the mean step time over the
    strata is the protocol's mean step time."""
    p = O.nsc_params(mu=sc.extra["mu"], beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=1,
                     friction_proj=sc.params.friction_proj, stop_rule=1, warm_start=1,
                     warm_walls=sc.params.warm_walls)
    W = O.world_from_scene(sc)
    per = []
    t_all = 0.0
    for step, st in snaps:
        W.set_state(x=st["x"], q=st["q"], v=st["v"], w=st["w"])
        W.nsc_step(p, 1, method=1)
        t0 = time.perf_counter()
        sweeps, _ = W.nsc_step(p, 1, method=1)
        dt = time.perf_counter() - t0
        t_all += dt
        per.append({"step": step + 2, "seconds": dt, "sweeps": sweeps})
        log(f"cpu baseline: step {step + 2}: {dt:.2f}s, {sweeps} sweeps")
    mean = sum(r["seconds"] for r in per) / len(per)
    return {"value": sc.n / mean, "unit": "rod-steps/s", "cores": cores(), "kind": "port",
            "sample": f"FP64 oracle (reference spatial hash + serial sorted-order PSOR to residual 1e-8, OpenMP "
                      f"{cores()} threads): one timed step (after one unmeasured step) from the GPU state at each of "
                      f"{len(per)} points spread evenly over the {total_steps}-step protocol (the starts of timed "
                      f"blocks); value = rods / mean step time",
            "seconds_measured": t_all, "per_snapshot": per}


def run_c2(a, L, scenes):
    """The settling protocol; --workload c5 runs it at N = 2e5 from an evaluation
    seed (the learned start is trained on other seeds) and then compares the
    starts on the settled bed.

    The protocol runs twice from the same start; the trajectory is
    deterministic (bitwise, checked on the counts of every timed block).  Run 1
    (untimed) times every kernel inside the timed blocks to rank them and takes
    the state snapshots of the CPU baseline.  Run 2 times the --steps steps as
    equal blocks at evenly spaced phases of the protocol (phase_blocks), with
    only the dominant kernel's HIP events recorded, and continues to the end of
    the protocol, where the learned-start comparison runs on the settled bed."""
    t0 = time.perf_counter()
    c5 = a.workload == "c5"
    if c5:
        from rod_dynamics_py import warmstart as WS

        sc = scenes.c2_settling(n=a.n, seed=WS.EVAL_SEED0)
    else:
        sc = scenes.c2_settling(n=a.n)
    log(f"scene {sc.name} built in {time.perf_counter() - t0:.1f}s (box {sc.extra['side']:.2f})")
    p = sc.params
    p.check_every = a.check_every
    if a.stall_sweeps is not None:
        p.stall_sweeps = a.stall_sweeps
    if a.hot_set is not None:
        p.hot_set = a.hot_set
    if a.friction_proj is not None:
        p.friction_proj = a.friction_proj
    if a.warm_mode is not None:
        p.warm_mode = a.warm_mode
    weights = None
    if p.warm_mode >= 2:
        import numpy as np

        weights = np.load(a.weights)
    total = a.protocol_steps

    # ---- run 1: warm-up (chunk length autotune), ranking and CPU snapshots
    ctx = L.Context(0)
    names = [ctx._L.rdk_kernel_name(k).decode() for k in range(ctx._L.rdk_num_kernels())]
    sc.load(ctx)
    if weights is not None:
        ctx.mlp_set_params(weights)
    if a.chunk_len > 0:
        best, tune, used = a.chunk_len, {}, 0
    else:  # --chunk-len 0: autotune inside the warm-up
        best, tune, used = autotune(ctx, lambda cl, k: ctx.step_nsc(L.detect_params(0.0, cl, 1.0), p, k),
                                    (2.0, 4.0, 8.0, 16.0), reps=min(3, max(1, a.warmup // 3 - 1)))
    det = L.detect_params(0.0, best, 1.0)
    warm = max(a.warmup, used)
    if warm > used:
        ctx.step_nsc(det, p, warm - used)
    blocks = phase_blocks(total, warm, a.steps)
    cpu_at = set()
    if not a.no_cpu:
        idx = range(len(blocks)) if len(blocks) <= 10 else [int((j + 0.5) * len(blocks) / 10) for j in range(10)]
        cpu_at = {blocks[j][0] for j in idx}
        if len(blocks) == 1:  # whole protocol timed: five points inside it
            cpu_at = {warm + int((j + 0.5) * (total - warm) / 5) for j in range(5)}
    log(f"warm-up {warm} steps; {len(blocks)} timed block(s) of {blocks[0][1]} step(s); run 1 (ranking)")
    starts = {s: k for s, k in blocks}
    ends = {s + k for s, k in blocks}
    marks = sorted(set(starts) | ends | cpu_at)
    rank_ms = {}
    rank_kt = {}  # kernel -> [timed launches, ms] inside the timed blocks
    run1 = []
    run1_wc = {}
    acc = None
    cur = warm
    snaps = []
    wc_names = ("manifolds", "sweep_contacts", "resid_contacts", "psor_timed_updates", "lds_manifold_sweeps",
                "contacts", "body_passes", "force_contacts")
    wcb = None
    # Outside the timed blocks the kernels with a byte model are timed too, so
    # run 1 also gives each one's per-launch average over the whole protocol
    # (the figure rocprofv3 statistics of the same command reproduce).
    cand_mask = sum(1 << names.index(k) for k in BYTES if k in names)
    all_mask = (1 << 64) - 1
    proto_kt = {}

    def absorb(into):
        for nm_, (n_, ms_) in ctx.kernel_times().items():
            kt_ = into.setdefault(nm_, [0, 0.0])
            kt_[0] += n_
            kt_[1] += ms_

    ctx.set_profiling(True)
    ctx.set_profile_mask(cand_mask)
    t1 = time.perf_counter()
    for m in marks:
        cur = advance(ctx, det, p, cur, m, t1, acc=acc)
        if m in ends:  # a timed block ends here
            for nm_, (n_, ms_) in ctx.kernel_times().items():
                rank_ms[nm_] = rank_ms.get(nm_, 0.0) + ms_
            absorb(rank_kt)
            absorb(proto_kt)
            wce = ctx.work_counters()
            for key in wc_names:
                run1_wc[key] = run1_wc.get(key, 0) + wce[key] - wcb[key]
            ctx.set_profiling(True)  # resets the counters
            ctx.set_profile_mask(cand_mask)
            run1.append(acc)
            acc = None
        if m in cpu_at:
            snaps.append((m, ctx.download_state()))
        if m in starts:  # a timed block starts here: every kernel timed (resets the counters)
            absorb(proto_kt)
            ctx.set_profiling(True)
            ctx.set_profile_mask(all_mask)
            wcb = ctx.work_counters()
            acc = {}
    absorb(proto_kt)
    ctx.set_profiling(False)
    ctx.close()
    ranked = sorted(((ms, nm_) for nm_, ms in rank_ms.items() if nm_ in BYTES), reverse=True)
    dom = ranked[0][1] if ranked else "broad_narrow"
    share = {nm_: round(ms, 3) for ms, nm_ in sorted(((ms, nm_) for nm_, ms in rank_ms.items()), reverse=True)[:8]}
    log(f"run 1 done in {time.perf_counter() - t1:.1f}s; dominant kernel in the timed blocks: {dom}; "
        f"kernel ms in the blocks: {share}")

    # ---- run 2: the same protocol, timed blocks
    mode = (0 if p.warm_mode >= 2 else 2) if a.step_graph is None else a.step_graph
    ctx = L.Context(0)
    sc.load(ctx)
    if tune:
        ctx.set_tuning(L.TUNE_BROAD_BLOCK, tune["broad_block"])
        ctx.set_tuning(L.TUNE_BROAD_STAGE, tune["broad_stage"])
    if weights is not None:
        ctx.mlp_set_params(weights)
    if mode == 0:
        ctx.set_profiling(True)
        ctx.set_profile_mask(1 << names.index(dom))
    ctx.step_nsc(det, p, warm)
    if mode != 0:  # the captured steps: the dominant kernel's launches were timed in run 1
        ctx.set_step_graph(mode)
    cur = warm
    acc_wc = {k: 0 for k in wc_names}
    acc_kt = [0, 0.0]
    el = 0.0
    phases = []
    identical = True
    t2 = time.perf_counter()
    log("run 2 (timed)")
    for (s0, k), ref in zip(blocks, run1):
        cur = advance(ctx, det, p, cur, s0, t2)
        kt0 = ctx.kernel_times()[dom]
        wc0 = ctx.work_counters()
        g0 = ctx.step_graph_info()
        ctx.sync()
        tb = time.perf_counter()
        st = ctx.step_nsc(det, p, k)
        ctx.sync()
        dt = time.perf_counter() - tb
        g1 = ctx.step_graph_info()
        el += dt
        cur += k
        kt1 = ctx.kernel_times()[dom]
        wc1 = ctx.work_counters()
        acc_kt[0] += kt1[0] - kt0[0]
        acc_kt[1] += kt1[1] - kt0[1]
        for key in wc_names:
            acc_wc[key] += wc1[key] - wc0[key]
        same = all(st[key] == ref[key] for key in ("manifolds_total", "sweeps_total", "contacts_total",
                                                    "nonconverged_steps"))
        identical &= same
        phases.append({"step": s0 + 1, "steps": k, "ms_per_step": 1e3 * dt / k,
                       "manifolds_per_step": st["manifolds_total"] / k, "sweeps_per_step": st["sweeps_total"] / k,
                       "rod_contacts_per_step": st["contacts_total"] / k, "colors_per_step": st["colors_total"] / k,
                       "nonconverged_steps": st["nonconverged_steps"], "residual_max": st["residual_max"],
                       "eager_steps": g1["fallbacks"] - g0["fallbacks"], "captures": g1["captures"] - g0["captures"]})
        log(f"timed block at step {s0 + 1}: {1e3 * dt / k:.2f} ms/step, manifolds/step "
            f"{st['manifolds_total'] / k:.0f}, sweeps/step {st['sweeps_total'] / k:.1f}, residual "
            f"{st['residual_max']:.2e}{'' if same else ' (differs from run 1)'}")
    cur = advance(ctx, det, p, cur, total, t2)  # the rest of the protocol (untimed)
    ginfo = ctx.step_graph_info()
    ctx.set_step_graph(0)
    kt_all = ctx.kernel_times()[dom]
    steps = sum(k for _, k in blocks)
    nonconv = sum(ph["nonconverged_steps"] for ph in phases)
    mpf = sum(ph["manifolds_per_step"] * ph["steps"] for ph in phases) / steps
    if mpf <= 0:
        raise SystemExit("bench.py: the timed blocks held no contact manifolds; no PSOR was measured")
    converged = steps - nonconv
    out = base_line("rod-steps/s at PSOR residual 1e-8", sc.n * converged / el, a, el, "f64", {
        "workload": ("learned-start bed: hard-contact settling, evaluation seed (C5)" if c5 else
                     "hard-contact PSOR settling (C2)"), "rods": sc.n, "L_over_d": 20, "phi0": 0.05,
        "box": round(sc.extra["side"], 3), "periodic": "x,y", "walls": "floor+ceiling", "mu": 0.5,
        "restitution": 0.0, "omega": 1.3, "residual_tol": 1e-8, "dt": 1e-3,
        "friction_projection": {0: "reference sequential", 1: "euclidean disc, scalar step",
                                2: "euclidean disc, block step in the tangential metric",
                                3: "euclidean disc, block step on sticking contacts, scalar step on sliding ones"}[
                                    p.friction_proj],
        "protocol_steps": total, "warmup_steps": warm,
        "timed_blocks": f"{len(blocks)} block(s) of {blocks[0][1]} step(s), centred in equal strata of steps "
                        f"{warm + 1}..{total}" if len(blocks) > 1 else f"steps {warm + 1}..{total}",
        "seed": sc.extra["seed"], "check_every": a.check_every, "chunk_len": best, "autotune": tune,
        "stall_sweeps": p.stall_sweeps, "stall_omega": p.stall_omega, "hot_set": p.hot_set,
        "warm_mode": p.warm_mode, "step_mode": mode,
        "hipgraph": ("run 2 replays one captured fixed-capacity step per step (every count on the device); steps "
                     "it cannot finish (an island beyond the LDS class, more contacts than reserved) are rolled back "
                     "and run eagerly, whose huge-island check blocks replay their own captured graphs"
                     if mode == 2 else "eager step: counts read back to the host; check blocks of the huge-island "
                     "solve captured per island and step, replayed"),
        "fixed_steps": ginfo["fixed_steps"], "fallback_steps": ginfo["fallbacks"],
        "graph_captures": ginfo["captures"]})
    out["value_counts"] = (f"rods x timed steps that reached residual 1e-8 ({converged} of {steps}) / wall time of "
                           f"all {steps} timed steps")
    out["psor"] = {"sweeps_per_step": sum(ph["sweeps_per_step"] * ph["steps"] for ph in phases) / steps,
                   "manifolds_per_step": mpf,
                   "rod_contacts_per_step": sum(ph["rod_contacts_per_step"] * ph["steps"] for ph in phases) / steps,
                   "residual_max_converged": max([ph["residual_max"] for ph in phases if ph["nonconverged_steps"] == 0]
                                                 or [0.0]),
                   "nonconverged_steps": nonconv,
                   "nonconverged_residuals": [ph["residual_max"] for ph in phases if ph["nonconverged_steps"] > 0],
                   "warm_start": {1: "previous step (rod-rod and plane-wall contacts)",
                                  4: f"learned for every manifold from the contact alone "
                                     f"({os.path.relpath(a.weights, ROOT)})"}.get(p.warm_mode, p.warm_mode)}
    out["phases"] = phases
    out["replay_identical"] = identical
    if mode == 0:
        out["roofline"] = roofline_from(dom, acc_kt, acc_wc, sc.n, a.workload)
        out["roofline"]["protocol_launches"] = kt_all[0]
        out["roofline"]["protocol_avg_launch_ms"] = kt_all[1] / max(kt_all[0], 1)
    else:  # run 1's eager launches of the same steps (counts checked equal above)
        out["roofline"] = roofline_from(dom, tuple(rank_kt.get(dom, (0, 0.0))), run1_wc, sc.n, a.workload)
        out["roofline"]["timed_in"] = ("run 1: HIP events around the dominant kernel's eager launches in the same "
                                       "timed blocks (a captured step cannot be split into kernels by events)")
        pk = proto_kt.get(dom, [0, 0.0])
        out["roofline"]["protocol_launches"] = pk[0]
        out["roofline"]["protocol_avg_launch_ms"] = pk[1] / max(pk[0], 1)
    out["kernel_ms_in_timed_blocks_run1"] = share
    if a.ab_steps > 0:
        out["learned_warm_start"] = learned_ab(ctx, sc, a, total)
    ctx.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c2(sc, snaps, total, oracle())
        out["cpu_baseline"]["ratio_gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def learned_ab(ctx, sc, a, total):
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
            log(f"learned-start A/B step {k + 1}/{a.ab_steps}")
    res = WS.ab_summary(rows)
    meta = os.path.splitext(a.weights)[0] + ".json"
    select = json.load(open(meta)).get("select", "uncached") if os.path.exists(meta) else "uncached"
    if select != "all":
        # the start for every manifold needs weights trained on every manifold
        for key in [k for k in res if "learned_all" in k]:
            del res[key]
        res["learned_all"] = {"status": "not reported: these weights were trained on manifolds without a "
                                        "previous-step impulse only"}
    res["weights"] = os.path.relpath(a.weights, ROOT)
    res["weights_trained_on"] = select
    res["bed"] = f"steps {total + 1}..{total + a.ab_steps} of the same run (after the protocol)"
    return res


# ---------------------------------------------------------------------- C1 / C3
def cpu_baseline_vv(sc, state, nsteps, O, note):
    W = O.world_from_scene(sc)
    W.set_state(x=state["x"], q=state["q"], v=state["v"], w=state["w"])
    hp = sc.params
    p = O.hm_params(E=hp.E, nu_poisson=hp.nu_poisson, restitution=hp.restitution, mu_k=hp.mu_k, mu_s=hp.mu_s,
                    v_c=hp.v_c, mu_roll=hp.mu_roll, enable_tangential=hp.enable_tangential,
                    enable_rolling=hp.enable_rolling, law=hp.law)
    W.vv_step(1, p, 1, method=1)  # history and allocations
    t0 = time.perf_counter()
    W.vv_step(1, p, nsteps, method=1)
    dt = time.perf_counter() - t0
    return {"value": sc.n * nsteps / dt, "unit": "rod-steps/s", "cores": cores(), "kind": "port",
            "sample": f"{nsteps} velocity-Verlet Hertz-Mindlin steps of the FP64 oracle (reference spatial hash, "
                      f"two force evaluations per step, OpenMP {cores()} threads) from {note}",
            "seconds": dt}


def run_vv(a, L, scenes, which):
    from rod_dynamics_py import protocols

    t0 = time.perf_counter()
    if which == "c1":
        sc = scenes.c1_thermal(n=a.n, model="hm")
    else:
        sc = scenes.c3_dense(n=a.n)
    sc.detect = L.detect_params(0.0, a.chunk_len, 1.0)
    log(f"scene {sc.name} built in {time.perf_counter() - t0:.1f}s (box {sc.extra['side']:.2f})")
    ctx = L.Context(0)
    sc.load(ctx)
    prep = {}
    if which == "c3":
        t0 = time.perf_counter()
        nsteps = protocols.compress_nsc(ctx, sc, sc.extra["side1"], log=sys.stderr)
        st0 = ctx.download_state()
        v, w = L.gen_thermal(st0["q"], sc.r, sc.h, sc.mass, 1.0, sc.extra["seed"] + 29)
        ctx.upload_state(v=v, w=w)
        prep = {"compression": f"affine increments of 0.998 with two overdamped hard-contact steps each "
                               f"({nsteps} steps, beta 0.8, dt 1e-2) from phi 0.05 to phi "
                               f"{protocols.packing_fraction_from_side(sc, sc.extra['side_now']):.4f}, then thermal "
                               f"velocities T=1", "prep_seconds": time.perf_counter() - t0}
        log(f"compressed to side {sc.extra['side_now']:.3f} in {prep['prep_seconds']:.1f}s")
    rate = 1e-3 if which == "c3" else 0.0
    tune, used = {}, 0
    if a.chunk_len <= 0:  # autotune inside the warm-up
        def run(cl, k):
            sc.detect = L.detect_params(0.0, cl, 1.0)
            protocols.strained_vv(ctx, sc, k, rate)

        best, tune, used = autotune(ctx, run, (1.0, 2.0, 4.0) if which == "c3" else (1.0, 2.0, 4.0, 10.0),
                                    reps=min(3, max(1, a.warmup // 4 - 1)))
        sc.detect = L.detect_params(0.0, best, 1.0)
    dom = rank_kernels(ctx, lambda: protocols.strained_vv(ctx, sc, max(1, a.warmup - used), rate))
    log(f"warmup {a.warmup} steps done; dominant kernel {dom}; timing {a.steps} steps")
    done = 0
    st = None
    t1 = time.perf_counter()
    while done < a.steps:
        k = min(a.block, a.steps - done)
        s_ = protocols.strained_vv(ctx, sc, k, rate) if rate != 0.0 else ctx.step_vv(sc.detect, 1, sc.params, k)
        done += k
        if st is None:
            st = dict(s_)
        else:
            for key in st:
                st[key] = max(st[key], s_[key]) if key == "residual_max" else st[key] + s_[key]
        log(f"step {a.warmup + done}: {time.perf_counter() - t1:.2f}s, contacts/evaluation "
            f"{s_['contacts_total'] / (2 * k):.0f}")
    ctx.sync()
    el = time.perf_counter() - t1
    if which == "c1":
        cfg = {"workload": "Hertz-Mindlin thermal gas (C1)", "rods": sc.n, "L_over_d": 10, "phi": 0.10,
               "box": round(sc.extra["side"], 3), "periodic": "x,y,z", "mu": 0.5, "E": 1e7, "nu": 0.25,
               "restitution": 0.5, "dt": 1e-4, "steps_total": max(a.warmup, used + 1) + a.steps,
               "seed": sc.extra["seed"], "force_evaluations_per_step": 2, "chunk_len": sc.detect.chunk_len,
               "autotune": tune}
    else:
        cfg = {"workload": "dense polydisperse Hertz-Mindlin (C3)", "rods": sc.n,
               "L": "lognormal median 16 sigma_log 0.3, truncated [4, 64]", "mean_L": round(sc.extra["mean_length"], 3),
               "phi": 0.30, "box": round(sc.extra["side_now"], 3), "periodic": "x,y,z", "mu": 0.3, "E": 1e7,
               "nu": 0.25, "restitution": 0.5, "dt": 5e-5, "strain_rate": 1e-3, "steps_timed": a.steps,
               "seed": sc.extra["seed"], "force_evaluations_per_step": 2, "chunk_len": sc.detect.chunk_len,
               "autotune": tune}
        cfg.update(prep)
    out = base_line("rod-steps/s", sc.n * a.steps / el, a, el, "f64", cfg)
    out["contacts_per_evaluation"] = st["contacts_total"] / (2 * max(st["steps"], 1))
    wc = ctx.work_counters()
    out["narrow_phase_candidates_per_evaluation"] = wc["candidates"] / max(wc["body_passes"], 1)
    out["roofline"] = roofline(ctx, dom, sc.n, which)
    state = ctx.download_state()
    ctx.close()
    if not a.no_cpu:
        log(f"cpu baseline: {a.cpu_steps} oracle steps")
        out["cpu_baseline"] = cpu_baseline_vv(sc, state, a.cpu_steps, oracle(),
                                              "the GPU state after the timed steps of the same run")
    return out


# ---------------------------------------------------------------------------- C4
def cpu_baseline_c4(sc, state, nsample, O):
    """Systems run one after another on the FP64 oracle with all cores, as the
    reference would: a size-stratified sample of systems, one unmeasured and
    one timed step each from the GPU state after the run."""
    import numpy as np

    sizes = np.diff(sc.offsets)
    order = np.argsort(sizes)
    pick = [int(order[int((k + 0.5) * len(order) / nsample)]) for k in range(nsample)]
    rods, secs, per = 0, 0.0, []
    for k in pick:
        a0, a1 = int(sc.offsets[k]), int(sc.offsets[k + 1])
        d = sc.descs[k]
        W = O.World(state["x"][a0:a1], state["q"][a0:a1], state["v"][a0:a1], state["w"][a0:a1], sc.r[a0:a1],
                    sc.h[a0:a1], sc.mass[a0:a1], sc.rest[a0:a1])
        W.set_box(list(d.lo), list(d.hi), [int(v) for v in d.periodic], 0, 0)
        W.set_physics(tuple(sc.phys.gravity[:]), sc.phys.dt)
        p = O.nsc_params(mu=d.mu, beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=1,
                         friction_proj=sc.params.friction_proj, stop_rule=1, warm_start=1)
        W.nsc_step(p, 1, method=1)
        t0 = time.perf_counter()
        sw, _ = W.nsc_step(p, 1, method=1)
        dt = time.perf_counter() - t0
        rods += a1 - a0
        secs += dt
        per.append({"system": k, "rods": a1 - a0, "seconds": dt, "sweeps": sw})
        log(f"cpu baseline: system {k} ({a1 - a0} rods): {dt:.3f}s")
    return {"value": rods / secs, "unit": "rod-steps/s", "cores": cores(), "kind": "port",
            "sample": f"{nsample} systems (size-stratified) run one after another on the FP64 oracle (reference "
                      f"spatial hash + serial sorted-order PSOR to 1e-8, OpenMP {cores()} threads), one timed step "
                      f"each after one unmeasured step, from the GPU state after the run",
            "seconds": secs, "per_system": per}


def run_c4(a, L, scenes):
    t0 = time.perf_counter()
    sc = scenes.c4_batched(nsys=a.n, threads=min(16, os.cpu_count() or 1))
    log(f"scene {sc.name} built in {time.perf_counter() - t0:.1f}s ({sc.n} rods, {sc.extra['redraws']} redraws)")
    p = sc.params
    if a.stall_sweeps is not None:
        p.stall_sweeps = a.stall_sweeps
    if a.hot_set is not None:
        p.hot_set = a.hot_set
    if a.friction_proj is not None:
        p.friction_proj = a.friction_proj
    ctx = L.Context(0)
    sc.load(ctx)
    best, tune, used = autotune(ctx, lambda cl, k: ctx.step_nsc(L.detect_params(0.0, cl, 1.0), p, k),
                                (2.0, 4.0, 8.0, 16.0), reps=3)
    det = L.detect_params(0.0, best, 1.0)
    dom = rank_kernels(ctx, lambda: ctx.step_nsc(det, p, a.warmup))
    mode = 2 if a.step_graph is None else a.step_graph
    if mode != 0:
        # captured steps cannot be split into kernels by events: the dominant
        # kernel is timed in eager steps after the timed run
        ctx.set_profiling(False)
        ctx.set_step_graph(mode)
        ctx.step_nsc(det, p, 1)  # capture (or first fixed step) outside the timed region
        ctx.sync()
    log(f"warmup {a.warmup} steps done; dominant kernel {dom}; timing {a.steps} steps (step mode {mode})")
    st = None
    done = 0
    t1 = time.perf_counter()
    while done < a.steps:
        k = min(a.block, a.steps - done)
        s_ = ctx.step_nsc(det, p, k)
        done += k
        if st is None:
            st = dict(s_)
        else:
            for key in st:
                st[key] = max(st[key], s_[key]) if key == "residual_max" else st[key] + s_[key]
        log(f"step {done}: {time.perf_counter() - t1:.2f}s, manifolds/step {s_['manifolds_total'] / k:.0f}, "
            f"sweeps/step {s_['sweeps_total'] / k:.1f}")
    ctx.sync()
    el = time.perf_counter() - t1
    tuned_steps = used
    ginfo = ctx.step_graph_info()
    roof_note = "timed launches inside the timed run"
    eager_ms = None
    if mode != 0:
        ctx.set_step_graph(0)
        names = [ctx._L.rdk_kernel_name(k).decode() for k in range(ctx._L.rdk_num_kernels())]
        ctx.set_profiling(True)
        ctx.set_profile_mask(1 << names.index(dom))
        t2 = time.perf_counter()
        ctx.step_nsc(det, p, a.roof_steps)
        ctx.sync()
        eager_ms = 1e3 * (time.perf_counter() - t2) / max(a.roof_steps, 1)
        roof_note = (f"timed launches of {a.roof_steps} eager steps after the timed run (same kernels; a captured "
                     f"step cannot be split into kernels by events)")
    out = base_line("rod-steps/s at PSOR residual 1e-8", sc.n * a.steps / el, a, el, "f64", {
        "workload": "batched hard-contact packings (C4)", "systems": len(sc.descs), "rods": sc.n,
        "N": "floor(exp(U[ln500, ln20000]))", "L_over_d": "{4,8,16,32,64}", "mu": "U[0,1]",
        "phi": "U[min(0.05,cap), cap], cap = min(0.25, 0.85 x RSA saturation)", "periodic": "x,y,z",
        "restitution": sc.extra["restitution"], "kT": sc.extra["kT"], "gravity": 0.0, "omega": 1.3,
        "residual_tol": 1e-8, "dt": 1e-3, "redraws": sc.extra["redraws"], "seed": sc.extra["seed"],
        "chunk_len_autotuned": best, "autotune": tune,
        "steps_total": tuned_steps + a.warmup + a.steps + (1 + a.roof_steps if mode != 0 else 0),
        "stall_sweeps": p.stall_sweeps,
        "stall_omega": p.stall_omega,
        "step_mode": mode,
        "hipgraph": ("each step replays one captured fixed-capacity step (every count on the device)"
                     if mode == 2 else "fixed-capacity launches, not captured" if mode == 1 else
                     "eager: the step reads contact and island counts back to size its launches"),
        "fixed_steps": ginfo["fixed_steps"], "fallback_steps": ginfo["fallbacks"],
        "graph_captures": ginfo["captures"], "eager_ms_per_step_after": eager_ms,
        "roofline_sample": roof_note})
    out["psor"] = {"sweeps_per_step": st["sweeps_total"] / max(st["steps"], 1),
                   "manifolds_per_step": st["manifolds_total"] / max(st["steps"], 1),
                   "residual_max": st["residual_max"], "nonconverged_steps": st["nonconverged_steps"]}
    out["roofline"] = roofline(ctx, dom, sc.n, "c4")
    state = ctx.download_state()
    ctx.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c4(sc, state, a.cpu_steps, oracle())
    return out


# ---------------------------------------------------------------------- populate
def run_populate(a, L):
    import math

    import numpy as np

    length, dia, phi, seed, att = 20.0, 1.0, 0.05, 1592590338, 1000
    vol = math.pi * (dia / 2) ** 2 * length + 4.0 / 3.0 * math.pi * (dia / 2) ** 3

    def side(n):
        return (n * vol / phi) ** (1.0 / 3.0)

    ctx = L.Context(0)
    ctx.populate_nonoverlap(1000, length, dia, [0, 0, 0], [side(1000)] * 3, 1, 0.0, seed, att)  # warm-up
    ctx.set_profiling(True)
    box = side(a.n)
    t0 = time.perf_counter()
    c, u, q, drawn = ctx.populate_nonoverlap(a.n, length, dia, [0, 0, 0], [box] * 3, 1, 0.0, seed, att)
    el = time.perf_counter() - t0
    log(f"placed {len(c)} of {a.n} rods from {drawn} candidates in {el:.2f}s")
    kt = ctx.kernel_times()["populate"]
    out = {"metric": "rods placed/s (non-overlapping populate)", "value": len(c) / el, "unit": "rods/s",
           "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * el, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "reference populate nonoverlap", "rods": a.n, "placed": len(c),
                      "candidates": drawn, "length": length, "diameter": dia, "phi": phi, "box": box,
                      "periodic": "x,y,z", "cell_size": "max(L/4, 1.25 d)", "seed": seed, "max_attempts": att}}
    achieved = (BYTES["populate"][1] * drawn / 1e9) / (kt[1] / 1e3) if kt[1] > 0 else 0.0
    out["roofline"] = {"kernel": "populate", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None, "launches": kt[0],
                       "avg_launch_ms": kt[1] / max(kt[0], 1), "bytes_per_unit": BYTES["populate"][1],
                       "unit_name": "candidates", "units": drawn}
    if not a.no_cpu:
        O = oracle()
        ns = a.cpu_steps
        t1 = time.perf_counter()
        k, oc, ou, odrawn = O.populate_nonoverlap(ns, length, dia, [0, 0, 0], [side(ns)] * 3, 1, 0.0, seed, att)
        cel = time.perf_counter() - t1
        gc, gu, _, gdrawn = ctx.populate_nonoverlap(ns, length, dia, [0, 0, 0], [side(ns)] * 3, 1, 0.0, seed, att)
        same = bool(len(gc) == k and gdrawn == odrawn and np.array_equal(gc, oc) and np.array_equal(gu, ou))
        log(f"oracle: {k} rods in {cel:.2f}s; GPU result on the sample equal: {same}")
        out["cpu_baseline"] = {"value": k / cel, "unit": "rods/s", "cores": 1, "kind": "port",
                               "sample": f"serial oracle of the same populate at {ns} rods, same density and seed "
                                         f"({cel:.1f} s)", "sample_gpu_bitwise_equal": same}
    ctx.close()
    return out


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

    if a.workload == "populate":
        out = run_populate(a, L)
    elif a.workload in ("c2", "c5"):
        out = run_c2(a, L, scenes)
    elif a.workload == "c4":
        out = run_c4(a, L, scenes)
    else:
        out = run_vv(a, L, scenes, a.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
