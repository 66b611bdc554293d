"""Learned PSOR warm start: training data from the solver itself, GPU training
and the A/B measurement against the plain starts.

The network (12 -> 64 -> 64 -> 3, SiLU; forward and backward kernels in
rdk_mlp.hip) predicts the converged impulses, in units of
lambda_s = |g| dt / g_n, of the manifolds that have no previous-step normal
impulse (new contacts and contacts that carried no load); the others keep the
previous-step start.  Training pairs are (features, lambda* / lambda_s) of such
manifolds, recorded from hard-contact settling runs; lambda* is what PSOR
converges to at the residual tolerance.
The reference has no learned start (its warm start is the previous-step cache,
src/physics/nsc_contact.cpp:305-317).

Command line (one GPU):
    python -m rod_dynamics_py.warmstart train --out weights.npy [--n 100000]
    python -m rod_dynamics_py.warmstart eval --weights weights.npy
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import scenes

TRAIN_SEED0 = 0x5EED1000
EVAL_SEED0 = 0x5EED2000
TARGET_CLIP = 1e4  # |target| in units of lambda_s; drops the rare pathological impacts


def _wo_warm(p, warm_mode):
    q = L.NscParams()
    for name, _ in L.NscParams._fields_:
        setattr(q, name, getattr(p, name))
    q.warm_mode = warm_mode
    return q


def _det0(det):
    return L.detect_params(0.0, det.chunk_len, det.cell_scale)


def record_step(ctx, det, p, select="uncached"):
    """One hard-contact step in stages, as rdk_step_nsc runs it, returning the
    training pairs of its manifolds (or None when the solve did not converge):
    those without a previous-step normal impulse (select="uncached", the ones
    the learned start is applied to) or all of them (select="all"; the
    previous-step impulse is one of the features)."""
    ctx.integrate(L.RDK_NSC_FREE)
    ctx.detect(_det0(det))
    p1 = _wo_warm(p, 1)
    nm = ctx.nsc_build(p1)
    out = None
    if nm > 0:
        X, ls = ctx.mlp_features(nm)
        _, _, lam0 = ctx.nsc_manifolds(nm)
        st = ctx.nsc_solve(p1, allow_noconv=True)
        _, _, lam = ctx.nsc_manifolds(nm)
        if st["residual"] <= p.tol:
            sel = lam0[:, 0] <= 0.0 if select == "uncached" else np.ones(len(lam0), bool)
            T = lam[sel] / ls[sel, None]
            out = (X[sel], T.astype(np.float32))
    ctx.integrate(L.RDK_NSC_POS)
    return out


def collect(ctx, sc, steps, record_from, every, max_pairs, log=None, select="uncached"):
    """Run the settling protocol of scene `sc` for `steps` steps, recording the
    pairs of every `every`-th step from `record_from` on."""
    p = _wo_warm(sc.params, 1)
    Xs, Ts = [], []
    have = 0
    s = 0
    while s < steps and have < max_pairs:
        if s < record_from:
            k = record_from - s
            ctx.step_nsc(_det0(sc.detect), p, k)
            s += k
            continue
        r = record_step(ctx, sc.detect, p, select)
        s += 1
        if r is not None:
            Xs.append(r[0])
            Ts.append(r[1])
            have += len(r[0])
        k = min(every - 1, steps - s)
        if k > 0:
            ctx.step_nsc(_det0(sc.detect), p, k)
            s += k
        if log and len(Xs) % 20 == 0:
            print(f"[warmstart] step {s}: {have} pairs", file=log, flush=True)
    if not Xs:
        return np.zeros((0, L.MLP_IN), np.float32), np.zeros((0, L.MLP_OUT), np.float32)
    return np.concatenate(Xs), np.concatenate(Ts)


def normalisation(X, T):
    """Per-feature / per-output mean and scale (scale >= 1e-3)."""
    mx, sx = X.mean(0), np.maximum(X.std(0), 1e-3)
    mt, st = T.mean(0), np.maximum(T.std(0), 1e-3)
    return mx, sx, mt, st


def fold(params, mx, sx, mt, st):
    """Fold input standardisation and output scaling into the first and last
    layers: the network trained on (X - mx)/sx -> (T - mt)/st becomes one that
    maps raw X to raw T, so inference needs no extra pass."""
    p = np.asarray(params, np.float64).copy()
    H, IN, OUT = L.MLP_H, L.MLP_IN, L.MLP_OUT
    o_w1, o_b1 = 0, IN * H
    o_w2 = o_b1 + H
    o_b2 = o_w2 + H * H
    o_w3 = o_b2 + H
    o_b3 = o_w3 + H * OUT
    W1 = p[o_w1:o_b1].reshape(IN, H)
    b1 = p[o_b1:o_w2]
    W1n = W1 / sx[:, None]
    b1n = b1 - (mx / sx) @ W1
    W3 = p[o_w3:o_b3].reshape(H, OUT)
    b3 = p[o_b3:]
    W3n = W3 * st[None, :]
    b3n = b3 * st + mt
    p[o_w1:o_b1] = W1n.ravel()
    p[o_b1:o_w2] = b1n
    p[o_w3:o_b3] = W3n.ravel()
    p[o_b3:] = b3n
    return p.astype(np.float32)


def train(ctx, X, T, epochs=20, batch=2048, lr=2e-3, weight_decay=1e-4, seed=0, val_frac=0.05):
    """Train on (X, T) with the GPU kernels; returns folded parameters and a
    report with train/validation losses in raw target units."""
    keep = np.all(np.abs(T) <= TARGET_CLIP, axis=1) & np.all(np.isfinite(X), axis=1)
    X, T = X[keep], T[keep]
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    nval = max(1, int(val_frac * len(X)))
    va, tr = idx[:nval], idx[nval:]
    mx, sx, mt, st = normalisation(X[tr], T[tr])
    Xn = ((X - mx) / sx).astype(np.float32)
    Tn = ((T - mt) / st).astype(np.float32)
    ctx.mlp_init(seed)
    t0 = time.perf_counter()
    loss = ctx.mlp_train(Xn[tr], Tn[tr], epochs=epochs, batch=batch, lr=lr, weight_decay=weight_decay, seed=seed)
    t_train = time.perf_counter() - t0
    p = fold(ctx.mlp_get_params(), mx, sx, mt, st)
    ctx.mlp_set_params(p)
    yv = ctx.mlp_infer(X[va])
    mse_val = float(np.mean((yv - T[va]) ** 2))
    mse_zero = float(np.mean(T[va] ** 2))  # the plain start of these manifolds is zero
    return p, {"pairs_used": int(len(tr)), "pairs_dropped": int((~keep).sum()), "epochs": epochs, "batch": batch,
               "lr": lr, "train_loss_normalised": loss, "val_mse": mse_val, "val_mse_zero_start": mse_zero,
               "train_seconds": t_train}


def ab_step(ctx, det, p):
    """One step of the protocol in which the same manifolds are solved from the
    cold, learned (new contacts only: warm_mode 2; every manifold: warm_mode 3;
    every manifold from the contact alone: warm_mode 4) and previous-step
    starts (same pre-solve velocities); the trajectory continues from the
    previous-step solve, as the plain protocol."""
    ctx.integrate(L.RDK_NSC_FREE)
    ctx.detect(_det0(det))
    p1 = _wo_warm(p, 1)
    nm = ctx.nsc_build(p1)
    res = None
    if nm > 0:
        _, _, lam_prev = ctx.nsc_manifolds(nm)
        ctx.nsc_build(_wo_warm(p, 2))
        _, _, lam_learn = ctx.nsc_manifolds(nm)
        ctx.nsc_build(_wo_warm(p, 3))
        _, _, lam_all = ctx.nsc_manifolds(nm)
        ctx.nsc_build(_wo_warm(p, 4))
        _, _, lam_fresh = ctx.nsc_manifolds(nm)
        s0 = ctx.download_state()
        res = {"manifolds": nm}
        for name, lam0 in (("cold", np.zeros_like(lam_prev)), ("learned", lam_learn), ("learned_all", lam_all),
                           ("learned_fresh", lam_fresh), ("previous", lam_prev)):
            ctx.upload_state(v=s0["v"], w=s0["w"])
            ctx.nsc_set_lambda(lam0)
            w0 = ctx.work_counters()["sweep_contacts"]
            ctx.sync()
            t0 = time.perf_counter()
            st = ctx.nsc_solve(p1, allow_noconv=True)
            ctx.sync()
            res[name] = {"sweeps": st["sweeps"], "residual": st["residual"],
                         "manifold_updates": ctx.work_counters()["sweep_contacts"] - w0,
                         "ms": 1e3 * (time.perf_counter() - t0)}
    ctx.integrate(L.RDK_NSC_POS)
    return res


def ab_summary(rows):
    rows = [r for r in rows if r is not None]
    out = {"steps": len(rows), "manifolds_per_step": float(np.mean([r["manifolds"] for r in rows])) if rows else 0.0}
    for name in ("cold", "previous", "learned", "learned_all", "learned_fresh"):
        if not rows:
            continue
        out[name] = {k: float(np.mean([r[name][k] for r in rows])) for k in ("sweeps", "manifold_updates", "ms")}
        out[name]["residual_max"] = float(max(r[name]["residual"] for r in rows))
    if rows:
        for base in ("cold", "previous"):
            out[f"sweeps_ratio_{base}_over_learned"] = out[base]["sweeps"] / max(out["learned"]["sweeps"], 1e-9)
            out[f"updates_ratio_{base}_over_learned"] = (out[base]["manifold_updates"] /
                                                        max(out["learned"]["manifold_updates"], 1e-9))
            out[f"sweeps_ratio_{base}_over_learned_all"] = (out[base]["sweeps"] /
                                                           max(out["learned_all"]["sweeps"], 1e-9))
            out[f"sweeps_ratio_{base}_over_learned_fresh"] = (out[base]["sweeps"] /
                                                             max(out["learned_fresh"]["sweeps"], 1e-9))
            out[f"updates_ratio_{base}_over_learned_fresh"] = (out[base]["manifold_updates"] /
                                                              max(out["learned_fresh"]["manifold_updates"], 1e-9))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rod_dynamics_py.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--n", type=int, default=100_000)
    t.add_argument("--seeds", type=int, default=2)
    t.add_argument("--steps", type=int, default=3600)
    t.add_argument("--record-from", type=int, default=1500)
    t.add_argument("--every", type=int, default=5)
    t.add_argument("--pairs", type=int, default=1_000_000)
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--save-data", default="")
    t.add_argument("--select", choices=["uncached", "all"], default="uncached",
                   help="training pairs: manifolds without a previous-step impulse, or all manifolds")
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--n", type=int, default=100_000)
    e.add_argument("--settle", type=int, default=3500)
    e.add_argument("--steps", type=int, default=20)
    a = ap.parse_args(argv)
    ctx = L.Context(0)
    if a.cmd == "train":
        Xs, Ts = [], []
        per_seed = a.pairs // a.seeds + 1
        for k in range(a.seeds):
            sc = scenes.c2_settling(n=a.n, seed=TRAIN_SEED0 + k)
            sc.load(ctx)
            X, T = collect(ctx, sc, a.steps, a.record_from, a.every, per_seed, log=sys.stderr, select=a.select)
            print(f"[warmstart] seed {k}: {len(X)} pairs", file=sys.stderr, flush=True)
            Xs.append(X)
            Ts.append(T)
        X, T = np.concatenate(Xs), np.concatenate(Ts)
        if a.save_data:
            np.savez_compressed(a.save_data, X=X, T=T)
        p, rep = train(ctx, X, T, epochs=a.epochs)
        np.save(a.out, p)
        rep["pairs_collected"] = int(len(X))
        rep["select"] = a.select
        rep["rods"] = a.n
        print(json.dumps(rep), flush=True)
    else:
        p = np.load(a.weights)
        ctx.mlp_set_params(p)
        sc = scenes.c2_settling(n=a.n, seed=EVAL_SEED0)
        sc.load(ctx)
        ctx.step_nsc(_det0(sc.detect), _wo_warm(sc.params, 1), a.settle)
        rows = [ab_step(ctx, sc.detect, sc.params) for _ in range(a.steps)]
        print(json.dumps(ab_summary(rows)), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
