"""The fixed-capacity hard-contact step (every count on the device, no host
synchronisation inside a step) and its graph-captured replay against the
eager step that reads its counts back to the host (App::physicsStep,
main.cpp:4333-4858; nsc_contact.cpp:239-655 for the solve).

Same start, same steps: positions, orientations, velocities and the step
statistics must be equal bit for bit in mode 1 (fixed launches) and mode 2
(one captured step replayed).  Steps the fixed path cannot finish (more
contacts or islands than reserved, an island with more bodies than the LDS
class holds) are rolled back and redone eagerly; those runs must be equal
too.
"""
import numpy as np
import pytest

from rod_dynamics_py import _lib as L
from rod_dynamics_py import scenes
from test_gpu_nsc import big_clusters, clusters, gpu_load, pile_and_clusters

pytestmark = pytest.mark.gpu

STAT_KEYS = ("steps", "contacts_total", "manifolds_total", "sweeps_total", "colors_total", "nonconverged_steps",
             "residual_max")


def _settling(n=3000):
    sc = scenes.c2_settling(n=n, seed=scenes.SEED0 + 2)
    return sc


def _load(ctx, kind):
    """Loads scene `kind` into ctx; returns (detect params, nsc params)."""
    if kind == "batched":
        sc = scenes.c4_batched(nsys=8, seed=scenes.SEED0 + 44, kT=25.0, threads=8, n_range=(300, 1500),
                               aspects=(4, 8, 16, 32))
        sc.load(ctx)
        return L.detect_params(0.0, 4.0, 1.0), sc.params
    if kind == "settling":
        sc = _settling()
        sc.load(ctx)
        ctx.upload_state(v=np.tile([0.0, 0.0, -8.0], (sc.n, 1)))
        return sc.detect, sc.params
    B = {"clusters": clusters, "big_clusters": big_clusters, "pile": pile_and_clusters}[kind]()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=4000, check_every=4, friction_proj=3,
                     warm_mode=1, stop_rule=1, warm_walls=1, stall_sweeps=1000, hot_set=2)
    return L.detect_params(0.0, 2.0, 1.0), p


def _run(ctx, kind, mode, calls, pair_capacity=0, small_m=None):
    dp, p = _load(ctx, kind)
    if small_m is not None:
        ctx.set_tuning(L.TUNE_SMALL_M, small_m)
    ctx.set_step_graph(mode, pair_capacity)
    tot = {k: 0 for k in STAT_KEYS}
    for k in calls:
        st = ctx.step_nsc(dp, p, k)
        for key in STAT_KEYS:
            tot[key] = max(tot[key], st[key]) if key == "residual_max" else tot[key] + st[key]
    return ctx.download_state(), tot, ctx.step_graph_info()


def _same(a, b):
    for k in ("x", "q", "v", "w"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("kind,calls", [("batched", (12, 18)), ("settling", (25, 15)), ("clusters", (10, 10)),
                                        ("big_clusters", (8, 8))])
def test_fixed_and_captured_steps_equal_eager_bitwise(gpu_ctx_factory, kind, calls):
    print(f"step graph: {kind}, calls {calls}", flush=True)
    s0, t0, _ = _run(gpu_ctx_factory(), kind, 0, calls)
    s1, t1, i1 = _run(gpu_ctx_factory(), kind, 1, calls)
    s2, t2, i2 = _run(gpu_ctx_factory(), kind, 2, calls)
    print(f"  eager {t0}\n  fixed {i1}, graph {i2}", flush=True)
    assert t0["manifolds_total"] > 0
    _same(s0, s1)
    _same(s0, s2)
    assert t1 == t0 and t2 == t0
    # the first step of a context runs eagerly and sets the capacities; a step
    # whose counts outgrow them (contacts, islands per class) is redone eagerly
    for i in (i1, i2):
        assert i["fixed_steps"] + i["fallbacks"] <= sum(calls) - 1
        assert i["fixed_steps"] >= sum(calls) // 2
    assert i2["captures"] >= 1


def test_colouring_rounds_in_the_fixed_step(gpu_ctx_factory):
    """Manifold counts above the single-workgroup limit are coloured by the
    enqueued rounds (limit lowered to 64 on every context)."""
    print("step graph: colouring rounds (small-colouring limit 64)", flush=True)
    s0, t0, _ = _run(gpu_ctx_factory(), "big_clusters", 0, (10,), small_m=64)
    s2, t2, i2 = _run(gpu_ctx_factory(), "big_clusters", 2, (10,), pair_capacity=8192, small_m=64)
    assert t0["manifolds_total"] > 10 * 64
    _same(s0, s2)
    assert t2 == t0 and i2["fixed_steps"] >= 5


def test_capacity_overflow_rolls_back_to_the_eager_step(gpu_ctx_factory):
    """A pair capacity below the contact count: the failed steps are redone
    eagerly from their start, with the same result."""
    print("step graph: pair capacity 256 on dense clusters", flush=True)
    s0, t0, _ = _run(gpu_ctx_factory(), "big_clusters", 0, (12,))
    s2, t2, i2 = _run(gpu_ctx_factory(), "big_clusters", 2, (12,), pair_capacity=256)
    print(f"  {i2}, contacts {t0['contacts_total']}", flush=True)
    assert t0["contacts_total"] > 12 * 256
    assert i2["fallbacks"] > 0
    _same(s0, s2)
    assert t2 == t0


def test_huge_island_rolls_back_to_the_eager_step(gpu_ctx_factory):
    """A log pile of 3600 rods (more bodies than the LDS island class) next to
    small clusters: every step falls back to the eager solve of the huge
    island, with the same result."""
    print("step graph: log pile + clusters", flush=True)
    s0, t0, _ = _run(gpu_ctx_factory(), "pile", 0, (6,))
    s2, t2, i2 = _run(gpu_ctx_factory(), "pile", 2, (6,))
    print(f"  {i2}", flush=True)
    assert i2["fallbacks"] >= 1
    _same(s0, s2)
    assert t2 == t0
