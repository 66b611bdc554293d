"""GPU entanglement observer (rdk_entanglement) against the oracle's pair sum
(orc_entanglement): the pair counts are equal, the sums agree to FP64
summation-order rounding, on free and periodic systems, with and without a
cutoff, one system and several."""
import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu


def random_rods(n, side, seed, hmax=2.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q, np.full(n, 0.05), rng.uniform(0.2, hmax, n)


def oracle_sum(x, q, r, h, lo, hi, periodic, cutoff):
    W = O.World(x, q, None, None, r, h, np.ones(len(r)))
    W.set_box(lo, hi, periodic)
    return W.entanglement(cutoff)


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 1, 1), (1, 0, 1)])
@pytest.mark.parametrize("cutoff", [-1.0, 0.8])
def test_entanglement_matches_oracle(gpu_ctx_factory, periodic, cutoff):
    n, side = 700, 8.0
    x, q, r, h = random_rods(n, side, 4)
    print(f"entanglement n={n} periodic={periodic} cutoff={cutoff}", flush=True)
    ctx = gpu_ctx_factory()
    ctx.set_bodies(x, q, np.zeros((n, 3)), np.zeros((n, 3)), r, h, np.ones(n), np.zeros(n))
    lo, hi = np.zeros(3), np.full(3, side)
    ctx.set_systems([0, n], [L.system_desc(lo, hi, periodic, 0, 0, 0.5)])
    ctx.set_physics(L.physics())
    sums, pairs = ctx.entanglement(cutoff)
    s, m = oracle_sum(x, q, r, h, lo, hi, periodic, cutoff)
    assert int(pairs[0]) == m
    if cutoff < 0:
        assert m == n * (n - 1) // 2
    assert abs(sums[0] - s) <= 1e-12 * s, (sums[0], s)
    again, _ = ctx.entanglement(cutoff)
    assert again[0] == sums[0]  # fixed reduction order


def test_entanglement_per_system(gpu_ctx_factory):
    """Several systems of different sizes (one with a single rod, one spanning
    several 64-rod tiles with a ragged last tile) in one context."""
    sizes = [1, 130, 64, 257, 2]
    periodic = [(1, 1, 1), (0, 0, 0), (1, 1, 0), (1, 1, 1), (0, 0, 0)]
    xs, qs, rs, hs, descs, boxes = [], [], [], [], [], []
    for k, (ns, per) in enumerate(zip(sizes, periodic)):
        side = 3.0 + k
        x, q, r, h = random_rods(ns, side, 20 + k)
        off = np.array([100.0 * k, 0, 0])
        xs.append(x + off), qs.append(q), rs.append(r), hs.append(h)
        lo, hi = off, off + side
        descs.append(L.system_desc(lo, hi, per, 0, 0, 0.5))
        boxes.append((x + off, q, r, h, lo, hi, per))
    offs = np.concatenate([[0], np.cumsum(sizes)])
    n = int(offs[-1])
    ctx = gpu_ctx_factory()
    ctx.set_bodies(np.vstack(xs), np.vstack(qs), np.zeros((n, 3)), np.zeros((n, 3)), np.concatenate(rs),
                   np.concatenate(hs), np.ones(n), np.zeros(n))
    ctx.set_systems(offs, descs)
    ctx.set_physics(L.physics())
    for cutoff in (-1.0, 1.0):
        print(f"entanglement per system cutoff={cutoff}", flush=True)
        sums, pairs = ctx.entanglement(cutoff)
        for k, (x, q, r, h, lo, hi, per) in enumerate(boxes):
            s, m = oracle_sum(x, q, r, h, lo, hi, per, cutoff)
            assert int(pairs[k]) == m, k
            assert abs(sums[k] - s) <= 1e-12 * max(s, 1e-300), k


def test_reference_packing_ebar0_on_gpu(gpu_ctx_factory):
    """The reference's frozen packings: the GPU mean |Lk| equals its published
    ebar0 (6 digits) and pair count."""
    from packings import free_box, rods_from_endpoints

    import os

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_packings.npz"))
    for k in range(int(z["count"])):
        x, q, r, h = rods_from_endpoints(z[f"ends_{k}"], float(z[f"radius_{k}"]))
        n = len(r)
        print(f"packing {k} n={n}", flush=True)
        lo, hi = free_box(x)
        ctx = gpu_ctx_factory()
        ctx.set_bodies(x, q, np.zeros((n, 3)), np.zeros((n, 3)), r, h, np.ones(n), np.zeros(n))
        ctx.set_systems([0, n], [L.system_desc(lo, hi, (0, 0, 0), 0, 0, 0.5)])
        ctx.set_physics(L.physics())
        sums, pairs = ctx.entanglement(-1.0)
        assert int(pairs[0]) == int(z[f"ent_pairs_{k}"])
        assert abs(sums[0] / pairs[0] - float(z[f"ebar0_{k}"])) < 1.5e-6
