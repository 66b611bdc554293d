"""GPU non-overlapping populate (rdk_populate_nonoverlap) against the serial
oracle (orc_populate_nonoverlap, the reference's loop of
src/app/main.cpp:1857-2140): the same rods, bit for bit, in the same order,
with the same number of candidates drawn, whatever the parallel window."""
import json

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

CASES = [
    # n, length, diameter, lo, hi, periodic, cell_size, seed, max_attempts
    (1500, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 1, 0.6, 11, 1000),
    (1500, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 0, 0.6, 11, 1000),
    (3000, 1.0, 0.05, [0, 0, 0], [3, 2, 4], 1, 0.0, 7, 1000),      # cell size max(L/4, 1.25 d)
    (5000, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 0, 0.6, 5, 50),     # saturates: stops after 50 failures
    (5000, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 1, 0.3, 9, 30),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("window", [0, 64, 1000])
def test_populate_equals_serial_oracle(gpu_ctx_factory, case, window):
    n, L, d, lo, hi, per, cs, seed, att = CASES[case]
    print(f"populate case {case} window {window}", flush=True)
    k, c, u, drawn = O.populate_nonoverlap(n, L, d, lo, hi, per, cs, seed, att)
    ctx = gpu_ctx_factory()
    gc, gu, gq, gdrawn = ctx.populate_nonoverlap(n, L, d, lo, hi, per, cs, seed, att, window=window)
    assert len(gc) == k and gdrawn == drawn
    assert np.array_equal(gc, c) and np.array_equal(gu, u)
    # quaternions rotate +Y onto the direction
    q = gq.astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    w, x, y, z = q.T
    axis = np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], 1)
    assert np.allclose(axis, u, atol=2e-6)


def test_empty_and_invalid(gpu_ctx_factory):
    from rod_dynamics_py import RdkError

    ctx = gpu_ctx_factory()
    c, u, q, drawn = ctx.populate_nonoverlap(0, 0.5, 0.05, [0, 0, 0], [1, 1, 1], 1, 0.6, 1, 1000)
    assert len(c) == 0 and drawn == 0
    with pytest.raises(RdkError):
        ctx.populate_nonoverlap(10, 0.5, 0.05, [0, 0, 0], [0, 1, 1], 1, 0.6, 1, 1000)
    with pytest.raises(RdkError):
        ctx.populate_nonoverlap(10, -0.5, 0.05, [0, 0, 0], [1, 1, 1], 1, 0.6, 1, 1000)


def test_simulator_populate_nonoverlap_is_the_reference_packing(tmp_path):
    """A scene with populate.mode "nonoverlap" places the rods the reference
    places for the same seed (positions and axes in its frame)."""
    from rod_dynamics_py.simulator import Simulator

    scene = {"physics": {"dt": 1e-3, "gravity": [0, 0, 0], "nsc": {"enabled": True}},
             "scene": {"periodic": {"enabled": True, "min": [-1, -1, -1], "max": [1, 1, 1], "cellSize": 0.6},
                       "populate": {"count": 800, "mode": "nonoverlap", "seed": 17, "maxAttempts": 10},
                       "bodies": [{"length": 0.5, "diameter": 0.05}]}}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"))
    k, c, u, _ = O.populate_nonoverlap(800, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 1, 0.6, 17, 1000)
    rods = sim.rods()
    assert len(rods) == k
    pos = np.array([r["position"] for r in rods])
    ax = np.array([np.subtract(r["endpoint_b"], r["endpoint_a"]) for r in rods]) / 0.5
    assert np.allclose(pos, c.astype(np.float64), atol=1e-6)
    assert np.allclose(ax, u.astype(np.float64), atol=2e-6)
    sim.close()
