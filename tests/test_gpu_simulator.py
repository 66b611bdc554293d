"""The reference's Python session API on the GPU: a scene JSON with explicit
bodies (the elastic exchange of smoke_test.py:165-195 through Simulator), and
a frozen relaxed packing of the reference loaded from its endpoint file."""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py.simulator import Simulator

pytestmark = pytest.mark.gpu


def test_elastic_exchange_through_the_session_api(tmp_path):
    # crossed rods (one along y, one along z) touching at their centres: the
    # impulse passes through both centres of mass, so an elastic frictionless
    # contact swaps the velocities of equal masses
    scene = {
        "physics": {"dt": 1e-3, "gravity": [0, 0, 0], "lin_damp": 0.0, "ang_damp": 0.0,
                    "nsc": {"enabled": True, "mu": 0.0, "beta": 0.0, "cfm": 0.0, "omega": 1.0, "velocity_iters": 200,
                            "position_stabilization": False}},
        "scene": {"bodies": [
            {"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 1.0, "pos": [-0.249, 0.0, 0.0],
             "v_lin": [1.0, 0.0, 0.0]},
            {"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 1.0, "pos": [0.249, 0.0, 0.0],
             "rot_quat": [0.7071067811865476, 0.7071067811865476, 0.0, 0.0], "v_lin": [-1.0, 0.0, 0.0]}]}}
    path = tmp_path / "elastic.json"
    path.write_text(json.dumps(scene))
    sim = Simulator(str(path))
    d0 = sim.diagnostics()
    assert d0["frame_index"] == 0 and d0["rod_count"] == 2
    ke0 = d0["last_ke"]
    sim.step()
    d = sim.diagnostics()
    assert d["frame_index"] == 1 and d["last_hit_count"] == 1 and d["dt"] == 1e-3
    r = sim.rods()
    assert np.allclose(r[0]["linear_velocity"], [-1, 0, 0], atol=1e-9)
    assert np.allclose(r[1]["linear_velocity"], [1, 0, 0], atol=1e-9)
    assert d["last_ke"] == pytest.approx(ke0, rel=1e-9)
    m = 1000.0 * np.pi * 0.25 ** 2 * 1.0  # density * pi r^2 * L (cylinder part), rigid_body.cpp:19-24
    assert r[0]["mass"] == pytest.approx(m) and r[0]["inverse_mass"] == pytest.approx(1 / m)
    assert r[0]["shape"] == "capsule" and r[0]["radius"] == 0.25 and r[0]["half_height"] == 0.5
    # the rod axes in the reference frame: +Y (identity) and +Z (90 degrees about x)
    assert np.allclose(np.subtract(r[0]["endpoint_b"], r[0]["endpoint_a"]), [0, 1.0, 0], atol=1e-12)
    assert np.allclose(np.subtract(r[1]["endpoint_b"], r[1]["endpoint_a"]), [0, 0, 1.0], atol=1e-12)
    assert np.allclose(r[0]["angular_velocity"], 0, atol=1e-12) and np.allclose(r[1]["angular_velocity"], 0, atol=1e-12)
    sim.close()


def test_reference_packing_from_its_endpoint_file(tmp_path):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_packings.npz"))
    ends, radius = z["ends_5"], float(z["radius_5"])  # a 200-rod packing, L/d = 10
    lines = [f"# rod_radius={radius}"] + [" ".join(f"{v:.17g}" for v in e) for e in ends]
    (tmp_path / "x_relaxed.txt").write_text("\n".join(lines) + "\n")
    scene = {"physics": {"dt": 1e-4, "gravity": [0, 0, 0], "nsc": {"enabled": True, "position_stabilization": False}}}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"), str(tmp_path / "x_relaxed.txt"))
    r = sim.rods()
    assert len(r) == len(ends)
    a = np.array([x["endpoint_a"] for x in r])
    b = np.array([x["endpoint_b"] for x in r])
    assert np.allclose(a, ends[:, :3], atol=1e-12) and np.allclose(b, ends[:, 3:], atol=1e-12)
    assert np.allclose([x["radius"] for x in r], radius)
    st = sim.ctx.download_state()  # the library frame: the oracle sees the same FP64 rods
    sim.step()
    n = len(ends)
    W = O.World(st["x"], st["q"], None, None, sim.r, sim.h, np.ones(n))
    W.set_box(st["x"].min(0) - 10, st["x"].max(0) + 10, [0, 0, 0])
    assert sim.diagnostics()["last_hit_count"] == W.detect(0, 0.0)  # every overlapping pair is a manifold
    sim.close()


def test_state_files_round_trip_through_the_simulator(tmp_path):
    """A populated scene runs a few steps; its per-rod log restarts a second
    simulator at the logged frame (init-state CSV) with the FP32-rounded
    state; the NDJSON snapshot and the endpoint file carry the same rods."""
    from rod_dynamics_py import sceneio

    scene = {"physics": {"dt": 1e-3, "gravity": [0, -10, 0], "nsc": {"enabled": True}},
             "scene": {"periodic": {"enabled": True, "min": [-2, -2, -2], "max": [2, 2, 2]},
                       "populate": {"count": 60, "mode": "nonoverlap", "seed": 3, "radius": 0.025,
                                    "density": 1000.0},
                       "bodies": [{"length": 0.5, "diameter": 0.05, "density": 1000.0, "restitution": 0.5,
                                   "friction": 0.5}]}}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"))
    sim.step(5)
    log = tmp_path / "per_rod.csv"
    with open(log, "w") as f:
        sim.write_per_rod(f, header=True)
    with open(tmp_path / "snap.ndjson", "w") as f:
        sim.write_snapshot(f)
    sim.save_endpoints(str(tmp_path / "ends.txt"))
    ref = sim.rods()
    snap = json.loads((tmp_path / "snap.ndjson").read_text())
    assert snap["frame"] == 5 and len(snap["bodies"]) == 60
    sim2 = Simulator(str(tmp_path / "s.json"), init_state_path=str(log))
    assert sim2.diagnostics()["frame_index"] == 5 and sim2.diagnostics()["rod_count"] == 60
    r2 = sim2.rods()
    for a, b, s in zip(ref, r2, snap["bodies"]):
        for key, skey in (("position", "pos"), ("linear_velocity", "vel"), ("angular_velocity", "omega"),
                          ("orientation_wxyz", "quat")):
            assert np.allclose(a[key], b[key], rtol=1e-5, atol=1e-6), key
            assert np.allclose(a[key], s[skey], rtol=1e-5, atol=1e-6), skey
        assert b["radius"] == 0.025 and b["half_height"] == 0.25
    E = np.loadtxt(tmp_path / "ends.txt", delimiter=",", comments="#", skiprows=5)
    got = np.array([r["endpoint_a"] + r["endpoint_b"] for r in ref])
    assert np.allclose(E, got, atol=2e-6)
    sim2.step(2)
    assert sim2.diagnostics()["frame_index"] == 7
    sim.close()
    sim2.close()


def test_network_log_of_a_hard_contact_step(tmp_path):
    """Two crossed rods meet head-on with restitution 0 and position
    stabilisation on (the projection reuses the manifold buffers): the network
    log holds the step's rod-rod manifold with a positive normal impulse, the
    approach velocity before the solve and zero normal velocity after it, and
    matches the kept velocity-solve manifolds."""
    import io

    from rod_dynamics_py import sceneio

    scene = {
        "physics": {"dt": 1e-3, "gravity": [0, 0, 0], "lin_damp": 0.0, "ang_damp": 0.0,
                    "nsc": {"enabled": True, "mu": 0.0, "beta": 0.0, "cfm": 0.0, "omega": 1.0, "velocity_iters": 200,
                            "position_stabilization": True}},
        "scene": {"bodies": [
            {"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 0.0, "pos": [-0.249, 0.0, 0.0],
             "v_lin": [1.0, 0.0, 0.0]},
            {"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 0.0, "pos": [0.249, 0.0, 0.0],
             "rot_quat": [0.7071067811865476, 0.7071067811865476, 0.0, 0.0], "v_lin": [-1.0, 0.0, 0.0]}]}}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"))
    sim.step()
    buf = io.StringIO()
    sim.write_network(buf, header=True)
    lines = buf.getvalue().splitlines()
    assert lines[0] + "\n" == sceneio.NETWORK_HEADER
    assert len(lines) == 2
    row = lines[1].split(",")
    assert row[:3] == ["1", "0", "1"]
    vals = [float(t) for t in row[3:]]
    lam_n, vn_pre, vn_post = vals[7], vals[10], vals[11]
    assert lam_n > 0 and vn_pre == pytest.approx(-2.0, rel=1e-6) and abs(vn_post) < 1e-9
    assert abs(vals[3]) == pytest.approx(1.0, rel=1e-6)  # normal along x in the reference frame
    rec, ab, lam = sim.ctx.kept_manifolds()
    rr = ab[:, 1] >= 0
    assert lam[rr][0, 0] == pytest.approx(lam_n, rel=1e-5)  # printed with 6 significant digits
    sim.close()


def test_network_log_of_a_soft_contact_step(tmp_path):
    """Two overlapping rods under soft contact: the network log has the soft
    header and one row per contact of the last force evaluation, with
    force_a = -force_b along the normal and friction_b = -friction_a; a
    Hertz-Mindlin scene writes the header only."""
    import io

    from rod_dynamics_py import sceneio

    bodies = [{"length": 1.0, "diameter": 0.5, "density": 1000.0, "pos": [-0.24, 0.0, 0.0], "v_lin": [0.0, 0.0, 0.3]},
              {"length": 1.0, "diameter": 0.5, "density": 1000.0, "pos": [0.24, 0.0, 0.0],
               "rot_quat": [0.7071067811865476, 0.7071067811865476, 0.0, 0.0], "v_lin": [0.0, 0.0, -0.3]}]
    for model in ("soft_contact", "hertz_mindlin"):
        scene = {"physics": {"dt": 1e-4, "gravity": [0, 0, 0], "lin_damp": 0.0, "ang_damp": 0.0,
                             model: {"enabled": True}},
                 "scene": {"bodies": bodies}}
        (tmp_path / "s.json").write_text(json.dumps(scene))
        sim = Simulator(str(tmp_path / "s.json"))
        sim.step()
        buf = io.StringIO()
        sim.write_network(buf, header=True)
        lines = buf.getvalue().splitlines()
        assert lines[0] + "\n" == sceneio.NETWORK_SOFT_HEADER
        if model == "hertz_mindlin":
            assert len(lines) == 1
            sim.close()
            continue
        raw = sim.ctx.contacts_raw()
        assert len(raw) >= 1 and len(lines) == 1 + len(raw)
        for ln in lines[1:]:
            row = ln.split(",")
            assert row[0] == "1" and {row[1], row[2]} == {"0", "1"}
            v = np.array([float(t) for t in row[3:]])
            nrm, fa, fb, fra, frb = v[3:6], v[7:10], v[10:13], v[13:16], v[16:19]
            assert np.array_equal(fa, -fb) and np.array_equal(fra, -frb)
            assert np.linalg.norm(fb) > 0 and abs(abs(np.dot(fb, nrm)) / np.linalg.norm(fb) - 1) < 1e-5
        sim.close()


def test_profile_log_of_a_hard_contact_step(tmp_path):
    """The profile row of the head-on hard-contact step: one manifold, one
    impact, one island, the kinetic energy 2 x m v^2 / 2 after the free update
    (no gravity or damping), less after the inelastic solve, the normal impulse
    sum equal to the kept manifold's, and the slim header for soft contact."""
    import io

    from rod_dynamics_py import sceneio

    bodies = [{"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 0.0, "pos": [-0.249, 0.0, 0.0],
               "v_lin": [1.0, 0.0, 0.0]},
              {"length": 1.0, "diameter": 0.5, "density": 1000.0, "restitution": 0.0, "pos": [0.249, 0.0, 0.0],
               "rot_quat": [0.7071067811865476, 0.7071067811865476, 0.0, 0.0], "v_lin": [-1.0, 0.0, 0.0]}]
    scene = {"physics": {"dt": 1e-3, "gravity": [0, 0, 0], "lin_damp": 0.0, "ang_damp": 0.0,
                         "nsc": {"enabled": True, "mu": 0.0, "beta": 0.0, "cfm": 0.0, "omega": 1.0,
                                 "velocity_iters": 200}},
             "scene": {"bodies": bodies}}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"))
    sim.step()
    buf = io.StringIO()
    sim.write_profile(buf, header=True)
    head, row = buf.getvalue().splitlines()
    assert head + "\n" == sceneio.PROFILE_HEADER_NSC
    cols = dict(zip(head.split(","), row.split(",")))
    assert len(row.split(",")) == 35
    assert cols["frame"] == "1" and cols["rods"] == "2"
    assert cols["contacts"] == "1" and cols["collisions"] == "1" and cols["islands"] == "1"
    m = 1000.0 * math.pi * 0.25 ** 2 * 1.0
    ke_int, ke_sol = float(cols["KE_after_integrate"]), float(cols["KE_after_solve"])
    assert ke_int == pytest.approx(m, rel=1e-5) and ke_sol < 0.5 * ke_int
    assert float(cols["KE"]) == pytest.approx(float(cols["KE_after_posCorrect"]), rel=1e-5)
    rec, ab, lam = sim.ctx.kept_manifolds()
    assert float(cols["jn_sum"]) == pytest.approx(np.abs(lam[:, 0]).sum(), rel=1e-5)
    assert float(cols["step_ms"]) > 0.0
    sim.close()
    scene["physics"] = {"dt": 1e-4, "gravity": [0, 0, 0], "lin_damp": 0.0, "ang_damp": 0.0,
                        "soft_contact": {"enabled": True}}
    bodies[0]["pos"], bodies[1]["pos"] = [-0.24, 0.0, 0.0], [0.24, 0.0, 0.0]
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sim = Simulator(str(tmp_path / "s.json"))
    sim.step()
    buf = io.StringIO()
    sim.write_profile(buf, header=True)
    head, row = buf.getvalue().splitlines()
    assert head + "\n" == sceneio.PROFILE_HEADER_SOFT
    cols = dict(zip(head.split(","), row.split(",")))
    assert len(row.split(",")) == 21 and cols["collisions"] == "-1" and float(cols["soft_PE"]) > 0.0
    sim.close()


def test_reference_paper_scene_runs_unchanged():
    """The reference's Table S1 scene (assets/scenes/default_entangled_nsc.json,
    committed as tests/golden/default_entangled_nsc.json) through the session
    API unchanged: 2000 rods of populate "nonoverlap" (seed 12345) equal to
    the oracle's serial FP32 placement, velocities of randomInit "gaussian"
    (seed 42, parallel spin projected) equal to the oracle's g++ restatement
    of main.cpp:2213-2294 bit for bit, then 30 hard-contact steps with the
    scene's solver settings (mu 0, cfm 0.05, 200 PGS iterations, omega 1)
    against the oracle world started from the same FP64 state."""
    path = os.path.join(os.path.dirname(__file__), "golden", "default_entangled_nsc.json")
    scene = json.load(open(path))
    sim = Simulator(path)
    n = scene["scene"]["populate"]["count"]
    assert sim.diagnostics()["rod_count"] == n
    pb = scene["scene"]["periodic"]
    body = scene["scene"]["bodies"][0]
    k, c, u, _ = O.populate_nonoverlap(n, body["length"], body["diameter"], pb["min"], pb["max"], 0, pb["cellSize"],
                                       scene["scene"]["populate"]["seed"], scene["scene"]["populate"]["maxAttempts"])
    assert k == n
    r = sim.rods()
    x = np.array([q["position"] for q in r])
    assert np.array_equal(x, c.astype(np.float64))
    ax = np.array([np.subtract(q["endpoint_b"], q["endpoint_a"]) for q in r])
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    assert np.abs(ax - u.astype(np.float64)).max() < 1e-6
    ri = scene["scene"]["randomInit"]
    vo, wo = O.random_init(sim.q32, 0, v_sigma=ri["vSigma"], w_sigma=ri["wSigma"], seed=ri["seed"],
                           project=ri["projectParallelSpin"])
    v = np.array([q["linear_velocity"] for q in r])
    w = np.array([q["angular_velocity"] for q in r])
    assert np.array_equal(v, vo.astype(np.float64)) and np.array_equal(w, wo.astype(np.float64))
    assert np.any(v != 0.0) and np.abs(np.sum(w * ax, axis=1)).max() < 1e-6
    st0 = sim.ctx.download_state()  # library frame
    p = sim.params
    m = 4.0 * float(np.max(sim.h + sim.r))
    W = O.World(st0["x"], st0["q"], st0["v"], st0["w"], sim.r, sim.h, sim.mass, sim.rest)
    W.set_box(st0["x"].min(0) - m, st0["x"].max(0) + m, [0, 0, 0])
    W.set_physics((0.0, 0.0, 0.0), sim.dt)
    op = O.nsc_params(mu=0.0, beta=p.beta, cfm=p.cfm, omega=p.omega, tol=0.0, max_sweeps=p.max_sweeps, check_every=1,
                      friction_proj=0, stop_rule=0, position_stabilization=p.position_stabilization,
                      warm_start=p.warm_mode, position_iters=p.position_iters,
                      position_psor_iters=p.position_psor_iters, slop=p.slop)
    print(f"paper scene: {n} rods, cfm {p.cfm}, {p.max_sweeps} sweeps max", flush=True)
    sim.step(30)
    W.nsc_step(op, 30, method=1)
    sg, so = sim.ctx.download_state(), W.state()
    assert np.abs(sg["x"] - so["x"]).max() < 1e-9
    assert np.abs(sg["v"] - so["v"]).max() < 1e-9
    assert np.abs(sg["w"] - so["w"]).max() < 1e-9
    assert sim.diagnostics()["frame_index"] == 30
    sim.close()
