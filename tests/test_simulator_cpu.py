"""The reference's Python session API (src/python/rod_dynamics_py.cpp): scene
JSON parsing with the reference's defaults (src/config/config.cpp), the
endpoint-file loader (App::loadInitialConfigCSV) and the frame rotation, on
the CPU."""
import json
import math
import os

import numpy as np
import pytest

from rod_dynamics_py import simulator as S


def write(tmp_path, name, obj):
    p = tmp_path / name
    p.write_text(json.dumps(obj) if not isinstance(obj, str) else obj)
    return str(p)


def test_defaults_match_reference_config(tmp_path):
    cfg = S.load_scene(write(tmp_path, "s.json", {"physics": {"nsc": {"enabled": True}}}))
    assert cfg["physics"]["dt"] == pytest.approx(1.0 / 600.0)
    assert cfg["physics"]["gravity"] == [0.0, -10.0, 0.0]
    assert (cfg["physics"]["lin_damp"], cfg["physics"]["ang_damp"], cfg["physics"]["w_max"]) == (0.08, 0.12, 80.0)
    n = cfg["nsc"]
    assert (n["mu"], n["beta"], n["cfm"], n["omega"], n["velocity_iters"]) == (0.3, 0.2, 0.0, 1.0, 40)
    assert n["position_stabilization"] and (n["position_iters"], n["position_psor_iters"], n["slop"]) == (5, 50, 1e-4)
    assert cfg["periodic"]["min"] == [-3.0, -1.0, -3.0] and cfg["periodic"]["max"] == [3.0, 3.0, 3.0]
    assert cfg["hm"]["youngs_modulus"] == 7e10 and cfg["soft"]["delta"] == 0.005


def test_overrides_contact_model_and_gravity_rule(tmp_path):
    cfg = S.load_scene(write(tmp_path, "s.json", {
        "physics": {"dt": 1e-3, "contact_model": "hertz-mindlin", "nsc": {"enabled": True, "mu": 0.7}},
        "scene": {"periodic": {"enabled": True}, "randomInit": {"enabled": True, "mode": "gaussian"},
                  "populate": {"count": 10, "grid": True},
                  "bodies": [{"length": 2.0, "radius": 0.1}]}}))
    assert cfg["hm"]["enabled"] and not cfg["nsc"]["enabled"] and not cfg["soft"]["enabled"]
    assert cfg["nsc"]["mu"] == 0.7
    assert cfg["physics"]["gravity"] == [0.0, 0.0, 0.0]  # periodic + randomInit, gravity not given
    assert cfg["populate"]["mode"] == "grid"
    assert cfg["bodies"][0]["diameter"] == pytest.approx(0.2)  # radius given without diameter
    with pytest.raises(NotImplementedError):
        S.load_scene(write(tmp_path, "m.json", {"physics": {"contact_model": "mujoco"}}))
    kept = S.load_scene(write(tmp_path, "g.json", {"physics": {"gravity": [0, -10, 0]},
                                                   "scene": {"periodic": {"enabled": True},
                                                             "randomInit": {"enabled": True}}}))
    assert kept["physics"]["gravity"] == [0, -10, 0]


def test_endpoint_loader(tmp_path):
    rng = np.random.default_rng(3)
    e = rng.normal(size=(7, 6))
    txt = "# rod_radius=0.02\n# pbc=true\n# box_size=8\n" + "\n".join(" ".join(f"{v:.17g}" for v in r) for r in e) + "\n"
    E = S.load_endpoints(write(tmp_path, "x.txt", txt), S.load_scene(write(tmp_path, "s.json", {})))
    assert np.array_equal(E["ends"], e) and np.allclose(E["diameter"], 0.04)
    assert E["meta"] == {"pbc": "true", "box_size": "8"}
    csv = "x0,y0,z0,x1,y1,z1\n" + "\n".join(",".join(f"{v:.17g}" for v in r) + ",0.3" for r in e) + "\n"
    E2 = S.load_endpoints(write(tmp_path, "x.csv", csv), S.load_scene(write(tmp_path, "s.json", {})))
    assert np.array_equal(E2["ends"], e) and np.allclose(E2["diameter"], 0.3)  # schema header: 7th column is d


def test_endpoint_quaternion_maps_y_onto_direction():
    rng = np.random.default_rng(4)
    d = rng.normal(size=(50, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.vstack([d, [[0, 1, 0], [0, -1, 0], [1e-9, -1, 0]]])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for u in d:
        q = S.endpoint_quaternion(u)
        assert abs(np.linalg.norm(q) - 1) < 1e-12
        assert np.allclose(S._axis(q)[0], u, atol=1e-6)


def test_frame_rotation_round_trip():
    rng = np.random.default_rng(5)
    q = rng.normal(size=(20, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x = rng.normal(size=(20, 3))
    ql = S._qmul(S._QR, q)
    back = S._qmul(np.array([S._QR[0], -S._QR[1], 0, 0]), ql)
    assert np.allclose(back, q, atol=1e-15)
    # the rotated body axis is the rotated world axis; +Y (reference up) becomes +Z
    assert np.allclose(S._axis(ql), S._axis(q) @ S._R.T, atol=1e-14)
    assert np.allclose(S._R @ [0, 1, 0], [0, 0, 1]) and np.allclose((x @ S._R.T) @ S._R, x)
