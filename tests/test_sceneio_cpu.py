"""Scene and state files in the reference's formats (rod_dynamics_py.sceneio):
number formatting against the C++ standard library itself, the init-state
CSV parser's rules (App::loadInitialStateCSV, main.cpp:2935-3055), and round
trips of the snapshot, endpoint and per-rod writers through the readers."""
import json
import math
import shutil
import subprocess

import numpy as np
import pytest

from rod_dynamics_py import sceneio
from rod_dynamics_py.simulator import _axis, load_endpoints

CXX_FORMAT = r"""
#include <cstdint>
#include <cstring>
#include <iomanip>
#include <iostream>
int main() {
  uint32_t bits;
  while (std::cin >> std::hex >> bits) {
    float f;
    std::memcpy(&f, &bits, 4);
    std::cout << f << ' ';
    std::cout << std::fixed << std::setprecision(6) << f << '\n';
    std::cout.unsetf(std::ios::floatfield);
    std::cout << std::setprecision(6);
  }
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_number_formatting_matches_iostream(tmp_path):
    """cxx_g / cxx_fixed6 print FP32 values exactly as std::ostream does with
    default flags and with std::fixed << setprecision(6)."""
    src = tmp_path / "fmt.cpp"
    src.write_text(CXX_FORMAT)
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.normal(size=300) * 10.0 ** rng.integers(-8, 9, 300),
        rng.integers(-100000, 100000, 100).astype(float),
        [0.0, -0.0, 1.0, -1.0, 0.5, 1e-5, 123456.0, 1234567.0, 0.1, 2.0 / 3.0, 9.9999995, 999999.5, 1e20, -3e-30],
    ]).astype(np.float32)
    bits = vals.view(np.uint32)
    out = subprocess.run([str(exe)], input="\n".join("%x" % b for b in bits), capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert len(out) == len(vals)
    for v, line in zip(vals, out):
        g, fx = line.split(" ")
        assert sceneio.cxx_g(v) == g, (float(v), g)
        assert sceneio.cxx_fixed6(v) == fx, (float(v), fx)


def test_stod_prefix_semantics():
    assert sceneio.stod_prefix(" 1.5") == 1.5
    assert sceneio.stod_prefix("1.5abc") == 1.5
    assert sceneio.stod_prefix("-2e3x") == -2000.0
    assert sceneio.stod_prefix(".5") == 0.5
    assert sceneio.stod_prefix("abc") == 0.0
    assert sceneio.stod_prefix("") == 0.0
    assert math.isinf(sceneio.stod_prefix("inf"))


def test_state_csv_rules(tmp_path):
    """Header columns are trimmed and looked up by name (first match); rows
    shorter than the header are skipped; the largest frame number wins,
    whatever the row order; unparsable fields read 0."""
    p = tmp_path / "state.csv"
    p.write_text(
        "# comment line\n"
        " frame , rod ,px,py,pz,vx,vy,vz,wx,wy,wz,qw,qx,qy,qz\n"
        "7,0,1,2,3,0,0,0,0,0,0,1,0,0,0\n"
        "3,0,9,9,9,9,9,9,9,9,9,1,0,0,0\n"
        "7,1,4,5,x,0.5,0,0,0,0,1,0,1,0,0\n"
        "7,2,1,1\n"
        "\n")
    frame, pos, v, w, q = sceneio.parse_state_csv(p)
    assert frame == 7
    assert pos.tolist() == [[1, 2, 3], [4, 5, 0]]
    assert v.tolist() == [[0, 0, 0], [0.5, 0, 0]]
    assert w.tolist() == [[0, 0, 0], [0, 0, 1]]
    assert q.tolist() == [[1, 0, 0, 0], [0, 1, 0, 0]]


def test_state_csv_missing_columns(tmp_path):
    p = tmp_path / "bad.csv"
    p.write_text("rod,px,py,pz,qw\n0,1,2,3,1\n")
    with pytest.raises(ValueError, match="Missing required columns"):
        sceneio.parse_state_csv(p)
    p.write_text("")
    with pytest.raises(ValueError):
        sceneio.parse_state_csv(p)


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-5, 5, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q, rng.normal(size=(n, 3)), rng.normal(size=(n, 3))


def test_per_rod_log_round_trip(tmp_path):
    """The per-rod log written here reads back through the init-state parser
    (its comment lines skipped) to the FP32-rounded state."""
    n = 50
    x, q, v, w = _random_state(n, 2)
    r, h, m = np.full(n, 0.025), np.full(n, 0.25), np.full(n, 2.0)
    p = tmp_path / "per_rod.csv"
    p.write_text(sceneio.per_rod_header(r[0], h[0]) + sceneio.per_rod_text(4, x, v, w, q, m, r, h)
                 + sceneio.per_rod_text(9, x + 1.0, v, w, q, m, r, h))
    head = p.read_text().splitlines()[:3]
    assert head[0] == "# rod_radius=0.025" and head[1] == "# rod_length=0.5"
    assert head[2].startswith("frame,rod,px,py,pz,vx,vy,vz,wx,wy,wz,qw,qx,qy,qz,KE_lin,KE_rot,KE_total")
    frame, px, pv, pw, pq = sceneio.parse_state_csv(p)
    assert frame == 9
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(px, f32(x + 1.0), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pv, f32(v), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pw, f32(w), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pq, f32(q), rtol=1e-5, atol=1e-6)


def test_per_rod_kinetic_energy_closed_form():
    """A rod translating at speed s and spinning at rate a about an axis
    perpendicular to it: KE_lin = m s^2 / 2, KE_rot = I_perp a^2 / 2 with
    I_perp = m (3 r^2 + L^2) / 12; spin about its own axis: I_ax = m r^2 / 2."""
    m, r, h = 3.0, 0.1, 1.0
    q = np.array([[1.0, 0, 0, 0], [1.0, 0, 0, 0]])  # axis +Y
    x = np.zeros((2, 3))
    v = np.array([[2.0, 0, 0], [0, 0, 0]])
    w = np.array([[0, 0, 4.0], [0, 4.0, 0]])
    rows = sceneio.per_rod_text(0, x, v, w, q, [m, m], [r, r], [h, h]).splitlines()
    ke = [[float(t) for t in row.split(",")[-3:]] for row in rows]
    i_perp = m * (3 * r * r + 4 * h * h) / 12
    assert ke[0][0] == pytest.approx(0.5 * m * 4.0, rel=1e-6)
    assert ke[0][1] == pytest.approx(0.5 * i_perp * 16.0, rel=1e-5)
    assert ke[1][1] == pytest.approx(0.5 * 0.5 * m * r * r * 16.0, rel=1e-5)
    assert ke[0][2] == pytest.approx(ke[0][0] + ke[0][1], rel=1e-5)


def test_glm_axis_matches_rotation_matrix():
    _, q, _, _ = _random_state(500, 3)
    u32 = sceneio.glm_rotate_y(q)
    assert u32.dtype == np.float32
    np.testing.assert_allclose(u32, _axis(q), atol=3e-7)


def test_endpoint_file_round_trip(tmp_path):
    """save -> load: the endpoint loader reads back the rods (6 decimals) and
    the box metadata."""
    n = 40
    x, q, _, _ = _random_state(n, 4)
    r, h = np.full(n, 0.05), np.full(n, 0.5)
    p = tmp_path / "ends.txt"
    p.write_text(sceneio.endpoints_text(x, q, r, h, True, 12.0))
    lines = p.read_text().splitlines()
    assert lines[:5] == ["# rod_length=1.000000", "# rod_diameter=0.100000", "# pbc=true", "# box_size=12.000000",
                         "x0,y0,z0,x1,y1,z1"]
    cfg = {"populate": dict(count=0, density=0.0, radius=0.0), "bodies": []}
    E = load_endpoints(p, cfg)
    u = _axis(q)
    np.testing.assert_allclose(E["ends"][:, :3], x - u * h[:, None], atol=2e-6)
    np.testing.assert_allclose(E["ends"][:, 3:], x + u * h[:, None], atol=2e-6)
    assert np.allclose(E["diameter"], 0.1)
    assert E["meta"]["pbc"] == "true" and E["meta"]["box_size"] == "12.000000"


def test_snapshot_line_fields():
    n = 5
    x, q, v, w = _random_state(n, 5)
    r, h = np.full(n, 0.05), np.full(n, 0.5)
    line = sceneio.snapshot_line(120, 1.0 / 600.0, x, q, v, w, r, h)
    assert line.endswith("]}\n") and line.count("\n") == 1
    d = json.loads(line)
    assert d["frame"] == 120 and d["time"] == pytest.approx(0.2, rel=1e-5)
    assert [b["id"] for b in d["bodies"]] == list(range(n))
    for i, b in enumerate(d["bodies"]):
        assert b["shape"] == "capsule"
        assert list(b) == ["id", "shape", "pos", "quat", "vel", "omega", "radius", "halfHeight"]
        np.testing.assert_allclose(b["pos"], x[i], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(b["quat"], q[i], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(b["vel"], v[i], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(b["omega"], w[i], rtol=1e-5, atol=1e-6)
        assert b["radius"] == 0.05 and b["halfHeight"] == 0.5
    # float32 time as the reference computes it: double(frame) * double(float dt)
    assert d["time"] == float("%g" % (120 * float(np.float32(1.0 / 600.0))))


def test_profile_log_packing_columns_closed_forms():
    """COM, gyration and relative displacement of the profile log
    (main.cpp:938-1022, 3500-3560) on configurations with known answers: the
    mass-weighted mean off a periodic box, the circular mean across a periodic
    face (the plain mean would sit mid-box), minimum images, and displacements
    relative to the COM's."""
    from rod_dynamics_py import sceneio as S

    lo, hi = [0.0, 0.0, 0.0], [10.0, 10.0, 10.0]
    x = np.array([[0.0, 1.0, 1.0], [4.0, 1.0, 1.0]])
    assert np.allclose(S.com(x, [1.0, 3.0], lo, hi, False), [3.0, 1.0, 1.0])
    assert S.gyration_sq(x, [1.0, 3.0], lo, hi, False) == pytest.approx(5.0, rel=1e-6)
    xp = np.array([[0.1, 5.0, 5.0], [9.9, 5.0, 5.0]])
    c = S.com(xp, [1.0, 1.0], lo, hi, True)
    assert min(abs(c[0]), abs(c[0] - 10.0)) < 1e-5 and abs(c[1] - 5.0) < 1e-5
    assert S.gyration_sq(xp, [1.0, 1.0], lo, hi, True) == pytest.approx(0.01, rel=1e-4)
    R = S.RelDisp()
    assert R(x, [1.0, 1.0], lo, hi, False) == 0.0
    assert R(x + [1.0, 2.0, 3.0], [1.0, 1.0], lo, hi, False) == pytest.approx(0.0, abs=1e-10)
    x2 = x + [1.0, 2.0, 3.0]
    x2[0, 0] += 1.0  # one of two equal rods moves: the COM moves half as far
    assert R(x2, [1.0, 1.0], lo, hi, False) == pytest.approx(0.25, rel=1e-6)
    Rp = S.RelDisp()
    a = np.array([[9.95, 5.0, 5.0], [3.0, 5.0, 5.0]])
    Rp(a, [1.0, 1.0], lo, hi, True)
    b = a.copy()
    b[0, 0] = 0.05  # crosses the face: a step of 0.1, not -9.9
    assert Rp(b, [1.0, 1.0], lo, hi, True) == pytest.approx(0.05 ** 2, rel=1e-3)


def test_profile_log_headers_and_row_format():
    from rod_dynamics_py import sceneio as S

    assert len(S.PROFILE_HEADER_NSC.strip().split(",")) == 35
    assert len(S.PROFILE_HEADER_SOFT.strip().split(",")) == 21
    assert S.profile_row([3, 10, 0.0, -1, 1.5e-7, 123456789.0]) == "3,10,0,-1,1.5e-07,1.23457e+08\n"
