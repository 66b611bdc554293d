"""Oracle pins of the entanglement observer (oracle/rd_oracle.cpp
orc_linking_number / orc_entanglement, following App::computeEntanglement,
main.cpp:4287-4331, and entanglement_adapter.cpp:56-117).

The linking number of two straight segments is pinned by
  * the closed form of perpendicular segments,
  * Gauss-Legendre quadrature of the Gauss double integral on random skew pairs
    (sign included),
  * its symmetries (exchange, reversal, mirror image, coplanar -> 0);
the pair sum by the reference's published ebar0 of its frozen packings
(assets/packings_metadata.csv, committed in tests/golden/reference_packings.npz)
and by the periodic wrap of endpoints."""
import math
import os

import numpy as np
import pytest

import oracle_py as O


def gauss_lk(a0, a1, b0, b1, n=160):
    """(1/4pi) int int (dr1 x dr2).(r1 - r2) / |r1 - r2|^3 by Gauss-Legendre."""
    xs, ws = np.polynomial.legendre.leggauss(n)
    s = 0.5 * (xs + 1.0)
    w = 0.5 * ws
    da, db = a1 - a0, b1 - b0
    r1 = a0[None, :] + s[:, None] * da[None, :]
    r2 = b0[None, :] + s[:, None] * db[None, :]
    d = r1[:, None, :] - r2[None, :, :]
    cr = np.cross(da, db)
    f = (d @ cr) / np.linalg.norm(d, axis=2) ** 3
    return float(w @ f @ w) / (4.0 * math.pi)


@pytest.mark.parametrize("a,b,delta", [(1.0, 1.0, 0.3), (0.5, 2.0, 1.0), (3.0, 0.2, 0.05), (1.0, 1.0, 10.0)])
def test_perpendicular_segments_closed_form(a, b, delta):
    """Segments of half lengths a (along x) and b (along y, at height delta),
    crossing over their midpoints: |Lk| = atan(ab / (delta sqrt(a^2 + b^2 +
    delta^2))) / pi (four quadrants of the solid angle of a rectangle)."""
    lk = O.linking_number(np.array([-a, 0, 0.0]), np.array([a, 0, 0.0]),
                          np.array([0, -b, delta]), np.array([0, b, delta]))
    ref = math.atan(a * b / (delta * math.sqrt(a * a + b * b + delta * delta))) / math.pi
    assert abs(abs(lk) - ref) < 1e-5 * max(ref, 1e-3), (lk, ref)


def test_corner_segments_closed_form():
    """Segments from the origin's projections: one quadrant of the rectangle,
    |Lk| = atan(ab / (delta sqrt(a^2 + b^2 + delta^2))) / (4 pi)."""
    a, b, d = 1.3, 0.7, 0.4
    lk = O.linking_number(np.zeros(3), np.array([a, 0, 0.0]), np.array([0, 0, d]), np.array([0, b, d]))
    ref = math.atan(a * b / (d * math.sqrt(a * a + b * b + d * d))) / (4.0 * math.pi)
    assert abs(abs(lk) - ref) < 1e-6


def test_linking_number_against_gauss_integral():
    """Random skew segment pairs kept apart (closest distance above 0.2 of the
    lengths, so the quadrature converges): the solid-angle formula equals the
    Gauss double integral, sign included, to the 1e-6 normalisation offset."""
    rng = np.random.default_rng(11)
    done = 0
    while done < 60:
        a0, a1, b0, b1 = rng.normal(size=(4, 3))
        s, t = O.segseg(a0, a1, b0, b1)
        gap = np.linalg.norm(a0 + s * (a1 - a0) - b0 - t * (b1 - b0))
        if gap < 0.2 * max(np.linalg.norm(a1 - a0), np.linalg.norm(b1 - b0)):
            continue
        lk = O.linking_number(a0, a1, b0, b1)
        g = gauss_lk(a0, a1, b0, b1)
        assert abs(lk - g) < 2e-5 * max(1.0, abs(g) * 10), (done, lk, g)
        done += 1


def test_linking_number_symmetries():
    rng = np.random.default_rng(3)
    mirror = np.diag([1.0, 1.0, -1.0])
    for _ in range(200):
        a0, a1, b0, b1 = rng.normal(size=(4, 3))
        lk = O.linking_number(a0, a1, b0, b1)
        # exchange of the two segments
        assert abs(O.linking_number(b0, b1, a0, a1) - lk) < 1e-12
        # reversing one segment flips the sign, reversing both keeps it
        assert abs(O.linking_number(a1, a0, b0, b1) + lk) < 1e-12
        assert abs(O.linking_number(a1, a0, b1, b0) - lk) < 1e-12
        # a mirror image flips the sign; a rigid motion keeps it
        assert abs(O.linking_number(mirror @ a0, mirror @ a1, mirror @ b0, mirror @ b1) + lk) < 1e-12
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        sh = rng.normal(size=3)
        assert abs(O.linking_number(Q @ a0 + sh, Q @ a1 + sh, Q @ b0 + sh, Q @ b1 + sh) - lk) < 1e-9
        assert abs(lk) <= 0.5


def test_coplanar_and_parallel_segments_do_not_link():
    assert O.linking_number(np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([1.0, 1.0, 0])) == 0.0
    assert O.linking_number(np.zeros(3), np.array([1.0, 0, 0]), np.array([0.5, -1, 0]), np.array([0.5, 1, 0])) == 0.0


def _world(ends, radius, lo=None, hi=None, periodic=(0, 0, 0)):
    from packings import free_box, rods_from_endpoints

    x, q, r, h = rods_from_endpoints(ends, radius)
    W = O.World(x, q, None, None, r, h, np.ones(len(r)))
    if lo is None:
        lo, hi = free_box(x)
    W.set_box(lo, hi, periodic)
    return W


def test_reference_packings_ebar0():
    """Frozen relaxed packings (aspect ratios 10 to 1000): the mean |Lk| over
    all pairs equals the reference's ebar0 (rounded to 6 digits in its
    metadata; its own arithmetic differs in the last digits) and the pair count
    equals its ent_pairs."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_packings.npz"))
    for k in range(int(z["count"])):
        W = _world(z[f"ends_{k}"], float(z[f"radius_{k}"]))
        s, n = W.entanglement(-1.0)
        assert n == int(z[f"ent_pairs_{k}"])
        assert abs(s / n - float(z[f"ebar0_{k}"])) < 1.5e-6, (k, s / n, float(z[f"ebar0_{k}"]))


@pytest.mark.skipif(not os.path.exists("/root/reference/assets/packings_metadata.csv"),
                    reason="the reference tree is not present")
def test_all_reference_packings_ebar0():
    import csv

    ref_root = "/root/reference"
    rows = list(csv.DictReader(open(os.path.join(ref_root, "assets", "packings_metadata.csv"))))
    assert len(rows) >= 60
    for row in rows:
        radius, ends = None, []
        for line in open(os.path.join(ref_root, row["path"])):
            line = line.strip()
            if line.startswith("#"):
                if "rod_radius" in line:
                    radius = float(line.split("=")[1].split()[0])
            elif line:
                ends.append([float(t) for t in line.split()[:6]])
        s, n = _world(np.asarray(ends), radius).entanglement(-1.0)
        assert n == int(row["ent_pairs"])
        assert abs(s / n - float(row["ebar0"])) < 1.5e-6, row["packing_id"]


def test_cutoff_selects_pairs_by_segment_distance():
    """With a cutoff the pair set is exactly the pairs whose axis segments are
    at most cutoff apart (brute force over segseg)."""
    rng = np.random.default_rng(8)
    n = 60
    p0 = rng.uniform(0, 4, size=(n, 3))
    ends = np.hstack([p0, p0 + rng.normal(size=(n, 3))])
    W = _world(ends, 0.05)
    for cut in (0.2, 0.7, 1.5):
        s, m = W.entanglement(cut)
        want_s, want_m = 0.0, 0
        for i in range(n):
            for j in range(i + 1, n):
                a0, a1, b0, b1 = ends[i, :3], ends[i, 3:], ends[j, :3], ends[j, 3:]
                u, v = O.segseg(a0, a1, b0, b1)
                if np.linalg.norm(a0 + u * (a1 - a0) - b0 - v * (b1 - b0)) <= cut:
                    want_s += abs(O.linking_number(a0, a1, b0, b1))
                    want_m += 1
        assert m == want_m and 0 < m < n * (n - 1) // 2
        assert abs(s - want_s) < 1e-9


def test_periodic_endpoints_wrapped_into_box():
    """On a periodic world each endpoint is wrapped into the box on its own
    before the pair sum (main.cpp:4294-4300): a rod straddling the boundary
    becomes a segment across the box."""
    ends = np.array([[9.5, 5.0, 5.0, 10.5, 5.2, 5.1],     # crosses x = 10
                     [9.9, 4.0, 4.6, 9.9, 6.0, 5.4],
                     [2.0, 2.0, 2.0, 3.0, 3.0, 3.5]])
    lo, hi = np.zeros(3), np.full(3, 10.0)
    s_per, n_per = _world(ends, 0.05, lo, hi, (1, 1, 1)).entanglement(-1.0)
    s_free, _ = _world(ends, 0.05, lo, hi, (0, 0, 0)).entanglement(-1.0)
    wrapped = ends.copy()
    wrapped[0, 3] -= 10.0
    want = sum(abs(O.linking_number(wrapped[i, :3], wrapped[i, 3:], wrapped[j, :3], wrapped[j, 3:]))
               for i in range(3) for j in range(i + 1, 3))
    assert n_per == 3
    assert abs(s_per - want) < 1e-12
    assert abs(s_per - s_free) > 1e-3
