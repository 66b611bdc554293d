"""Oracle pins of the reference's non-overlapping populate
(oracle/rsa_oracle.cpp, following App::resetScene populate.mode "nonoverlap",
src/app/main.cpp:1857-2140).

Pinned by: the C++ standard's known answer of std::mt19937 and libstdc++'s
float canonical mapping (the candidate stream); closed forms of the FP32
segment distance, including the periodic variant's end-point rule; the placed
rods never overlapping (FP64 detection); results independent of the placement
grid's cell size on free boxes (the grid only prunes pairs); uniform centre
and direction statistics."""
import math

import numpy as np
import pytest

import oracle_py as O


def test_mt19937_known_answer():
    """The C++ standard fixes the 10000th output of a default-seeded
    std::mt19937 (seed 5489) at 4123659995."""
    assert O.mt_raw(5489, 10000) == 4123659995


def test_uniform_float_is_raw_over_2_32():
    """uniform_real_distribution<float>(0, 1) on a 32-bit engine: one engine
    output per number, float(raw) / 2^32, kept below 1."""
    seed = 1234
    got = O.mt_uniform_float(seed, 2000)
    raw = np.array([O.mt_raw(seed, k + 1) for k in range(2000)], np.uint64)
    want = np.float32(raw.astype(np.float32)) / np.float32(2.0 ** 32)
    want = np.where(want >= 1.0, np.nextafter(np.float32(1.0), np.float32(0.0)), want)
    assert np.array_equal(got, want.astype(np.float32))


def _seg_dist_fp64(p0, p1, q0, q1):
    """Squared FP64 segment distance at the oracle's Lumelsky closest points
    (pinned in test_oracle_geometry.py)."""
    s, t = O.segseg(np.asarray(p0, float), np.asarray(p1, float), np.asarray(q0, float), np.asarray(q1, float))
    a = np.asarray(p0, float) + s * (np.asarray(p1, float) - np.asarray(p0, float))
    b = np.asarray(q0, float) + t * (np.asarray(q1, float) - np.asarray(q0, float))
    return float(np.dot(a - b, a - b))


def test_segment_distance_free_rule_matches_fp64():
    """Without the periodic rule the FP32 squared distance equals the FP64
    segment distance to FP32 rounding."""
    rng = np.random.default_rng(0)
    for _ in range(500):
        p0, p1, q0, q1 = rng.normal(size=(4, 3)).astype(np.float32)
        d2 = O.rsa_seg_d2(p0, p1, q0, q1, 0)
        ref = _seg_dist_fp64(p0, p1, q0, q1)
        assert abs(d2 - ref) <= 2e-5 * max(ref, 1e-3), (d2, ref)


def test_segment_distance_periodic_rule():
    """The periodic variant (main.cpp:1978-1986) leaves s unclamped above 1 at
    an end point of q: the distance is then that end point's distance to p's
    infinite line, below the true segment distance; otherwise it equals the
    free rule."""
    p0, p1 = np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32)
    q0, q1 = np.array([3, 0.5, 0], np.float32), np.array([3, 2.0, 0], np.float32)  # beyond p1, t clamps to 0
    free = O.rsa_seg_d2(p0, p1, q0, q1, 0)
    per = O.rsa_seg_d2(p0, p1, q0, q1, 1)
    assert free == pytest.approx(2.0 ** 2 + 0.5 ** 2, rel=1e-6)   # end point p1 to q0
    assert per == pytest.approx(0.5 ** 2, rel=1e-6)                 # q0 to the x axis
    rng = np.random.default_rng(1)
    same = 0
    for _ in range(500):
        p0, p1, q0, q1 = rng.normal(size=(4, 3)).astype(np.float32)
        a, b = O.rsa_seg_d2(p0, p1, q0, q1, 0), O.rsa_seg_d2(p0, p1, q0, q1, 1)
        assert b <= a * (1 + 1e-6) + 1e-12
        same += a == b
    assert same > 250  # most random pairs meet at interior points or below s = 1


def _overlaps(c, u, length, diameter, lo, hi, periodic):
    """FP64 detection of the placed capsules; returns the smallest axis
    distance found among pairs closer than one diameter (inf if none)."""
    from packings import rods_from_endpoints

    c = c.astype(np.float64)
    u = u.astype(np.float64)
    ends = np.hstack([c - u * length / 2, c + u * length / 2])
    x, q, r, h = rods_from_endpoints(ends, diameter / 2)
    W = O.World(x, q, None, None, r, h, np.ones(len(r)))
    W.set_box(np.asarray(lo, float), np.asarray(hi, float), (1, 1, 1) if periodic else (0, 0, 0))
    nc = W.detect(0, 0.0)
    if nc == 0:
        return math.inf
    return float(W.contacts(nc)["dist"].min())


@pytest.mark.parametrize("periodic", [0, 1])
def test_placed_rods_do_not_overlap(periodic):
    n, L, d = 1500, 0.5, 0.05
    lo, hi = [-1, -1, -1], [1, 1, 1]
    k, c, u, att = O.populate_nonoverlap(n, L, d, lo, hi, periodic, 0.6, 11, 1000)
    assert 0 < k <= n and att >= k
    assert np.all(c >= np.float32(-1)) and np.all(c < np.float32(1))
    assert np.allclose(np.linalg.norm(u.astype(np.float64), axis=1), 1.0, atol=3e-7)
    dmin = _overlaps(c, u, L, d, lo, hi, periodic)
    assert dmin >= d * (1 - 1e-5), dmin


def test_stop_after_max_attempts():
    """Near saturation a rod fails max_attempts times in a row and the
    populate ends there: the last max_attempts candidates were all rejected."""
    k, c, u, att = O.populate_nonoverlap(5000, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 0, 0.6, 5, 50)
    assert k < 5000
    k2, c2, u2, att2 = O.populate_nonoverlap(k, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 0, 0.6, 5, 50)
    assert k2 == k and np.array_equal(c2, c) and att - att2 == 50


def test_free_box_result_independent_of_cell_size():
    """On a free box the grid only prunes pairs whose AABBs are apart, which
    never collide: any cell size (including one cell for the whole box)
    places the same rods."""
    args = (900, 0.5, 0.05, [-1, -1, -1], [1, 1, 1], 0)
    ref = O.populate_nonoverlap(*args, 0.6, 21, 1000)
    for cs in (0.0, 0.25, 2.0, 5.0):
        got = O.populate_nonoverlap(*args, cs, 21, 1000)
        assert got[0] == ref[0] and np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]), cs


def test_prefix_property():
    """Asking for fewer rods returns a prefix of the same placement."""
    a = O.populate_nonoverlap(400, 0.5, 0.05, [0, 0, 0], [2, 2, 2], 1, 0.6, 3, 1000)
    b = O.populate_nonoverlap(250, 0.5, 0.05, [0, 0, 0], [2, 2, 2], 1, 0.6, 3, 1000)
    assert b[0] == 250 and np.array_equal(a[1][:250], b[1]) and np.array_equal(a[2][:250], b[2])


def test_uniform_statistics():
    """Dilute packing: centres uniform in the box, directions uniform on S^2."""
    n = 4000
    k, c, u, _ = O.populate_nonoverlap(n, 0.2, 0.02, [0, 0, 0], [10, 10, 10], 1, 0.6, 8, 1000)
    assert k == n
    c, u = c.astype(float), u.astype(float)
    assert np.all(np.abs(c.mean(0) - 5.0) < 4 * 10 / math.sqrt(12 * n))
    assert np.all(np.abs(u.mean(0)) < 4 / math.sqrt(3 * n))
    m2 = (u ** 2).mean(0)
    assert np.all(np.abs(m2 - 1 / 3) < 4 * math.sqrt(4 / 45 / n))
