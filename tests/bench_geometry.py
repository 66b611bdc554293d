"""The reference broad-phase benchmark geometry, regenerated in numpy
(src/tools/benchmark_broadphase.cpp:53-74,105-113): std::mt19937(42) through
libstdc++'s uniform_real_distribution<float> (one 32-bit draw per value,
float(x) / 2^32), positions first then three uniforms for a Shoemake
quaternion, everything in float32.  GCC evaluates the three pos(rng) arguments
of the position constructor right to left, so the first draw is z."""
import numpy as np


def _canonical(n, seed=42):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)  # init_genrand, the std::mt19937(seed) state
    raw = bg.random_raw(n).astype(np.uint64)
    f = raw.astype(np.float32) / np.float32(4294967296.0)
    return np.where(f >= 1, np.nextafter(np.float32(1), np.float32(0)), f).astype(np.float32)


def bench_rods(n, seed=42):
    r, hh = np.float32(0.05), np.float32(0.35)
    pi = np.float32(np.pi)
    vrod = pi * r * r * np.float32(2.0) * hh + (np.float32(4.0) / np.float32(3.0)) * pi * r * r * r
    box = np.float32(np.cbrt(np.float32(n) * vrod / np.float32(0.30)))
    c = _canonical(6 * n, seed).reshape(n, 6)
    a, b = -box * np.float32(0.5), box * np.float32(0.5)
    p = (c[:, 2::-1] * (b - a) + a).astype(np.float32)
    u1, u2, u3 = c[:, 3], c[:, 4], c[:, 5]
    s1, s2 = np.sqrt(np.float32(1) - u1), np.sqrt(u1)
    t1, t2 = np.float32(2) * pi * u2, np.float32(2) * pi * u3
    q = np.stack([s1 * np.sin(t1), s1 * np.cos(t1), s2 * np.sin(t2), s2 * np.cos(t2)], 1).astype(np.float32)
    q = (q / np.sqrt((q * q).sum(1, keepdims=True))).astype(np.float32)
    return (p.astype(np.float64), q.astype(np.float64), np.full(n, float(r)), np.full(n, float(hh)),
            float(a), float(b))
