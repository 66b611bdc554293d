"""Reference packings (assets/initial-configs/*/x_relaxed.txt) as rod bodies.

Each rod of a packing is given by its two endpoints and the packing's rod
radius.  A capsule of half length h along the body's local +Y axis with
centre c and orientation q reproduces the segment: c is the midpoint,
h = |p1 - p0| / 2 and q is the shortest rotation taking +Y onto the unit
direction u, q = normalize(1 + u_y, u_z, 0, -u_x) in (w, x, y, z) order.
The capsule is symmetric under u -> -u, so u is flipped into u_y >= 0 and
the rotation never degenerates.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_packings.npz")


def rods_from_endpoints(ends, radius):
    p0, p1 = ends[:, :3], ends[:, 3:6]
    c = 0.5 * (p0 + p1)
    d = p1 - p0
    length = np.linalg.norm(d, axis=1)
    u = d / length[:, None]
    u[u[:, 1] < 0] *= -1.0
    q = np.column_stack([1.0 + u[:, 1], u[:, 2], np.zeros(len(u)), -u[:, 0]])
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    n = len(c)
    return c, q, np.full(n, float(radius)), 0.5 * length


def golden_packings():
    """Yields (packing_id, x, q, r, h, n_contacts) for the committed fixture."""
    z = np.load(GOLDEN)
    for k in range(int(z["count"])):
        x, q, r, h = rods_from_endpoints(z[f"ends_{k}"], float(z[f"radius_{k}"]))
        yield str(z[f"id_{k}"]), x, q, r, h, int(z[f"n_contacts_{k}"])


def contact_delta(r):
    """The reference's contact criterion for these packings is a segment
    distance below 1.01 d (repro/build_packing_metadata.py), i.e. a capsule
    gap below 0.01 d."""
    return 0.01 * 2.0 * float(r[0])


def free_box(x):
    """A non-periodic box enclosing the packing with room to spare."""
    return x.min(0) - 10.0, x.max(0) + 10.0
