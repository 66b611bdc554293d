"""Scene and state files in the reference's formats (host side).

* ``parse_state_csv`` reads an init-state CSV the way App::loadInitialStateCSV
  does (src/app/main.cpp:2935-3055): the first non-empty line is the header
  (columns trimmed, blank columns dropped), ``frame``, ``px`` and ``qw`` are
  required, every field is read with ``std::stod`` prefix semantics (a field
  that does not start with a number reads 0), rows with fewer fields than
  header columns are skipped, and the rods of the largest frame number are
  returned in file order.  Two differences, both on inputs the reference
  cannot handle: lines starting with ``#`` are skipped (the reference's own
  per-rod log, main.cpp:3412-3449, begins with two such lines, which its
  loader takes as the header and then rejects), and missing velocity columns
  read as zero (the reference indexes column -1).
* ``snapshot_line`` formats one NDJSON snapshot line like App::writeSnapshotLine
  (main.cpp:640-692, capsule branch), ``endpoints_text`` an endpoint file like
  App::saveInitialConfigCSV (main.cpp:2905-2933) and ``per_rod_text`` the rows
  of the per-rod log like App::logPerRodFrame (main.cpp:3412-3449), and
  ``network_text`` the hard-contact network log like App::logNetworkFrame
  (main.cpp:3624-3730, its hard-contact branch; ``network_soft_text`` its
  soft-contact branch), and ``profile_row`` a row of the profile log like
  App::logCsvFrame (main.cpp:3276-3355) with ``com``, ``gyration_sq`` and
  ``RelDisp`` for its packing columns (main.cpp:938-1022, 3500-3560).  The
  reference stores state in FP32 (glm::vec3 / glm::quat) and prints it with
  iostream defaults (``%g``, 6 significant digits) or ``std::fixed`` with 6
  decimals; the library's FP64 state is rounded to FP32 first and printed the
  same way, so a file written here reads back exactly as the reference's would.
"""
import math
import re

import numpy as np

_STOD = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.IGNORECASE)


def stod_prefix(s):
    """``std::stod`` on a field: leading blanks skipped, the longest numeric
    prefix parsed; no numeric prefix reads 0.0 (the loader's catch-all)."""
    m = _STOD.match(s)
    if not m:
        return 0.0
    try:
        return float(m.group(1))
    except ValueError:
        return 0.0


def _split_fields(line):
    """``std::getline(ss, tok, ',')`` fields: a trailing comma adds no field."""
    parts = line.split(",")
    if parts and parts[-1] == "":
        parts.pop()
    return parts


def parse_state_csv(path):
    """Last frame of an init-state CSV: (frame, p, v, w, q) with (n, 3) arrays
    and q as (w, x, y, z).  Raises ValueError where the reference's loader
    returns false."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    lines = [ln for ln in lines if ln and not ln.startswith("#")]
    if not lines:
        raise ValueError(f"[init-state] no lines in {path}")
    headers = [c.strip(" \t\r\n") for c in lines[0].split(",")]
    headers = [c for c in headers if c]
    col = {name: k for k, name in reversed(list(enumerate(headers)))}
    if "frame" not in col or "px" not in col or "qw" not in col:
        raise ValueError("[init-state] Missing required columns in CSV (need frame, px, qw, etc).")
    frames = {}
    for ln in lines[1:]:
        row = [stod_prefix(t) for t in _split_fields(ln)]
        if len(row) < len(headers):
            continue

        def get(name):
            k = col.get(name)
            return row[k] if k is not None else 0.0

        fr = int(row[col["frame"]])
        frames.setdefault(fr, []).append(
            ([get("px"), get("py"), get("pz")], [get("vx"), get("vy"), get("vz")],
             [get("wx"), get("wy"), get("wz")], [get("qw"), get("qx"), get("qy"), get("qz")]))
    if not frames:
        raise ValueError(f"[init-state] no complete rows in {path}")
    last = max(frames)
    rows = frames[last]
    p, v, w, q = (np.array([r[k] for r in rows], np.float64) for k in range(4))
    return last, p, v, w, q


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def cxx_g(v):
    """``std::ostream << v`` with default flags (6 significant digits, %g)."""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    return "%g" % v


def cxx_fixed6(v):
    """``std::ostream << std::fixed << std::setprecision(6) << v``."""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    return "%.6f" % v


def _vec(a):
    return "[" + ",".join(cxx_g(t) for t in a) + "]"


def snapshot_line(frame, dt, x, q, v, w, r, h):
    """One NDJSON snapshot line (with its newline) for capsules: positions,
    quaternions (w, x, y, z), velocities, angular velocities, radii and half
    heights in the reference's frame."""
    x, q, v, w, r, h = (_f32(a) for a in (x, q, v, w, r, h))
    t = float(frame) * float(np.float32(dt))
    out = ['{"frame":%d,"time":%s,"bodies":[' % (frame, cxx_g(t))]
    for i in range(len(r)):
        if i:
            out.append(",")
        out.append('{"id":%d,"shape":"capsule","pos":%s,"quat":%s,"vel":%s,"omega":%s,"radius":%s,"halfHeight":%s}'
                   % (i, _vec(x[i]), _vec(q[i]), _vec(v[i]), _vec(w[i]), cxx_g(r[i]), cxx_g(h[i])))
    out.append("]}\n")
    return "".join(out)


def glm_rotate_y(q):
    """FP32 ``q * vec3(0, 1, 0)`` in glm's operation order:
    v + 2 (w (u x v) + u x (u x v)) with u = (qx, qy, qz)."""
    q = np.asarray(q, np.float32)
    qw, qx, qy, qz = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    one, zero = np.float32(1.0), np.float32(0.0)
    # uv = u x (0, 1, 0) = (-qz, 0, qx)
    uvx, uvy, uvz = -qz, np.zeros_like(qx), qx
    # uuv = u x uv
    uuvx = qy * uvz - qz * uvy
    uuvy = qz * uvx - qx * uvz
    uuvz = qx * uvy - qy * uvx
    two = np.float32(2.0)
    return np.stack([zero + (uvx * qw + uuvx) * two, one + (uvy * qw + uuvy) * two, zero + (uvz * qw + uuvz) * two],
                    1).astype(np.float32)


def endpoints_text(x, q, r, h, pbc, box_size=None):
    """Endpoint file as the reference saves it: '# rod_length', '# rod_diameter'
    (of rod 0), '# pbc', '# box_size' (periodic only), the 'x0,...' header and
    one row of endpoints c -/+ u h per rod, all with 6 decimals."""
    x = np.asarray(x, np.float32)
    r = np.asarray(r, np.float32)
    h = np.asarray(h, np.float32)
    out = []
    if len(r):
        out.append("# rod_length=%s\n" % cxx_fixed6(h[0] * np.float32(2.0)))
        out.append("# rod_diameter=%s\n" % cxx_fixed6(r[0] * np.float32(2.0)))
    out.append("# pbc=%s\n" % ("true" if pbc else "false"))
    if pbc:
        out.append("# box_size=%s\n" % cxx_fixed6(np.float32(box_size)))
    out.append("x0,y0,z0,x1,y1,z1\n")
    u = glm_rotate_y(q)
    d = u * h[:, None]
    p0 = (x - d).astype(np.float32)
    p1 = (x + d).astype(np.float32)
    for a, b in zip(p0, p1):
        out.append(",".join(cxx_fixed6(t) for t in (*a, *b)) + "\n")
    return "".join(out)


NETWORK_HEADER = ("frame,rod_i,rod_j,contact_x,contact_y,contact_z,normal_x,normal_y,normal_z,phi,lambda_n,"
                  "lambda_t1,lambda_t2,vn_pre,vn_post\n")


def network_text(frame, a, b, point, normal, phi, lam, vn_pre, vn_post):
    """Rows of the hard-contact network log for one frame (App::logNetworkFrame,
    main.cpp:3699-3720): rod-rod manifolds only, contact point x_a + r_a,
    normal, gap, the three impulses and the normal relative velocity before and
    after the solve."""
    point, normal, lam = _f32(point), _f32(normal), _f32(lam)
    phi, vn_pre, vn_post = _f32(phi), _f32(vn_pre), _f32(vn_post)
    out = []
    for i in range(len(a)):
        vals = [*point[i], *normal[i], phi[i], *lam[i], vn_pre[i], vn_post[i]]
        out.append("%d,%d,%d," % (frame, a[i], b[i]) + ",".join(cxx_g(t) for t in vals) + "\n")
    return "".join(out)


NETWORK_SOFT_HEADER = ("frame,rod_i,rod_j,contact_x,contact_y,contact_z,normal_x,normal_y,normal_z,distance,"
                       "force_a_x,force_a_y,force_a_z,force_b_x,force_b_y,force_b_z,"
                       "friction_a_x,friction_a_y,friction_a_z,friction_b_x,friction_b_y,friction_b_z\n")


def network_soft_text(frame, a, b, point_a, point_b, normal, dist, force_b, friction_a):
    """Rows of the soft-contact network log for one frame (App::logNetworkFrame,
    main.cpp:3674-3688): the FP32 midpoint 0.5 (p_a + p_b), normal, distance,
    and the contact's forces; force_a = -force_b and friction_b = -friction_a
    (soft_contact.cpp:115-147 and its friction laws)."""
    pa, pb = np.asarray(point_a, np.float32), np.asarray(point_b, np.float32)
    mid = (np.float32(0.5) * (pa + pb)).astype(np.float64)
    normal, dist = _f32(normal), _f32(dist)
    fb, fra = _f32(force_b), _f32(friction_a)
    out = []
    for i in range(len(a)):
        vals = [*mid[i], *normal[i], dist[i], *(-fb[i]), *fb[i], *fra[i], *(-fra[i])]
        out.append("%d,%d,%d," % (frame, a[i], b[i]) + ",".join(cxx_g(t) for t in vals) + "\n")
    return "".join(out)


PROFILE_HEADER_SOFT = ("frame,rods,integrate_ms,sleep_ms,broadphase_ms,bpCount_ms,bpPrefix_ms,bpFill_ms,bpPairs_ms,"
                       "solve_ms,pbcWrap_ms,render_ms,contacts,collisions,KE,soft_PE,gyration_sq,reldisp_sq,"
                       "ent_pairs,ent_sum,step_ms\n")
PROFILE_HEADER_NSC = ("frame,rods,integrate_ms,sleep_ms,broadphase_ms,bpCount_ms,bpPrefix_ms,bpFill_ms,bpPairs_ms,"
                      "bpLongLong_ms,warmstart_ms,buildIslands_ms,solve_ms,floorSolve_ms,posCorrect_ms,pbcWrap_ms,"
                      "render_ms,contacts,collisions,islands,KE,KE_after_integrate,KE_after_warmstart,KE_after_solve,"
                      "KE_after_posCorrect,KE_after_pbcWrap,soft_PE,gyration_sq,reldisp_sq,jn_sum,jt_sum,"
                      "nsc_residual,ent_pairs,ent_sum,step_ms\n")


def profile_row(values):
    """One profile-log row: ints as integers, everything else with iostream
    defaults."""
    return ",".join(str(v) if isinstance(v, (int, np.integer)) else cxx_g(v) for v in values) + "\n"


def _box(lo, hi, pbc):
    return (np.asarray(hi, np.float32) - np.asarray(lo, np.float32)) if pbc else None


def _min_image(d, box):
    """d - box floor(d / box + 1/2) in FP32 on the axes with a positive box."""
    d = np.asarray(d, np.float32).copy()
    for k in range(3):
        if box is not None and box[k] > np.float32(0):
            d[..., k] -= box[k] * np.floor(d[..., k] / box[k] + np.float32(0.5))
    return d


def com(x, mass, lo, hi, pbc):
    """App::computeCOM: mass-weighted mean, on each periodic axis the circular
    mean of 2 pi (x - lo) / L mapped back into [lo, hi); FP32 result."""
    x = np.asarray(x, np.float32).astype(np.float64)
    m = np.asarray(mass, np.float32).astype(np.float64)
    out = np.zeros(3)
    if len(m) == 0:
        return out.astype(np.float32)
    box = _box(lo, hi, pbc)
    for k in range(3):
        if box is None or not box[k] > 0:
            out[k] = np.sum(m * x[:, k]) / np.sum(m)
        else:
            th = 2.0 * math.pi * ((x[:, k] - float(np.float32(lo[k]))) / float(box[k]))
            a = math.atan2(np.sum(m * np.sin(th)) / np.sum(m), np.sum(m * np.cos(th)) / np.sum(m))
            if a < 0.0:
                a += 2.0 * math.pi
            out[k] = float(np.float32(lo[k])) + (a / (2.0 * math.pi)) * float(box[k])
    return out.astype(np.float32)


def gyration_sq(x, mass, lo, hi, pbc):
    """App::computeGyrationSq: mean squared (minimum-image) distance to the COM."""
    if len(x) == 0:
        return 0.0
    d = _min_image(np.asarray(x, np.float32) - com(x, mass, lo, hi, pbc), _box(lo, hi, pbc))
    return float(np.sum(np.sum(d * d, axis=1, dtype=np.float32).astype(np.float64)) / len(d))


class RelDisp:
    """App::computeRelativeMotionSq: mean squared displacement since the last
    call relative to the COM's (minimum image); 0 on the first call."""

    def __init__(self):
        self.last = None

    def __call__(self, x, mass, lo, hi, pbc):
        x = np.asarray(x, np.float32)
        c = com(x, mass, lo, hi, pbc)
        if self.last is None or len(self.last[0]) != len(x):
            self.last = (x.copy(), c)
            return 0.0
        box = _box(lo, hi, pbc)
        dc = _min_image(c - self.last[1], box)
        u = _min_image(x - self.last[0], box) - dc
        self.last = (x.copy(), c)
        return float(np.sum(np.sum(u * u, axis=1, dtype=np.float32).astype(np.float64)) / len(x))


def per_rod_header(r0, h0):
    """The per-rod log's two comment lines and column header."""
    return ("# rod_radius=%s\n# rod_length=%s\n" % (cxx_g(np.float32(r0)), cxx_g(np.float32(h0) * np.float32(2.0)))
            + "frame,rod,px,py,pz,vx,vy,vz,wx,wy,wz,qw,qx,qy,qz,KE_lin,KE_rot,KE_total\n")


def per_rod_text(frame, x, v, w, q, mass, r, h):
    """Rows of the per-rod log for one frame: state, and the translational,
    rotational and total kinetic energy of each rod (solid-cylinder inertia
    m (3 r^2 + L^2) / 12 across the axis and m r^2 / 2 about it)."""
    xs, vs, ws, qs = (_f32(a) for a in (x, v, w, q))
    m = np.asarray(mass, np.float64)
    r = np.asarray(r, np.float64)
    h = np.asarray(h, np.float64)
    ke_lin = 0.5 * m * np.sum(vs * vs, axis=1)
    u = glm_rotate_y(qs).astype(np.float64)
    wa = np.sum(ws * u, axis=1)
    wp2 = np.sum(ws * ws, axis=1) - wa * wa
    i_perp = m * (3.0 * r * r + 4.0 * h * h) / 12.0
    i_ax = 0.5 * m * r * r
    ke_rot = 0.5 * (i_perp * wp2 + i_ax * wa * wa)
    out = []
    for i in range(len(m)):
        vals = [*xs[i], *vs[i], *ws[i], *qs[i], ke_lin[i], ke_rot[i], ke_lin[i] + ke_rot[i]]
        out.append("%d,%d," % (frame, i) + ",".join(cxx_g(t) for t in vals) + "\n")
    return "".join(out)
