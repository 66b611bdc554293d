"""The reference's Python session API on top of librdk.

``Simulator(scene_path, init_csv_path="", quiet=True)`` with ``step(steps=1)``,
``diagnostics()`` and ``rods()`` mirrors the pybind module of the reference
(src/python/rod_dynamics_py.cpp:56-97, include/app/python_api.hpp): a scene
JSON in the reference's schema (src/config/config.cpp:259-718, defaults
config.cpp:defaultAppCfg and include/config/config.hpp), optionally an
endpoint file (App::loadInitialConfigCSV, src/app/main.cpp), and one
physicsStep per step (main.cpp:4333-4858) run by the library's fused steps:

* ``physics.nsc.enabled``: the hard-contact step (free flight, detection,
  PSOR with the reference's stop rule max|dlambda| < 1e-6 or
  ``velocity_iters`` sweeps, sequential friction clamp, previous-step warm
  start, position update, position stabilisation when enabled);
* ``physics.hertz_mindlin.enabled`` / ``physics.soft_contact.enabled`` (or
  ``physics.contact_model``): velocity Verlet with the force evaluated twice
  per step, as the reference's soft paths do.

Beyond the pybind surface, the reference app's file I/O is kept:
``init_state_path`` loads the last frame of an init-state CSV (the CLI's
``--init-state``, App::loadInitialStateCSV, main.cpp:2935-3055), and
``write_snapshot``, ``save_endpoints``, ``write_per_rod`` and
``write_network`` write NDJSON snapshots, endpoint files, the per-rod log and
the hard-contact network log in the reference's formats (``sceneio``).

The reference is y-up (default gravity (0, -10, 0), floors are box top faces
in y); the library's plane walls are floors in z.  Scenes are therefore run in
the frame rotated by +90 degrees about x, (x, y, z) -> (x, -z, y), and every
quantity is rotated back in ``rods()``: a rotation leaves the dynamics
unchanged.

populate.mode "nonoverlap" places the reference's rods bit for bit (its FP32
serial placement with its std::mt19937 stream, evaluated in parallel on the
GPU by ``rdk_populate_nonoverlap``).

randomInit draws the reference's own stream (std::mt19937 with fresh
std::normal_distribution<float> objects per rod, FP32, main.cpp:2213-2294,
``rdk_random_init``) from the FP32 quaternions the reference's rods hold
(glm::normalize of the populate quaternion), so a scene populated with
"nonoverlap" starts from the reference's state bit for bit (given a reference
build without FMA contraction).

What differs from the reference, by design: populate "grid"/"uniform" come
from numpy, not libstdc++'s mt19937 + distributions, so those scenes are
statistically but not bitwise the reference's (and randomInit then sees their
orientations rounded to FP32); ``last_island_count`` reports the number of contact
islands of the last hard-contact solve (the reference never updates that
field); the soft/Hertz-Mindlin paths use the library's capsule cell list
instead of the reference's sphere grid.  Options the library does not model
(spheres, cylinder tube walls, rotated floors, random forces, fixed rods, the
MuJoCo-style model) raise ``NotImplementedError`` instead of being ignored.
"""
import json
import math
import time
import sys

import numpy as np

from . import _lib as L
from . import sceneio

# defaults of the reference's configuration (include/config/config.hpp,
# src/config/config.cpp:defaultAppCfg)
_PHYSICS = dict(dt=1.0 / 600.0, gravity=[0.0, -10.0, 0.0], lin_damp=0.08, ang_damp=0.12, w_max=80.0,
                contact_model="")
_NSC = dict(enabled=False, mu=0.3, beta=0.2, cfm=0.0, omega=1.0, velocity_iters=40, position_stabilization=True,
            position_iters=5, position_psor_iters=50, slop=1e-4, enable_warm_start=True)
_SOFT = dict(enabled=False, delta=0.005, k_scaler=1000.0, damping=0.0, mu=0.5, mu_static=0.0, nu=1e-5,
             enable_friction=True, friction_karnopp=False, friction_cundall=False, kt=1000.0, vel_deadband=1e-3)
_HM = dict(enabled=False, youngs_modulus=7e10, poisson_ratio=0.25, restitution_coeff=0.9, friction_coeff=0.3,
           friction_static_coeff=0.5, friction_transition_vel=1e-3, rolling_friction_coeff=0.01,
           enable_tangential=True, enable_rolling=True)
_FLOOR = dict(enabled=False, pos=[0.0, -0.8, 0.0], rot_quat=[1.0, 0.0, 0.0, 0.0], half_extents=[10.0, 0.1, 10.0],
              restitution=0.3, friction=0.9)
_PERIODIC = dict(enabled=False, min=[-3.0, -1.0, -3.0], max=[3.0, 3.0, 3.0], cellSize=0.6, longSpan=4)
_RANDOM_INIT = dict(enabled=False, mode="thermal", vSigma=0.0, wSpeed=0.0, wSigma=0.0, kBT=1.0, kBTTrans=-1.0,
                    kBTRot=-1.0, seed=0, projectParallelSpin=True)
_POPULATE = dict(count=0, grid=False, spacingMul=1.0, seed=0, mode="uniform", maxAttempts=200000,
                 shape="capsule", radius=0.05, density=2500.0)
_BODY = dict(shape="capsule", length=0.5, diameter=0.1, radius=0.1, density=1000.0, restitution=0.2, friction=0.7,
             is_static=False, pos=[0.0, 0.0, 0.0], rot_quat=[1.0, 0.0, 0.0, 0.0], v_lin=[0.0, 0.0, 0.0],
             v_ang=[0.0, 0.0, 0.0])

# +90 degrees about x: the reference's y-up frame -> the library's z-up frame
_R = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])
_QR = np.array([math.sqrt(0.5), math.sqrt(0.5), 0.0, 0.0])


def _merge(defaults, given):
    out = dict(defaults)
    for k, v in (given or {}).items():
        out[k] = v
    return out


def _qmul(a, b):
    """Hamilton product of (w, x, y, z) quaternions, row-wise."""
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    w1, x1, y1, z1 = a.T
    w2, x2, y2, z2 = b.T
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], 1)


def _axis(q):
    """Body +Y axis in the world for (w, x, y, z) quaternions (column 1 of R)."""
    w, x, y, z = np.atleast_2d(q).T
    return np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], 1)


def load_scene(path):
    """Scene JSON in the reference's schema -> dict of merged sections.
    Unknown keys are kept (the reference warns about them and ignores them)."""
    with open(path) as f:
        j = json.load(f)
    ph = j.get("physics", {})
    sc = j.get("scene", {})
    cfg = {"physics": _merge(_PHYSICS, {k: v for k, v in ph.items() if not isinstance(v, dict)}),
           "nsc": _merge(_NSC, ph.get("nsc")),
           "soft": _merge(_SOFT, ph.get("soft_contact")),
           "hm": _merge(_HM, ph.get("hertz_mindlin")),
           "floor": _merge(_FLOOR, sc.get("floor")),
           "periodic": _merge(_PERIODIC, sc.get("periodic")),
           "randomInit": _merge(_RANDOM_INIT, sc.get("randomInit")),
           "populate": _merge(_POPULATE, sc.get("populate")),
           "bodies": [], "initCsvPath": sc.get("initCsvPath", sc.get("initCsv", "")),
           "cylinder": sc.get("cylinder", {}), "randomForce": sc.get("randomForce", {}),
           "fixed": bool(sc.get("fixCentroidRod", False)) or int(sc.get("fixEveryExcept", -1)) >= 0}
    if "populate" in sc and "mode" not in sc["populate"]:
        cfg["populate"]["mode"] = "grid" if cfg["populate"]["grid"] else "uniform"
    for b in sc.get("bodies", []) or []:
        bb = _merge(_BODY, b)
        if bb["shape"] != "sphere" and "radius" in b and "diameter" not in b:
            bb["diameter"] = 2.0 * bb["radius"]
        cfg["bodies"].append(bb)
    if "use_soft_contact" in ph:
        cfg["soft"]["enabled"] = bool(ph["use_soft_contact"])
    model = cfg["physics"].get("contact_model") or ""
    if model:
        sel = {"nsc": "nsc", "harmonic": "soft", "hertz-mindlin": "hm"}.get(model)
        if sel is None:
            raise NotImplementedError(f"contact_model {model!r} is not implemented by this library")
        for k in ("nsc", "soft", "hm"):
            cfg[k]["enabled"] = k == sel
    # periodic + randomInit with the default gravity: the reference zeroes it
    if cfg["periodic"]["enabled"] and cfg["randomInit"]["enabled"] and "gravity" not in ph \
            and list(cfg["physics"]["gravity"]) == [0.0, -10.0, 0.0]:
        cfg["physics"]["gravity"] = [0.0, 0.0, 0.0]
    return cfg


def endpoint_quaternion(d):
    """Shortest rotation taking +Y onto the unit direction d, as the reference's
    endpoint loader builds it (180 degrees about z for d = -Y)."""
    up_d = d[1]
    if up_d > 0.999999:
        return np.array([1.0, 0.0, 0.0, 0.0])
    if up_d < -0.999999:
        return np.array([0.0, 0.0, 0.0, 1.0])
    c = np.array([d[2], 0.0, -d[0]])  # (0, 1, 0) x d
    s = math.sqrt((1.0 + up_d) * 2.0)
    q = np.array([0.5 * s, c[0] / s, c[1] / s, c[2] / s])
    return q / np.linalg.norm(q)


def load_endpoints(path, cfg):
    """Rods from an endpoint file (App::loadInitialConfigCSV): '#key=value'
    metadata (rod_length, rod_diameter, rod_radius, pbc, box_size), an optional
    'x0,y0,z0,x1,y1,z1' header, rows of six numbers separated by commas or
    blanks, an optional seventh column with the diameter.  Returns a dict of
    per-rod arrays in the reference's frame and the metadata."""
    pop = cfg["populate"]
    dia = 0.05
    dens, rest, fric = 1000.0, 0.15, 0.6
    if pop["count"] > 0 or pop["density"] > 0:
        if pop["radius"] > 0:
            dia = 2.0 * pop["radius"]
        if pop["density"] > 0:
            dens = pop["density"]
    if cfg["bodies"]:
        b = cfg["bodies"][0]
        dia, dens, rest, fric = b["diameter"], b["density"], b["restitution"], b["friction"]
    meta = {}
    ends, dias = [], []
    schema = False
    ignore_dia = False
    for raw in open(path):
        line = raw.strip()
        if not line:
            continue
        low = " ".join(line.replace(",", " ").replace("\t", " ").lower().split())
        if "x0 y0 z0 x1 y1 z1" in low:
            schema = True
            ignore_dia = "attempts" in low
            continue
        if line.startswith("#"):
            if "Rod configuration: x y z phi theta length" in line:
                raise NotImplementedError("centre/angle rod files are not supported; use endpoint files")
            if "=" in line:
                key, val = line[1:].split("=", 1)
                key, val = key.strip(), val.strip().split()[0] if val.strip() else ""
                try:
                    if key == "rod_diameter":
                        dia = float(val)
                    elif key == "rod_radius":
                        dia = 2.0 * float(val)
                    elif key in ("pbc", "box_size", "rod_length"):
                        meta[key] = val
                except ValueError:
                    pass
            continue
        toks = line.replace(",", " ").split()
        try:
            vals = [float(t) for t in toks]
        except ValueError:
            continue
        if len(vals) < 6:
            continue
        if pop["count"] > 0 and len(ends) >= pop["count"]:
            break
        d = dia
        if len(vals) >= 7 and not ignore_dia:
            v = vals[6]
            if schema:
                d = v
            elif not (abs(v - round(v)) < 1e-9 and v >= 1.0) and 0.0 < v < 0.5:
                d = v
        ends.append(vals[:6])
        dias.append(d)
    if not ends:
        raise ValueError(f"no rods in {path}")
    e = np.asarray(ends, np.float64)
    return dict(ends=e, diameter=np.asarray(dias), density=dens, restitution=rest, friction=fric, meta=meta)


class Simulator:
    """See the module docstring.  One GPU context per simulator."""

    def __init__(self, scene_path, init_csv_path="", quiet=True, init_state_path=""):
        self.cfg = load_scene(scene_path)
        self.quiet = quiet
        cfg = self.cfg
        if cfg["cylinder"].get("enabled"):
            raise NotImplementedError("cylinder tube walls are not implemented by this library")
        if cfg["randomForce"].get("enabled"):
            raise NotImplementedError("random forces are not implemented by this library")
        if cfg["fixed"]:
            raise NotImplementedError("fixed rods are not implemented by this library")
        if cfg["floor"]["enabled"] and list(cfg["floor"]["rot_quat"]) != [1.0, 0.0, 0.0, 0.0]:
            raise NotImplementedError("rotated floors are not implemented by this library")
        self.model = "nsc" if cfg["nsc"]["enabled"] else ("hm" if cfg["hm"]["enabled"] else
                                                          ("soft" if cfg["soft"]["enabled"] else None))
        if self.model is None:
            raise NotImplementedError("the scene selects no contact model this library implements "
                                      "(physics.nsc, physics.soft_contact or physics.hertz_mindlin)")
        self.dt = float(cfg["physics"]["dt"])
        self.ctx = L.Context(0)
        csv = init_csv_path or cfg["initCsvPath"]
        frame0 = 0
        loaded = False
        if init_state_path:  # the CLI's --init-state: highest priority (main.cpp:1630-1638)
            try:
                frame0 = self._rods_from_state(init_state_path)
                loaded = True
            except ValueError as e:
                print(f"{e}\n[init-state] Failed to load state from {init_state_path}", file=sys.stderr)
        elif csv:
            self._rods_from_endpoints(csv)
            loaded = True
        if not loaded:
            if cfg["populate"]["count"] > 0:
                self._populate()
            elif cfg["bodies"]:
                self._rods_from_bodies()
            else:
                raise ValueError("the scene has no rods (no populate.count, bodies or initCsvPath)")
        self._random_init()
        self.frame_index = frame0
        self.last_hit_count = 0
        self.last_island_count = 0
        self.last_residual = 0.0
        self.last_step_ms = 0.0
        self._reldisp = sceneio.RelDisp()
        self._load()
        if self.model == "nsc":
            self.ctx.keep_manifolds(True)  # for write_network (one copy of the manifolds per step)
            self.ctx.record_stage_energies(True)  # for write_profile (three KE reductions per step)
        elif self.model == "soft":
            self.ctx.keep_soft_forces(True)  # for write_network (6 doubles per contact per evaluation)
        self.last_ke = self.ctx.observables()[0]
        if not quiet:
            print(f"[simulator] {len(self.r)} rods, model {self.model}, dt {self.dt}", file=sys.stderr)

    # ---- scene construction (reference frame) --------------------------------
    def _rods_from_endpoints(self, path):
        E = load_endpoints(path, self.cfg)
        meta = E["meta"]
        if "pbc" in meta:
            self.cfg["periodic"]["enabled"] = meta["pbc"] in ("1", "true", "True")
        if "box_size" in meta:
            Lb = float(meta["box_size"])
            self.cfg["periodic"]["min"] = [-0.5 * Lb] * 3
            self.cfg["periodic"]["max"] = [0.5 * Lb] * 3
        e = E["ends"]
        p0, p1 = e[:, :3], e[:, 3:6]
        d = p1 - p0
        length = np.linalg.norm(d, axis=1)
        keep = length >= 1e-6
        p0, p1, d, length, dia = p0[keep], p1[keep], d[keep], length[keep], E["diameter"][keep]
        u = d / length[:, None]
        self.x = 0.5 * (p0 + p1)
        self.q = np.array([endpoint_quaternion(ui) for ui in u])
        self.r = 0.5 * dia
        self.h = 0.5 * length
        self._materials(E["density"], E["restitution"], E["friction"])

    def _rods_from_state(self, path):
        """Rods of the last frame of an init-state CSV (App::loadInitialStateCSV):
        length 0.5, radius 0.025, density 1000, restitution and friction 0.5,
        with radius and density taken from scene.populate when its count or
        density is set.  Returns the frame number."""
        frame, p, v, w, q = sceneio.parse_state_csv(path)
        pop = self.cfg["populate"]
        rad, dens = 0.025, 1000.0
        if pop["count"] > 0 or pop["density"] > 0:
            rad, dens = pop["radius"], pop["density"]
        if pop["shape"] != "capsule":
            raise NotImplementedError("spheres are not implemented by this library")
        n = len(p)
        self.x = p
        self.q = q / np.linalg.norm(q, axis=1, keepdims=True)
        self.r = np.full(n, float(rad))
        self.h = np.full(n, 0.25)
        self._materials(dens, 0.5, 0.5)
        self.v, self.w = v, w
        return frame

    def _body0(self):
        b = dict(_BODY)
        if self.cfg["bodies"]:
            b = self.cfg["bodies"][0]
        else:
            b["diameter"] = 2.0 * self.cfg["populate"]["radius"]
            b["density"] = self.cfg["populate"]["density"]
        return b

    def _populate(self):
        cfg, pop = self.cfg, self.cfg["populate"]
        if pop["shape"] == "sphere":
            raise NotImplementedError("spheres are not implemented by this library")
        b = self._body0()
        n = int(pop["count"])
        r, h = 0.5 * b["diameter"], 0.5 * b["length"]
        per = cfg["periodic"]
        lo, hi = np.array(per["min"], float), np.array(per["max"], float)
        if pop["mode"] == "uniform" and not per["enabled"]:
            lo, hi = np.full(3, -10.0), np.full(3, 10.0)
        rng = np.random.default_rng(pop["seed"] or None)
        rr, hh = np.full(n, r), np.full(n, h)
        if pop["mode"] == "nonoverlap":
            # the reference's serial placement, reproduced bit for bit on the GPU
            # (FP32, its mt19937 stream and grid; main.cpp:1857-2140)
            seed = int(pop["seed"]) or int(rng.integers(1, 2**32))
            att = max(1000, int(pop["maxAttempts"]))
            c, _, q, _ = self.ctx.populate_nonoverlap(n, b["length"], b["diameter"], lo, hi, per["enabled"],
                                                      float(per["cellSize"]), seed, att)
            if len(c) < n:
                print(f"[populate] nonoverlap: failed to place rod {len(c)}/{n} after attempts= {att}",
                      file=sys.stderr)
            self.x = c.astype(np.float64)
            self.q32 = L.glm_normalize_quat_f32(q)  # what RigidBody::makeCapsule stores
            q = q.astype(np.float64)
            self.q = q / np.linalg.norm(q, axis=1, keepdims=True)
            rr, hh = rr[:len(c)], hh[:len(c)]
        elif pop["mode"] == "grid":
            sp = pop["spacingMul"] * b["diameter"]
            ext = hi - lo
            dims = np.maximum(1, (ext / sp).astype(int))
            cap = int(np.prod(dims))
            if cap < n:
                dims = np.maximum(1, (dims / np.cbrt(n / max(1.0, cap))).astype(int))
            cells = np.array([(i, j, k) for i in range(dims[0]) for j in range(dims[1]) for k in range(dims[2])])[:n]
            self.x = lo + (cells + 0.5) * sp + (rng.random((len(cells), 3)) - 0.5) * 0.3 * sp
            ax = rng.random((len(cells), 3)) - 0.5
            ax /= np.linalg.norm(ax, axis=1, keepdims=True)
            ang = (rng.random(len(cells)) - 0.5) * math.pi
            self.q = np.column_stack([np.cos(ang / 2), ax * np.sin(ang / 2)[:, None]])
            rr, hh = rr[:len(cells)], hh[:len(cells)]
        else:  # uniform positions, uniform orientations, overlaps allowed
            self.x = lo + rng.random((n, 3)) * (hi - lo)
            qq = rng.normal(size=(n, 4))
            self.q = qq / np.linalg.norm(qq, axis=1, keepdims=True)
        self.r, self.h = rr, hh
        self._materials(b["density"], b["restitution"], b["friction"])

    def _rods_from_bodies(self):
        bs = self.cfg["bodies"]
        if any(b["shape"] == "sphere" for b in bs):
            raise NotImplementedError("spheres are not implemented by this library")
        self.x = np.array([b["pos"] for b in bs], float)
        q = np.array([b["rot_quat"] for b in bs], float)
        self.q = q / np.linalg.norm(q, axis=1, keepdims=True)
        self.r = np.array([0.5 * b["diameter"] for b in bs])
        self.h = np.array([0.5 * b["length"] for b in bs])
        self._materials(np.array([b["density"] for b in bs]), np.array([b["restitution"] for b in bs]),
                        np.array([b["friction"] for b in bs]))
        self.v = np.array([b["v_lin"] for b in bs], float)
        self.w = np.array([b["v_ang"] for b in bs], float)
        self.static = np.array([bool(b["is_static"]) for b in bs])

    def _materials(self, density, restitution, friction):
        n = len(self.r)
        self.density = np.broadcast_to(np.asarray(density, float), (n,)).copy()
        self.mass = self.density * math.pi * self.r ** 2 * (2.0 * self.h)
        self.rest = np.broadcast_to(np.asarray(restitution, float), (n,)).copy()
        self.fric = np.broadcast_to(np.asarray(friction, float), (n,)).copy()
        self.v = np.zeros((n, 3))
        self.w = np.zeros((n, 3))
        self.static = np.zeros(n, bool)

    def _random_init(self):
        """The reference's randomInit (App::resetScene, main.cpp:2213-2294):
        one std::mt19937(seed) over the rods in order, FP32, from the rods'
        FP32 quaternions and the FP32 mass and transverse inertia of
        RigidBody::makeCapsule (rigid_body.cpp:11-34)."""
        ri = self.cfg["randomInit"]
        if not ri["enabled"]:
            return
        n = len(self.r)
        q32 = getattr(self, "q32", None)
        if q32 is None or len(q32) != n:
            q32 = L.glm_normalize_quat_f32(self.q.astype(np.float32))
        f = np.float32
        r32, h32, d32 = self.r.astype(f), self.h.astype(f), self.density.astype(f)
        H = f(2.0) * h32
        vol = ((f(math.pi) * r32) * r32) * H
        m32 = np.maximum(f(1e-6), d32 * vol)
        ip32 = (m32 * ((f(3.0) * r32) * r32 + H * H)) / f(12.0)
        mode = L.RANDOM_INIT_MODES.get(ri["mode"], L.RANDOM_INIT_MODES["uniform"])
        kbt = f(ri["kBT"])
        kt = f(ri["kBTTrans"]) if ri["kBTTrans"] >= 0.0 else kbt
        kr = f(ri["kBTRot"]) if ri["kBTRot"] >= 0.0 else kbt
        seed = int(ri["seed"]) & 0xFFFFFFFF
        v, w = L.random_init(q32, mode, ri["vSigma"], ri["wSigma"], ri["wSpeed"], kt, kr, seed,
                             ri["projectParallelSpin"], m32, ip32)
        self.v = v.astype(np.float64)
        self.w = w.astype(np.float64)
        self.v[self.static] = 0.0
        self.w[self.static] = 0.0

    # ---- library frame --------------------------------------------------------
    @staticmethod
    def _box_lib(lo, hi):
        a, b = _R @ lo, _R @ hi
        return np.minimum(a, b), np.maximum(a, b)

    def _load(self):
        cfg = self.cfg
        ctx = self.ctx
        xl = self.x @ _R.T
        ql = _qmul(_QR, self.q)
        vl, wl = self.v @ _R.T, self.w @ _R.T
        mass = np.where(self.static, 0.0, self.mass)
        ctx.set_bodies(xl, ql, vl, wl, self.r, self.h, mass, self.rest)
        per = cfg["periodic"]
        fl = cfg["floor"]
        if per["enabled"]:
            lo, hi = self._box_lib(np.array(per["min"], float), np.array(per["max"], float))
            periodic = (1, 1, 1)
        else:  # free space: a box around the rods (bodies leaving it stay in its edge cells)
            m = 4.0 * float(np.max(self.h + self.r))
            lo, hi = xl.min(0) - m, xl.max(0) + m
            periodic = (0, 0, 0)
        floor = 0
        if fl["enabled"]:
            top = fl["pos"][1] + fl["half_extents"][1]
            lo = lo.copy()
            lo[2] = top
            floor = 1
            if per["enabled"]:
                periodic = (1, 1, 0)
        mu = {"nsc": cfg["nsc"]["mu"], "soft": cfg["soft"]["mu"], "hm": cfg["hm"]["friction_coeff"]}[self.model]
        ctx.set_systems([0, len(self.r)], [L.system_desc(lo, hi, periodic, floor, 0, mu)])
        ph = cfg["physics"]
        g = _R @ np.array(ph["gravity"], float)
        ctx.set_physics(L.physics(tuple(g), self.dt, ph["lin_damp"], ph["ang_damp"], ph["w_max"]))
        self.det = L.detect_params(0.0, 0.0, 1.0)
        n = cfg["nsc"]
        if self.model == "nsc":
            self.params = L.nsc_params(beta=n["beta"], cfm=n["cfm"], omega=n["omega"], tol=0.0,
                                       max_sweeps=int(n["velocity_iters"]), check_every=1, friction_proj=0,
                                       position_stabilization=int(bool(n["position_stabilization"])),
                                       warm_mode=1 if n["enable_warm_start"] else 0, stop_rule=0, warm_walls=0,
                                       position_iters=int(n["position_iters"]),
                                       position_psor_iters=int(n["position_psor_iters"]), slop=n["slop"])
        elif self.model == "hm":
            hm = cfg["hm"]
            self.params = L.hm_params(E=hm["youngs_modulus"], nu_poisson=hm["poisson_ratio"],
                                      restitution=hm["restitution_coeff"], mu_k=hm["friction_coeff"],
                                      mu_s=hm["friction_static_coeff"], v_c=hm["friction_transition_vel"],
                                      mu_roll=hm["rolling_friction_coeff"],
                                      enable_tangential=int(bool(hm["enable_tangential"])),
                                      enable_rolling=int(bool(hm["enable_rolling"])))
        else:
            s = cfg["soft"]
            fm = 1 if s["friction_cundall"] else (2 if s["friction_karnopp"] else 0)
            self.params = L.soft_params(delta=s["delta"], k_scaler=s["k_scaler"], damping=s["damping"], mu=s["mu"],
                                        mu_static=s["mu_static"], nu=s["nu"], kt=s["kt"],
                                        vel_deadband=s["vel_deadband"], enable_friction=int(bool(s["enable_friction"])),
                                        friction_model=fm)

    # ---- the reference's session API ------------------------------------------
    def step(self, steps=1):
        if steps < 0:
            raise ValueError("steps must be non-negative")
        for _ in range(int(steps)):
            t0 = time.perf_counter()
            if self.model == "nsc":
                st = self.ctx.step_nsc(self.det, self.params, 1)
                self.last_residual = float(st["residual_max"])
                self.last_hit_count = int(st["manifolds_total"])
                self.last_island_count = int(self.ctx.island_info()[0]) if st["manifolds_total"] > 0 else 0
            else:
                st = self.ctx.step_vv(self.det, 1 if self.model == "hm" else 0, self.params, 1)
                self.last_hit_count = int(st["contacts_total"] // 2)
            self.last_step_ms = 1e3 * (time.perf_counter() - t0)
            self.frame_index += 1
        self.last_ke = self.ctx.observables()[0]

    def diagnostics(self):
        return {"frame_index": self.frame_index, "dt": self.dt, "last_ke": self.last_ke,
                "last_hit_count": self.last_hit_count, "last_island_count": self.last_island_count,
                "rod_count": len(self.r)}

    def _state_ref(self):
        """Device state rotated back into the reference's frame."""
        s = self.ctx.download_state()
        q = _qmul(np.array([_QR[0], -_QR[1], 0.0, 0.0]), s["q"])
        return s["x"] @ _R, q, s["v"] @ _R, s["w"] @ _R, s["f"] @ _R, s["tau"] @ _R

    def rods(self):
        x, q, v, w, f, t = self._state_ref()
        u = _axis(q)
        out = []
        for i in range(len(self.r)):
            m = 0.0 if self.static[i] else float(self.mass[i])
            a = x[i] - u[i] * self.h[i]
            b = x[i] + u[i] * self.h[i]
            out.append({"position": x[i].tolist(), "orientation_wxyz": q[i].tolist(),
                        "linear_velocity": v[i].tolist(), "angular_velocity": w[i].tolist(),
                        "force": f[i].tolist(), "torque": t[i].tolist(), "mass": float(self.mass[i]),
                        "inverse_mass": (1.0 / m) if m > 0 else 0.0, "friction": float(self.fric[i]),
                        "restitution": float(self.rest[i]), "shape": "capsule", "radius": float(self.r[i]),
                        "half_height": float(self.h[i]), "endpoint_a": a.tolist(), "endpoint_b": b.tolist()})
        return out

    # ---- reference file formats (App logging / I/O) --------------------------
    def write_snapshot(self, stream):
        """Append one NDJSON snapshot line of the current frame
        (App::writeSnapshotLine)."""
        x, q, v, w, _, _ = self._state_ref()
        stream.write(sceneio.snapshot_line(self.frame_index, self.dt, x, q, v, w, self.r, self.h))

    def save_endpoints(self, path):
        """Write the rods as an endpoint file (App::saveInitialConfigCSV)."""
        x, q, _, _, _, _ = self._state_ref()
        per = self.cfg["periodic"]
        box = float(per["max"][0]) - float(per["min"][0])
        with open(path, "w") as f:
            f.write(sceneio.endpoints_text(x, q, self.r, self.h, per["enabled"], box))

    def write_per_rod(self, stream, header=False):
        """Append the per-rod rows of the current frame (App::logPerRodFrame),
        preceded by the log's header when ``header`` is set."""
        x, q, v, w, _, _ = self._state_ref()
        if header:
            stream.write(sceneio.per_rod_header(self.r[0], self.h[0]))
        stream.write(sceneio.per_rod_text(self.frame_index, x, v, w, q, self.mass, self.r, self.h))

    def write_network(self, stream, header=False):
        """Append the network rows of the last step (App::logNetworkFrame).
        Hard contact: one row per rod-rod manifold of the step's velocity solve
        with its impulses, the normal velocity before the solve and the normal
        velocity from the current body velocities.  Soft contact: one row per
        contact of the last force evaluation with its forces.  Hertz-Mindlin:
        the soft header and no rows, as the reference writes."""
        if self.model != "nsc":
            if header:
                stream.write(sceneio.NETWORK_SOFT_HEADER)
            if self.model != "soft":
                return
            raw = self.ctx.contacts_raw()
            f = self.ctx.soft_forces()
            if len(f) != len(raw):
                raise RuntimeError("soft forces and contact list out of step (%d against %d)" % (len(f), len(raw)))
            pa = np.stack([raw["px_a"], raw["py_a"], raw["pz_a"]], 1) @ _R.astype(np.float32)
            pb = np.stack([raw["px_b"], raw["py_b"], raw["pz_b"]], 1) @ _R.astype(np.float32)
            nrm = np.stack([raw["nx"], raw["ny"], raw["nz"]], 1).astype(np.float64) @ _R
            stream.write(sceneio.network_soft_text(self.frame_index, raw["a"], raw["b"], pa, pb, nrm, raw["dist"],
                                                   f[:, 0:3] @ _R, f[:, 3:6] @ _R))
            return
        rec, ab, lam = self.ctx.kept_manifolds()
        keep = ab[:, 1] >= 0
        rec, ab, lam = rec[keep], ab[keep], lam[keep]
        s = self.ctx.download_state()  # library frame
        a, b = ab[:, 0], ab[:, 1]
        n_l, ra_l, rb_l = rec[:, 0:3], rec[:, 9:12], rec[:, 12:15]
        va = s["v"][a] + np.cross(s["w"][a], ra_l)
        vb = s["v"][b] + np.cross(s["w"][b], rb_l)
        vn_post = np.sum(n_l * (vb - va), axis=1)
        point = (s["x"][a] + ra_l) @ _R  # back to the reference frame
        if header:
            stream.write(sceneio.NETWORK_HEADER)
        stream.write(sceneio.network_text(self.frame_index, a, b, point, n_l @ _R, rec[:, 38], lam, rec[:, 39],
                                          vn_post))

    def write_profile(self, stream, header=False, ent_pairs=0, ent_sum=0.0):
        """Append one row of the profile log for the last step
        (App::logCsvFrame): the hard-contact header with the kinetic energy
        after each stage, the impulse sums and the solve's residual, or the
        slim soft/Hertz-Mindlin header.  Stage timings are the reference's own
        CPU timers and are written as 0 (the library's per-kernel times come
        from its profiling counters); step_ms is the wall time of the last
        step; ent_pairs / ent_sum are the entanglement observer's values when
        the caller runs it."""
        per = self.cfg["periodic"]
        pbc = bool(per["enabled"])
        lo, hi = per["min"], per["max"]
        x, m = self._state_ref()[0], self.mass
        gyr = sceneio.gyration_sq(x, m, lo, hi, pbc)
        rel = self._reldisp(x, m, lo, hi, pbc)
        n = len(self.r)
        if self.model == "nsc":
            if header:
                stream.write(sceneio.PROFILE_HEADER_NSC)
            rec, ab, lam = self.ctx.kept_manifolds()
            impacts = int(np.sum(rec[:, 39] < -1e-6)) if len(rec) else 0
            ke = self.ctx.stage_energies()
            row = [self.frame_index, n] + [0.0] * 15 + [self.last_hit_count, impacts, self.last_island_count,
                                                        self.last_ke, ke[0], 0.0, ke[1], ke[2], ke[2], 0.0, gyr, rel,
                                                        float(np.sum(np.abs(lam[:, 0]))),
                                                        float(np.sum(np.abs(lam[:, 1]) + np.abs(lam[:, 2]))),
                                                        self.last_residual, int(ent_pairs), ent_sum, self.last_step_ms]
        else:
            if header:
                stream.write(sceneio.PROFILE_HEADER_SOFT)
            row = [self.frame_index, n] + [0.0] * 10 + [self.last_hit_count, -1, self.last_ke,
                                                        self.ctx.last_potential_energy(), gyr, rel, int(ent_pairs),
                                                        ent_sum, self.last_step_ms]
        stream.write(sceneio.profile_row(row))

    def close(self):
        if getattr(self, "ctx", None) is not None:
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
