"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MF_FIELDS = 42

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class NscParams(C.Structure):
    _fields_ = [("mu", C.c_double), ("beta", C.c_double), ("cfm", C.c_double), ("omega", C.c_double),
                ("tol", C.c_double), ("max_sweeps", C.c_int), ("check_every", C.c_int), ("friction_proj", C.c_int),
                ("stop_rule", C.c_int), ("position_stabilization", C.c_int), ("warm_start", C.c_int),
                ("warm_walls", C.c_int), ("position_iters", C.c_int), ("position_psor_iters", C.c_int),
                ("slop", C.c_double)]


class SoftParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("k_scaler", C.c_double), ("damping", C.c_double), ("mu", C.c_double),
                ("mu_static", C.c_double), ("nu", C.c_double), ("kt", C.c_double), ("vel_deadband", C.c_double),
                ("enable_friction", C.c_int), ("friction_model", C.c_int)]


class HmParams(C.Structure):
    _fields_ = [("E", C.c_double), ("nu_poisson", C.c_double), ("restitution", C.c_double), ("mu_k", C.c_double),
                ("mu_s", C.c_double), ("v_c", C.c_double), ("mu_roll", C.c_double),
                ("enable_tangential", C.c_int), ("enable_rolling", C.c_int), ("law", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] + [_dp] * 8
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_box.argtypes = [C.c_void_p, _dp, _dp, _ip, C.c_int, C.c_int]
        L.orc_set_physics.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_set_state.argtypes = [C.c_void_p] + [_dp] * 4
        L.orc_set_forces.argtypes = [C.c_void_p] + [_dp] * 2
        L.orc_get_state.argtypes = [C.c_void_p] + [_dp] * 6
        L.orc_axis.argtypes = [_dp, _dp]
        L.orc_segseg.argtypes = [_dp] * 6
        L.orc_detect.restype = C.c_int
        L.orc_detect.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_get_contacts.argtypes = [C.c_void_p, C.c_int, _ip, _ip] + [_dp] * 7 + [_ip]
        L.orc_nsc_build.argtypes = [C.c_void_p, C.POINTER(NscParams)]
        L.orc_nsc_num_manifolds.argtypes = [C.c_void_p]
        L.orc_nsc_get_manifolds.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _dp]
        L.orc_nsc_set_manifolds.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _dp]
        L.orc_nsc_apply_warm.argtypes = [C.c_void_p]
        L.orc_nsc_sweeps.argtypes = [C.c_void_p, C.POINTER(NscParams), _ip, C.c_int]
        L.orc_nsc_solve.argtypes = [C.c_void_p, C.POINTER(NscParams), _ip, _dp, _ip]
        L.orc_nsc_solve_island.restype = C.c_int
        L.orc_nsc_solve_island.argtypes = [C.c_void_p, C.POINTER(NscParams), _ip, C.c_int, C.c_int, C.c_double,
                                           C.c_int, _dp, C.POINTER(C.c_longlong), _ip]
        L.orc_disc_project_metric.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_nsc_residual.restype = C.c_double
        L.orc_nsc_residual.argtypes = [C.c_void_p, C.POINTER(NscParams)]
        L.orc_nsc_store_warm.argtypes = [C.c_void_p]
        L.orc_nsc_project.argtypes = [C.c_void_p, C.POINTER(NscParams), _ip, _ip]
        L.orc_nsc_project.restype = C.c_int
        L.orc_nsc_step.argtypes = [C.c_void_p, C.POINTER(NscParams), C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.orc_integrate.argtypes = [C.c_void_p, C.c_int]
        L.orc_soft_forces.argtypes = [C.c_void_p, C.POINTER(SoftParams), _dp]
        L.orc_hm_forces.argtypes = [C.c_void_p, C.POINTER(HmParams), _dp]
        L.orc_hm_history_size.argtypes = [C.c_void_p]
        L.orc_vv_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [C.c_void_p]
        L.orc_packing_fraction.restype = C.c_double
        L.orc_packing_fraction.argtypes = [C.c_void_p]
        L.orc_linking_number.restype = C.c_double
        L.orc_linking_number.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_entanglement.restype = C.c_double
        L.orc_entanglement.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_longlong)]
        _fp = C.POINTER(C.c_float)
        L.orc_populate_nonoverlap.restype = C.c_int
        L.orc_populate_nonoverlap.argtypes = [C.c_int, C.c_float, C.c_float, _fp, _fp, C.c_int, C.c_float,
                                              C.c_uint, C.c_int, _fp, _fp, C.POINTER(C.c_longlong)]
        L.orc_mt_uniform_float.argtypes = [C.c_uint, C.c_int, _fp]
        L.orc_rsa_seg_d2.restype = C.c_float
        L.orc_rsa_seg_d2.argtypes = [_fp] * 4 + [C.c_int]
        L.orc_mt_raw.restype = C.c_uint
        L.orc_mt_raw.argtypes = [C.c_uint, C.c_longlong]
        L.orc_random_init.restype = None
        L.orc_random_init.argtypes = [C.c_int, _fp, _fp, _fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_uint, C.c_int, C.c_void_p, _fp, _fp]
        L.orc_mt_normal_float.argtypes = [C.c_uint, C.c_int, _fp]
        _lib = L
    return _lib


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def nsc_params(mu=0.5, beta=0.0, cfm=0.0, omega=1.0, tol=1e-8, max_sweeps=10000, check_every=1, friction_proj=0,
               stop_rule=1, position_stabilization=0, warm_start=1, warm_walls=0, position_iters=5,
               position_psor_iters=50, slop=1e-4):
    p = NscParams()
    p.position_iters, p.position_psor_iters, p.slop = position_iters, position_psor_iters, slop
    p.warm_walls = warm_walls
    p.mu, p.beta, p.cfm, p.omega, p.tol = mu, beta, cfm, omega, tol
    p.max_sweeps, p.check_every, p.friction_proj, p.stop_rule = max_sweeps, check_every, friction_proj, stop_rule
    p.position_stabilization, p.warm_start = position_stabilization, warm_start
    return p


def disc_project_metric(a, b, c, r, y):
    """argmin over |x| <= r of (x - y)^T G (x - y), G = [[a, b], [b, c]]
    (the oracle's friction_proj 2 projection; |y| > r > 0)."""
    out = np.array(y, dtype=np.float64)
    lib().orc_disc_project_metric(a, b, c, r, _d(out))
    return out


def soft_params(delta=0.005, k_scaler=1000.0, damping=0.0, mu=0.5, mu_static=0.0, nu=1e-5, kt=1000.0,
                vel_deadband=1e-3, enable_friction=1, friction_model=0):
    p = SoftParams()
    p.delta, p.k_scaler, p.damping, p.mu, p.mu_static, p.nu = delta, k_scaler, damping, mu, mu_static, nu
    p.kt, p.vel_deadband, p.enable_friction, p.friction_model = kt, vel_deadband, enable_friction, friction_model
    return p


def hm_params(E=1e7, nu_poisson=0.25, restitution=0.5, mu_k=0.5, mu_s=0.5, v_c=1e-3, mu_roll=0.0,
              enable_tangential=1, enable_rolling=0, law=0):
    p = HmParams()
    p.E, p.nu_poisson, p.restitution, p.mu_k, p.mu_s, p.v_c, p.mu_roll = E, nu_poisson, restitution, mu_k, mu_s, v_c, mu_roll
    p.enable_tangential, p.enable_rolling, p.law = enable_tangential, enable_rolling, law
    return p


class World:
    def __init__(self, x, q, v, w, r, h, mass, rest=None):
        L = lib()
        arr = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (x, q, v, w, r, h, mass, rest)]
        self.n = len(arr[4])
        if arr[2] is None:
            arr[2] = np.zeros((self.n, 3))
        if arr[3] is None:
            arr[3] = np.zeros((self.n, 3))
        self._h = L.orc_create(self.n, *[_d(a) for a in arr])
        self._L = L

    def __del__(self):
        try:
            if self._h:
                self._L.orc_destroy(self._h)
        except Exception:
            pass

    def set_box(self, lo, hi, periodic, floor_wall=0, ceil_wall=0):
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        per = np.ascontiguousarray(periodic, np.int32)
        self._L.orc_set_box(self._h, _d(lo), _d(hi), _i(per), int(floor_wall), int(ceil_wall))

    def set_physics(self, gravity=(0, 0, 0), dt=1e-3, lin_damp=0.0, ang_damp=0.0, w_max=1e30):
        g = np.ascontiguousarray(gravity, np.float64)
        self._L.orc_set_physics(self._h, _d(g), dt, lin_damp, ang_damp, w_max)

    def set_state(self, x=None, q=None, v=None, w=None):
        arr = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (x, q, v, w)]
        self._L.orc_set_state(self._h, *[_d(a) for a in arr])

    def set_forces(self, f=None, tau=None):
        arr = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (f, tau)]
        self._L.orc_set_forces(self._h, *[_d(a) for a in arr])

    def state(self):
        n = self.n
        x, v, w, f, t = (np.zeros((n, 3)) for _ in range(5))
        q = np.zeros((n, 4))
        self._L.orc_get_state(self._h, _d(x), _d(q), _d(v), _d(w), _d(f), _d(t))
        return dict(x=x, q=q, v=v, w=w, f=f, tau=t)

    def detect(self, method=0, delta=0.0, cell_size=-1.0):
        return self._L.orc_detect(self._h, int(method), float(delta), float(cell_size))

    def contacts(self, nc):
        a = np.zeros(nc, np.int32)
        b = np.zeros(nc, np.int32)
        ty = np.zeros(nc, np.int32)
        pa, pb, nn, sh = (np.zeros((nc, 3)) for _ in range(4))
        dist, s, t = (np.zeros(nc) for _ in range(3))
        self._L.orc_get_contacts(self._h, nc, _i(a), _i(b), _d(pa), _d(pb), _d(nn), _d(dist), _d(sh), _d(s), _d(t),
                                 _i(ty))
        return dict(a=a, b=b, pa=pa, pb=pb, n=nn, dist=dist, shift=sh, s=s, t=t, type=ty)

    def nsc_build(self, p):
        return self._L.orc_nsc_build(self._h, C.byref(p))

    def manifolds(self):
        m = self._L.orc_nsc_num_manifolds(self._h)
        rec = np.zeros((m, MF_FIELDS))
        ab = np.zeros((m, 2), np.int32)
        lam = np.zeros((m, 3))
        self._L.orc_nsc_get_manifolds(self._h, m, _d(rec), _i(ab), _d(lam))
        return rec, ab, lam

    def set_manifolds(self, rec, ab, lam=None):
        rec = np.ascontiguousarray(rec, np.float64)
        ab = np.ascontiguousarray(ab, np.int32)
        lam = None if lam is None else np.ascontiguousarray(lam, np.float64)
        self._L.orc_nsc_set_manifolds(self._h, rec.shape[0], _d(rec), _i(ab), _d(lam))

    def apply_warm(self):
        self._L.orc_nsc_apply_warm(self._h)

    def sweeps(self, p, n, order=None):
        order = None if order is None else np.ascontiguousarray(order, np.int32)
        self._L.orc_nsc_sweeps(self._h, C.byref(p), _i(order), int(n))

    def solve(self, p, order=None):
        order = None if order is None else np.ascontiguousarray(order, np.int32)
        r = C.c_double()
        s = C.c_int()
        self._L.orc_nsc_solve(self._h, C.byref(p), _i(order), C.byref(r), C.byref(s))
        return r.value, s.value

    def solve_island(self, p, ix, stall_sweeps=0, stall_omega=0.0, hot=0):
        """One island (manifold indices ix in solve order) solved as the
        library's island kernels do: (sweeps, residual, updates, hot sweeps)."""
        ix = np.ascontiguousarray(ix, np.int32)
        r = C.c_double()
        u = C.c_longlong()
        hs = C.c_int()
        sw = self._L.orc_nsc_solve_island(self._h, C.byref(p), _i(ix), len(ix), int(stall_sweeps),
                                          float(stall_omega), int(hot), C.byref(r), C.byref(u), C.byref(hs))
        return sw, r.value, u.value, hs.value

    def residual(self, p):
        return self._L.orc_nsc_residual(self._h, C.byref(p))

    def project(self, p, order=None):
        """Position projection over the current contact list; returns (passes, first-pass count)."""
        npos = C.c_int()
        o = None if order is None else np.ascontiguousarray(order, np.int32)
        passes = self._L.orc_nsc_project(self._h, C.byref(p), None if o is None else o.ctypes.data_as(_ip),
                                         C.byref(npos))
        return passes, npos.value

    def store_warm(self):
        self._L.orc_nsc_store_warm(self._h)

    def nsc_step(self, p, nsteps, method=1, cell_size=-1.0):
        s = C.c_double()
        r = C.c_double()
        self._L.orc_nsc_step(self._h, C.byref(p), int(nsteps), int(method), float(cell_size), C.byref(s),
                             C.byref(r))
        return s.value, r.value

    def integrate(self, stage):
        self._L.orc_integrate(self._h, int(stage))

    def soft_forces(self, p):
        pe = C.c_double()
        self._L.orc_soft_forces(self._h, C.byref(p), C.byref(pe))
        return pe.value

    def hm_forces(self, p):
        pe = C.c_double()
        self._L.orc_hm_forces(self._h, C.byref(p), C.byref(pe))
        return pe.value

    def hm_history_size(self):
        return self._L.orc_hm_history_size(self._h)

    def vv_step(self, model, p, nsteps, method=1, cell_size=-1.0):
        self._L.orc_vv_step(self._h, int(model), C.cast(C.byref(p), C.c_void_p), int(nsteps), int(method),
                            float(cell_size))

    def kinetic_energy(self):
        return self._L.orc_kinetic_energy(self._h)

    def packing_fraction(self):
        return self._L.orc_packing_fraction(self._h)

    def entanglement(self, cutoff=-1.0):
        """(sum of |Lk| over rod pairs, pair count); all pairs for cutoff <= 0."""
        n = C.c_longlong()
        s = self._L.orc_entanglement(self._h, float(cutoff), C.byref(n))
        return s, n.value


def linking_number(a0, a1, b0, b1):
    """Gauss linking number of the segments a0a1 and b0b1."""
    v = [np.ascontiguousarray(x, np.float64) for x in (a0, a1, b0, b1)]
    return lib().orc_linking_number(*[_d(x) for x in v])


def populate_nonoverlap(n, length, diameter, bmin, bmax, pbc, cell_size, seed, max_attempts):
    """The reference's serial non-overlapping populate in FP32: (placed,
    centres (placed, 3), directions (placed, 3), candidates drawn)."""
    fp = C.POINTER(C.c_float)
    lo = np.ascontiguousarray(bmin, np.float32)
    hi = np.ascontiguousarray(bmax, np.float32)
    c = np.zeros((n, 3), np.float32)
    u = np.zeros((n, 3), np.float32)
    att = C.c_longlong()
    k = lib().orc_populate_nonoverlap(int(n), float(length), float(diameter), lo.ctypes.data_as(fp),
                                      hi.ctypes.data_as(fp), int(pbc), float(cell_size), int(seed) & 0xFFFFFFFF,
                                      int(max_attempts), c.ctypes.data_as(fp), u.ctypes.data_as(fp), C.byref(att))
    return k, c[:k], u[:k], att.value


def rsa_seg_d2(p0, p1, q0, q1, pbc_rule):
    fp = C.POINTER(C.c_float)
    v = [np.ascontiguousarray(a, np.float32) for a in (p0, p1, q0, q1)]
    return float(lib().orc_rsa_seg_d2(*[a.ctypes.data_as(fp) for a in v], int(pbc_rule)))


def mt_uniform_float(seed, k):
    out = np.zeros(k, np.float32)
    lib().orc_mt_uniform_float(int(seed) & 0xFFFFFFFF, int(k), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def mt_raw(seed, k):
    return lib().orc_mt_raw(int(seed) & 0xFFFFFFFF, int(k))


def axis(q):
    q = np.ascontiguousarray(q, np.float64)
    u = np.zeros(3)
    lib().orc_axis(_d(q), _d(u))
    return u


def segseg(a1, a2, b1, b2):
    arr = [np.ascontiguousarray(a, np.float64) for a in (a1, a2, b1, b2)]
    s = np.zeros(1)
    t = np.zeros(1)
    lib().orc_segseg(*[_d(a) for a in arr], _d(s), _d(t))
    return s[0], t[0]


def world_from_scene(sc):
    """Oracle world with the bodies, box and physics of a rod_dynamics_py Scene (single system)."""
    W = World(sc.x, sc.q, sc.v, sc.w, sc.r, sc.h, sc.mass, sc.rest)
    lo, hi, per, fw, cw = sc.box(0)
    W.set_box(lo, hi, per, fw, cw)
    p = sc.phys
    W.set_physics(tuple(p.gravity[:]), p.dt, p.lin_damp, p.ang_damp, p.w_max)
    return W


def random_init(q, mode, v_sigma=0.0, w_sigma=0.0, w_speed=0.0, kbt_trans=1.0, kbt_rot=1.0, seed=1, project=True,
                mass=None, i_perp=None, skip=None):
    """The reference's randomInit stream (main.cpp:2213-2294) for FP32
    quaternions q (n x 4, w x y z); mode 0 gaussian, 1 thermal, 2 uniform.
    Returns (v, w) as float32 n x 3 (rods with skip set keep zeros)."""
    import numpy as np

    q = np.ascontiguousarray(q, np.float32)
    n = len(q)
    m = np.ascontiguousarray(mass if mass is not None else np.ones(n), np.float32)
    ip = np.ascontiguousarray(i_perp if i_perp is not None else np.ones(n), np.float32)
    sk = None if skip is None else np.ascontiguousarray(skip, np.uint8)
    v = np.zeros((n, 3), np.float32)
    w = np.zeros((n, 3), np.float32)
    fp = C.POINTER(C.c_float)
    lib().orc_random_init(n, q.ctypes.data_as(fp), m.ctypes.data_as(fp), ip.ctypes.data_as(fp), int(mode),
                          float(v_sigma), float(w_sigma), float(w_speed), float(kbt_trans), float(kbt_rot), int(seed),
                          int(bool(project)), None if sk is None else sk.ctypes.data, v.ctypes.data_as(fp),
                          w.ctypes.data_as(fp))
    return v, w


def mt_normal_float(seed, k):
    """The first k outputs of std::normal_distribution<float>(0, 1) on std::mt19937(seed)."""
    import numpy as np

    out = np.zeros(k, np.float32)
    lib().orc_mt_normal_float(int(seed), int(k), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out
