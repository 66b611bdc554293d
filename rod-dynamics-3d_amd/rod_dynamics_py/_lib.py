"""ctypes binding of librdk (include/rdk.h).

The library is loaded from the package directory; there is no fallback: if
librdk.so is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "librdk.so")

RDK_OK, RDK_ERR_HIP, RDK_ERR_ARG, RDK_ERR_CAPACITY, RDK_ERR_STATE, RDK_ERR_NOCONV = range(6)
RDK_NSC_FREE, RDK_NSC_POS, RDK_VV_HALF1, RDK_VV_HALF2, RDK_CLEAR = range(5)
MF_FIELDS = 42  # RDK_MF_FIELDS
MLP_IN, MLP_OUT, MLP_H = 12, 3, 64  # RDK_MLP_IN, RDK_MLP_OUT, hidden width
TUNE_SMALL_M, TUNE_BROAD_BLOCK, TUNE_BROAD_STAGE, TUNE_CELL_BAND = 0, 1, 2, 3  # RDK_TUNE_SMALL_M, _BROAD_BLOCK, _BROAD_STAGE

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class SystemDesc(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int * 3),
                ("floor_wall", C.c_int), ("ceil_wall", C.c_int), ("mu", C.c_double)]


class Physics(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("dt", C.c_double), ("lin_damp", C.c_double),
                ("ang_damp", C.c_double), ("w_max", C.c_double)]


class DetectParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("chunk_len", C.c_double), ("cell_scale", C.c_double)]


class NscParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("cfm", C.c_double), ("omega", C.c_double), ("tol", C.c_double),
                ("max_sweeps", C.c_int), ("check_every", C.c_int), ("friction_proj", C.c_int),
                ("position_stabilization", C.c_int), ("warm_mode", C.c_int), ("stop_rule", C.c_int),
                ("warm_walls", C.c_int), ("stall_sweeps", C.c_int), ("position_iters", C.c_int),
                ("position_psor_iters", C.c_int), ("slop", C.c_double), ("stall_omega", C.c_double),
                ("hot_set", C.c_int)]


class SoftParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("k_scaler", C.c_double), ("damping", C.c_double), ("mu", C.c_double),
                ("mu_static", C.c_double), ("nu", C.c_double), ("kt", C.c_double), ("vel_deadband", C.c_double),
                ("enable_friction", C.c_int), ("friction_model", C.c_int)]


class HmParams(C.Structure):
    _fields_ = [("E", C.c_double), ("nu_poisson", C.c_double), ("restitution", C.c_double), ("mu_k", C.c_double),
                ("mu_s", C.c_double), ("v_c", C.c_double), ("mu_roll", C.c_double),
                ("enable_tangential", C.c_int), ("enable_rolling", C.c_int), ("law", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("manifolds", C.c_int), ("rod_contacts", C.c_int), ("wall_contacts", C.c_int), ("colors", C.c_int),
                ("sweeps", C.c_int), ("residual", C.c_double), ("residual_ref", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("contacts_total", C.c_int64), ("manifolds_total", C.c_int64),
                ("sweeps_total", C.c_int64), ("colors_total", C.c_int64), ("nonconverged_steps", C.c_int64),
                ("residual_max", C.c_double), ("ms_detect", C.c_double), ("ms_build", C.c_double),
                ("ms_color", C.c_double), ("ms_solve", C.c_double), ("ms_integrate", C.c_double)]


# name -> (restype, argtypes); must match include/rdk.h
_SIGS = {
    "rdk_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rdk_destroy": (C.c_int, [C.c_void_p]),
    "rdk_last_error": (C.c_char_p, [C.c_void_p]),
    "rdk_sync": (C.c_int, [C.c_void_p]),
    "rdk_version": (C.c_int, []),
    "rdk_set_bodies": (C.c_int, [C.c_void_p, C.c_int] + [_dp] * 8),
    "rdk_set_systems": (C.c_int, [C.c_void_p, C.c_int, _ip, C.POINTER(SystemDesc)]),
    "rdk_set_physics": (C.c_int, [C.c_void_p, C.POINTER(Physics)]),
    "rdk_upload_state": (C.c_int, [C.c_void_p] + [_dp] * 4),
    "rdk_upload_forces": (C.c_int, [C.c_void_p] + [_dp] * 2),
    "rdk_download_state": (C.c_int, [C.c_void_p] + [_dp] * 6),
    "rdk_detect": (C.c_int, [C.c_void_p, C.POINTER(DetectParams), _ip]),
    "rdk_get_contacts": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip] + [_dp] * 7 + [_ip]),
    "rdk_set_pair_capacity": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_detect_info": (C.c_int, [C.c_void_p, _ip, _ip, C.POINTER(C.c_int64), _ip]),
    "rdk_forces_soft": (C.c_int, [C.c_void_p, C.POINTER(SoftParams), _dp]),
    "rdk_forces_hm": (C.c_int, [C.c_void_p, C.POINTER(HmParams), _dp]),
    "rdk_history_size": (C.c_int, [C.c_void_p, _ip]),
    "rdk_nsc_build": (C.c_int, [C.c_void_p, C.POINTER(NscParams), _ip]),
    "rdk_nsc_get_manifolds": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip, _dp]),
    "rdk_nsc_set_lambda": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "rdk_nsc_solve": (C.c_int, [C.c_void_p, C.POINTER(NscParams), C.POINTER(SolveStats)]),
    "rdk_nsc_sweeps": (C.c_int, [C.c_void_p, C.POINTER(NscParams), C.c_int]),
    "rdk_nsc_get_coloring": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip]),
    "rdk_nsc_residual": (C.c_int, [C.c_void_p, C.POINTER(NscParams), _dp]),
    "rdk_set_solver_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_nsc_island_info": (C.c_int, [C.c_void_p, _ip, _ip]),
    "rdk_nsc_project": (C.c_int, [C.c_void_p, C.POINTER(DetectParams), C.POINTER(NscParams), _ip, _ip]),
    "rdk_nsc_get_solve_order": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "rdk_integrate": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_step_nsc": (C.c_int, [C.c_void_p, C.POINTER(DetectParams), C.POINTER(NscParams), C.c_int,
                               C.POINTER(StepStats)]),
    "rdk_step_vv": (C.c_int, [C.c_void_p, C.POINTER(DetectParams), C.c_int, C.c_void_p, C.c_int,
                              C.POINTER(StepStats)]),
    "rdk_observables": (C.c_int, [C.c_void_p, _dp, _dp]),
    "rdk_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_set_profile_mask": (C.c_int, [C.c_void_p, C.c_uint64]),
    "rdk_kernel_times": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64), _dp]),
    "rdk_kernel_name": (C.c_char_p, [C.c_int]),
    "rdk_num_kernels": (C.c_int, []),
    "rdk_work_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int]),
    "rdk_entanglement": (C.c_int, [C.c_void_p, C.c_double, _dp, C.POINTER(C.c_int64)]),
    "rdk_entanglement_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "rdk_get_contacts_raw": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int)]),
    "rdk_keep_manifolds": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_keep_soft_forces": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_record_stage_energies": (C.c_int, [C.c_void_p, C.c_int]),
    "rdk_get_stage_energies": (C.c_int, [C.c_void_p, _dp]),
    "rdk_last_potential_energy": (C.c_int, [C.c_void_p, _dp]),
    "rdk_get_soft_forces": (C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_int)]),
    "rdk_nsc_get_kept_manifolds": (C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_int), _dp,
                                             C.POINTER(C.c_int)]),
    "rdk_detect_capsule_pairs_raw": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_float)] * 4 +
                                     [C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p,
                                      C.POINTER(C.c_int)]),
    "rdk_populate_nonoverlap": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.POINTER(C.c_float),
                                          C.POINTER(C.c_float), C.c_int, C.c_float, C.c_uint32, C.c_int, C.c_int,
                                          C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "rdk_gen_rsa": (C.c_int, [C.c_int, _dp, _dp, C.POINTER(SystemDesc), C.c_double, C.c_uint64, C.c_int, _dp, _dp,
                              _ip]),
    "rdk_gen_thermal": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_uint64, _dp, _dp]),
    "rdk_affine_scale": (C.c_int, [C.c_void_p, C.c_double]),
    "rdk_sort_pairs_u32": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "rdk_sort_pair_keys_u64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]),
    "rdk_exclusive_scan_i32": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "rdk_set_step_graph": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "rdk_step_graph_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]),
    "rdk_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "rdk_random_init": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "rdk_mlp_num_params": (C.c_int, []),
    "rdk_mlp_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "rdk_mlp_set_params": (C.c_int, [C.c_void_p, _fp, C.c_int]),
    "rdk_mlp_get_params": (C.c_int, [C.c_void_p, _fp, C.c_int]),
    "rdk_mlp_features": (C.c_int, [C.c_void_p, C.c_int, _fp, _dp]),
    "rdk_mlp_infer": (C.c_int, [C.c_void_p, _fp, C.c_int, _fp]),
    "rdk_mlp_grad": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int, _fp, _fp]),
    "rdk_mlp_train": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64,
                                _fp]),
}

_lib = None


def lib():
    """Load librdk.so once; raises OSError if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"librdk.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class RdkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"librdk status {status}: {msg}")
        self.status = status


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float64"
    return a.ctypes.data_as(_dp)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float32"
    return a.ctypes.data_as(_fp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous int32"
    return a.ctypes.data_as(_ip)


def system_desc(lo, hi, periodic, floor_wall=False, ceil_wall=False, mu=0.0):
    d = SystemDesc()
    for k in range(3):
        d.lo[k] = float(lo[k])
        d.hi[k] = float(hi[k])
        d.periodic[k] = int(bool(periodic[k]))
    d.floor_wall = int(bool(floor_wall))
    d.ceil_wall = int(bool(ceil_wall))
    d.mu = float(mu)
    return d


def nsc_params(beta=0.0, cfm=0.0, omega=1.0, tol=1e-8, max_sweeps=10000, check_every=1, friction_proj=0,
               position_stabilization=0, warm_mode=1, stop_rule=1, warm_walls=0, stall_sweeps=0, position_iters=5,
               position_psor_iters=50, slop=1e-4, stall_omega=1.0, hot_set=0):
    p = NscParams()
    p.hot_set = hot_set
    p.position_iters, p.position_psor_iters, p.slop = position_iters, position_psor_iters, slop
    p.stall_omega = stall_omega
    p.warm_walls = warm_walls
    p.stall_sweeps = stall_sweeps
    p.beta, p.cfm, p.omega, p.tol = beta, cfm, omega, tol
    p.max_sweeps, p.check_every, p.friction_proj = max_sweeps, check_every, friction_proj
    p.position_stabilization, p.warm_mode, p.stop_rule = position_stabilization, warm_mode, stop_rule
    return p


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


def physics(gravity=(0.0, 0.0, 0.0), dt=1e-3, lin_damp=0.0, ang_damp=0.0, w_max=1e30):
    p = Physics()
    for k in range(3):
        p.gravity[k] = float(gravity[k])
    p.dt, p.lin_damp, p.ang_damp, p.w_max = dt, lin_damp, ang_damp, w_max
    return p


def detect_params(delta=0.0, chunk_len=0.0, cell_scale=1.0):
    p = DetectParams()
    p.delta, p.chunk_len, p.cell_scale = delta, chunk_len, cell_scale
    return p


# the reference plug-in's GpuContactRaw record (rdk_contact_raw)
CONTACT_RAW = np.dtype([("a", np.int32), ("b", np.int32), ("px_a", np.float32), ("py_a", np.float32),
                        ("pz_a", np.float32), ("px_b", np.float32), ("py_b", np.float32), ("pz_b", np.float32),
                        ("nx", np.float32), ("ny", np.float32), ("nz", np.float32), ("dist", np.float32),
                        ("surface_limit", np.float32), ("s", np.float32), ("t", np.float32),
                        ("shift_bx", np.float32), ("shift_by", np.float32), ("shift_bz", np.float32)])


class Context:
    """One device context (one HIP stream) owning the device copy of a packing."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        st = L.rdk_create(int(device), C.byref(h))
        if st != RDK_OK:
            raise RdkError(st, "rdk_create failed (no HIP device?)")
        self._h = h
        self._L = L
        self.n = 0
        self._nsys = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.rdk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow=()):
        if st != RDK_OK and st not in allow:
            raise RdkError(st, self._L.rdk_last_error(self._h).decode())
        return st

    def set_bodies(self, x, q, v, w, r, h, mass, rest=None):
        n = len(r)
        arr = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
               for a in (x, q, v, w, r, h, mass, rest)]
        self._check(self._L.rdk_set_bodies(self._h, n, *[dptr(a) for a in arr]))
        self.n = n

    def set_systems(self, offsets, descs):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        arr = (SystemDesc * len(descs))(*descs)
        self._check(self._L.rdk_set_systems(self._h, len(descs), iptr(off), arr))
        self._nsys = len(descs)

    def set_physics(self, p):
        self._check(self._L.rdk_set_physics(self._h, C.byref(p)))

    def upload_state(self, x=None, q=None, v=None, w=None):
        arr = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (x, q, v, w)]
        self._check(self._L.rdk_upload_state(self._h, *[dptr(a) for a in arr]))

    def upload_forces(self, f=None, tau=None):
        arr = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (f, tau)]
        self._check(self._L.rdk_upload_forces(self._h, *[dptr(a) for a in arr]))

    def download_state(self):
        n = self.n
        x, v, w, f, t = (np.zeros((n, 3)) for _ in range(5))
        q = np.zeros((n, 4))
        self._check(self._L.rdk_download_state(self._h, dptr(x), dptr(q), dptr(v), dptr(w), dptr(f), dptr(t)))
        return dict(x=x, q=q, v=v, w=w, f=f, tau=t)

    def detect(self, p):
        nc = C.c_int()
        self._check(self._L.rdk_detect(self._h, C.byref(p), C.byref(nc)))
        return nc.value

    def contacts(self, nc):
        a = np.zeros(nc, np.int32)
        b = np.zeros(nc, np.int32)
        ty = np.zeros(nc, np.int32)
        pa, pb, nn, sh = (np.zeros((nc, 3)) for _ in range(4))
        dist, s, t = (np.zeros(nc) for _ in range(3))
        self._check(self._L.rdk_get_contacts(self._h, nc, iptr(a), iptr(b), dptr(pa), dptr(pb), dptr(nn),
                                             dptr(dist), dptr(sh), dptr(s), dptr(t), iptr(ty)))
        return dict(a=a, b=b, pa=pa, pb=pb, n=nn, dist=dist, shift=sh, s=s, t=t, type=ty)

    def set_pair_capacity(self, cap):
        self._check(self._L.rdk_set_pair_capacity(self._h, int(cap)))

    def detect_info(self):
        a, b, r = C.c_int(), C.c_int(), C.c_int()
        c = C.c_int64()
        self._check(self._L.rdk_detect_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(r)))
        return dict(chunks=a.value, cells=b.value, candidates=c.value, regrows=r.value)

    def forces_soft(self, p):
        pe = C.c_double()
        self._check(self._L.rdk_forces_soft(self._h, C.byref(p), C.byref(pe)))
        return pe.value

    def forces_hm(self, p):
        pe = C.c_double()
        self._check(self._L.rdk_forces_hm(self._h, C.byref(p), C.byref(pe)))
        return pe.value

    def history_size(self):
        n = C.c_int()
        self._check(self._L.rdk_history_size(self._h, C.byref(n)))
        return n.value

    def nsc_build(self, p):
        nm = C.c_int()
        self._check(self._L.rdk_nsc_build(self._h, C.byref(p), C.byref(nm)))
        return nm.value

    def nsc_manifolds(self, nm):
        rec = np.zeros((nm, MF_FIELDS))
        ab = np.zeros((nm, 2), np.int32)
        lam = np.zeros((nm, 3))
        self._check(self._L.rdk_nsc_get_manifolds(self._h, nm, dptr(rec), iptr(ab), dptr(lam)))
        return rec, ab, lam

    def keep_manifolds(self, on=True):
        self._check(self._L.rdk_keep_manifolds(self._h, int(bool(on))))

    def record_stage_energies(self, on=True):
        self._check(self._L.rdk_record_stage_energies(self._h, int(bool(on))))

    def stage_energies(self):
        """Kinetic energy after the free-motion update, the velocity solve and
        the position update of the last fused hard-contact step."""
        out = np.zeros(3)
        self._check(self._L.rdk_get_stage_energies(self._h, dptr(out)))
        return out

    def last_potential_energy(self):
        out = np.zeros(1)
        self._check(self._L.rdk_last_potential_energy(self._h, dptr(out)))
        return float(out[0])

    def keep_soft_forces(self, on=True):
        self._check(self._L.rdk_keep_soft_forces(self._h, int(bool(on))))

    def soft_forces(self):
        """Per-contact terms of the last soft force evaluation, (n, 6) in
        contact-list order: normal force on b, friction force on a."""
        cnt = C.c_int()
        self._check(self._L.rdk_get_soft_forces(self._h, 0, None, C.byref(cnt)), allow=(RDK_ERR_CAPACITY,))
        m = cnt.value
        out = np.zeros((max(m, 1), 6))
        self._check(self._L.rdk_get_soft_forces(self._h, m, dptr(out), C.byref(cnt)))
        return out[:m]

    def kept_manifolds(self):
        """Velocity-solve manifolds of the last fused hard-contact step:
        (records, body pairs, impulses) as nsc_manifolds returns them."""
        cnt = C.c_int()
        self._check(self._L.rdk_nsc_get_kept_manifolds(self._h, 0, None, None, None, C.byref(cnt)),
                    allow=(RDK_ERR_CAPACITY,))
        m = cnt.value
        rec = np.zeros((max(m, 1), MF_FIELDS))
        ab = np.zeros((max(m, 1), 2), np.int32)
        lam = np.zeros((max(m, 1), 3))
        self._check(self._L.rdk_nsc_get_kept_manifolds(self._h, m, dptr(rec), iptr(ab), dptr(lam), C.byref(cnt)))
        return rec[:m], ab[:m], lam[:m]

    def nsc_set_lambda(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._check(self._L.rdk_nsc_set_lambda(self._h, lam.shape[0], dptr(lam)))

    def nsc_solve(self, p, allow_noconv=False):
        st = SolveStats()
        self._check(self._L.rdk_nsc_solve(self._h, C.byref(p), C.byref(st)),
                    allow=(RDK_ERR_NOCONV,) if allow_noconv else ())
        return {k: getattr(st, k) for k, _ in SolveStats._fields_}

    def nsc_sweeps(self, p, n):
        self._check(self._L.rdk_nsc_sweeps(self._h, C.byref(p), int(n)))

    def nsc_coloring(self, nm):
        col = np.zeros(nm, np.int32)
        order = np.zeros(nm, np.int32)
        nc = C.c_int()
        self._check(self._L.rdk_nsc_get_coloring(self._h, nm, iptr(col), iptr(order), C.byref(nc)))
        return col, order, nc.value

    def nsc_residual(self, p):
        r = C.c_double()
        self._check(self._L.rdk_nsc_residual(self._h, C.byref(p), C.byref(r)))
        return r.value

    def set_solver_mode(self, mode):
        self._check(self._L.rdk_set_solver_mode(self._h, int(mode)))

    def nsc_project(self, dp, p):
        passes, nfirst = C.c_int(), C.c_int()
        self._check(self._L.rdk_nsc_project(self._h, C.byref(dp), C.byref(p), C.byref(passes), C.byref(nfirst)))
        return passes.value, nfirst.value

    def island_info(self):
        n = C.c_int()
        cnt = np.zeros(4, np.int32)
        self._check(self._L.rdk_nsc_island_info(self._h, C.byref(n), iptr(cnt)))
        return n.value, cnt.tolist()

    def solve_order(self, nm):
        order = np.zeros(nm, np.int32)
        self._check(self._L.rdk_nsc_get_solve_order(self._h, nm, iptr(order)))
        return order

    def integrate(self, stage):
        self._check(self._L.rdk_integrate(self._h, int(stage)))

    def step_nsc(self, dp, np_, nsteps):
        st = StepStats()
        self._check(self._L.rdk_step_nsc(self._h, C.byref(dp), C.byref(np_), int(nsteps), C.byref(st)))
        return {k: getattr(st, k) for k, _ in StepStats._fields_}

    def step_vv(self, dp, model, params, nsteps):
        st = StepStats()
        self._check(self._L.rdk_step_vv(self._h, C.byref(dp), int(model), C.cast(C.byref(params), C.c_void_p),
                                        int(nsteps), C.byref(st)))
        return {k: getattr(st, k) for k, _ in StepStats._fields_}

    def observables(self):
        ke, phi = C.c_double(), C.c_double()
        self._check(self._L.rdk_observables(self._h, C.byref(ke), C.byref(phi)))
        return ke.value, phi.value

    def entanglement(self, cutoff=-1.0, nsys=None):
        """Per system: (sum of |linking number| over rod pairs, pair count);
        every pair for cutoff <= 0."""
        k = max(int(nsys or self._nsys), 1)
        sums = np.zeros(k)
        pairs = np.zeros(k, np.int64)
        self._check(self._L.rdk_entanglement(self._h, float(cutoff), dptr(sums),
                                             pairs.ctypes.data_as(C.POINTER(C.c_int64))))
        return sums, pairs

    def entanglement_noalloc(self, d_sums, d_pairs, cutoff=-1.0):
        """rdk_entanglement_noalloc: sums and pair counts written into device
        memory at d_sums (doubles) and d_pairs (uint64), one per system,
        asynchronously on the context's stream."""
        self._check(self._L.rdk_entanglement_noalloc(self._h, float(cutoff), C.c_void_p(d_sums), C.c_void_p(d_pairs)))

    def populate_nonoverlap(self, n, length, diameter, bmin, bmax, periodic, cell_size, seed, max_attempts,
                            window=0):
        """The reference's non-overlapping populate on the GPU: (centres, unit
        directions, quaternions (w,x,y,z), candidates drawn), FP32 arrays of the
        rods placed, in placement order."""
        fp = C.POINTER(C.c_float)
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        c = np.zeros((max(n, 1), 3), np.float32)
        u = np.zeros((max(n, 1), 3), np.float32)
        q = np.zeros((max(n, 1), 4), np.float32)
        placed = C.c_int()
        att = C.c_int64()
        self._check(self._L.rdk_populate_nonoverlap(
            self._h, int(n), float(length), float(diameter), lo.ctypes.data_as(fp), hi.ctypes.data_as(fp),
            int(bool(periodic)), float(cell_size), int(seed) & 0xFFFFFFFF, int(max_attempts), int(window),
            c.ctypes.data_as(fp), u.ctypes.data_as(fp), q.ctypes.data_as(fp), C.byref(placed), C.byref(att)))
        k = placed.value
        return c[:k], u[:k], q[:k], att.value

    def contacts_raw(self):
        """The current contact list as the plug-in's FP32 records."""
        cnt = C.c_int()
        self._check(self._L.rdk_get_contacts_raw(self._h, 0, None, C.byref(cnt)), allow=(RDK_ERR_CAPACITY,))
        out = np.zeros(max(cnt.value, 1), CONTACT_RAW)
        self._check(self._L.rdk_get_contacts_raw(self._h, cnt.value, out.ctypes.data_as(C.c_void_p),
                                                 C.byref(cnt)))
        return out[:cnt.value]

    def detect_capsule_pairs_raw(self, x, q, r, h, margin, pbc=False, box=(0.0, 0.0, 0.0), cap=None):
        """One call in place of the reference's cudaDetectCapsulePairsTwoPass:
        FP32 capsules in, FP32 GpuContactRaw-layout records out."""
        f = [np.ascontiguousarray(v, np.float32) for v in (x, q, r, h)]
        n = len(f[2])
        cnt = C.c_int()
        fp = C.POINTER(C.c_float)
        args = [self._h, n] + [v.ctypes.data_as(fp) for v in f] + [float(margin), int(bool(pbc))] + \
            [float(v) for v in box]
        if cap is None:
            st = self._L.rdk_detect_capsule_pairs_raw(*args, 0, None, C.byref(cnt))
            self._check(st, allow=(RDK_ERR_CAPACITY,))
            cap = cnt.value
        out = np.zeros(max(cap, 1), CONTACT_RAW)
        self._check(self._L.rdk_detect_capsule_pairs_raw(*args, int(cap), out.ctypes.data_as(C.c_void_p),
                                                         C.byref(cnt)))
        self.n = n
        self._nsys = 1
        return out[:cnt.value]

    def sync(self):
        self._check(self._L.rdk_sync(self._h))

    def set_profiling(self, on):
        self._check(self._L.rdk_set_profiling(self._h, int(bool(on))))

    def set_profile_mask(self, mask):
        self._check(self._L.rdk_set_profile_mask(self._h, C.c_uint64(mask)))

    def kernel_times(self):
        out = {}
        for k in range(self._L.rdk_num_kernels()):
            n = C.c_int64()
            ms = C.c_double()
            self._check(self._L.rdk_kernel_times(self._h, k, C.byref(n), C.byref(ms)))
            out[self._L.rdk_kernel_name(k).decode()] = (n.value, ms.value)
        return out

    def work_counters(self):
        arr = (C.c_int64 * 16)()
        self._check(self._L.rdk_work_counters(self._h, arr, 16))
        names = ["steps", "body_passes", "chunks", "candidates", "contacts", "manifolds", "sweep_contacts",
                 "resid_contacts", "integrate_bodies", "force_contacts", "narrow_calls", "psor_timed_updates",
                 "lds_manifold_sweeps", "hot_sweeps"]
        return {nm: arr[i] for i, nm in enumerate(names)}

    # ---- fixed-capacity / graph-captured hard-contact step ----
    def set_step_graph(self, mode, pair_capacity=0):
        """0: eager step; 1: fixed-capacity launches; 2: the fixed step captured in a graph."""
        self._check(self._L.rdk_set_step_graph(self._h, int(mode), int(pair_capacity)))

    def step_graph_info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self._L.rdk_step_graph_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"fixed_steps": a.value, "fallbacks": b.value, "captures": c.value}

    def set_tuning(self, what, value):
        self._check(self._L.rdk_set_tuning(self._h, int(what), int(value)))

    # ---- the library's device sort and scan on host arrays ----
    def sort_pairs_u32(self, keys, vals, key_bits, cap=0):
        k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
        v = np.ascontiguousarray(vals, dtype=np.int32).copy()
        assert k.shape == v.shape and k.ndim == 1
        self._check(self._L.rdk_sort_pairs_u32(self._h, k.size, int(cap), k.ctypes.data, v.ctypes.data,
                                               int(key_bits)))
        return k, v

    def sort_pair_keys_u64(self, keys, index_bits, cap=0):
        k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
        self._check(self._L.rdk_sort_pair_keys_u64(self._h, k.size, int(cap), k.ctypes.data, int(index_bits)))
        return k

    def exclusive_scan(self, x, cap=0):
        a = np.ascontiguousarray(x, dtype=np.int32)
        out = np.empty_like(a)
        self._check(self._L.rdk_exclusive_scan_i32(self._h, a.size, int(cap), a.ctypes.data, out.ctypes.data))
        return out

    def affine_scale(self, s):
        self._check(self._L.rdk_affine_scale(self._h, float(s)))

    # ---- learned warm start ----
    def mlp_init(self, seed=0):
        self._check(self._L.rdk_mlp_init(self._h, int(seed)))

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self._L.rdk_mlp_set_params(self._h, fptr(p), int(p.size)))

    def mlp_get_params(self):
        p = np.zeros(self._L.rdk_mlp_num_params(), np.float32)
        self._check(self._L.rdk_mlp_get_params(self._h, fptr(p), int(p.size)))
        return p

    def mlp_features(self, nm):
        X = np.zeros((max(nm, 1), MLP_IN), np.float32)
        ls = np.zeros(max(nm, 1), np.float64)
        self._check(self._L.rdk_mlp_features(self._h, int(nm), fptr(X), dptr(ls)))
        return X[:nm], ls[:nm]

    def mlp_infer(self, X):
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        Y = np.zeros((max(n, 1), MLP_OUT), np.float32)
        self._check(self._L.rdk_mlp_infer(self._h, fptr(X), int(n), fptr(Y)))
        return Y[:n]

    def mlp_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self._L.rdk_mlp_num_params(), np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self._L.rdk_mlp_grad(self._h, fptr(X), fptr(T), int(X.shape[0]), fptr(g), fptr(loss)))
        return g, float(loss[0])

    def mlp_train(self, X, T, epochs=10, batch=256, lr=1e-3, weight_decay=0.0, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self._L.rdk_mlp_train(self._h, fptr(X), fptr(T), int(X.shape[0]), int(epochs), int(batch),
                                          float(lr), float(weight_decay), int(seed), fptr(loss)))
        return float(loss[0])


def gen_rsa(r, h, desc, gap=0.0, seed=0, max_attempts=2000):
    r = np.ascontiguousarray(r, np.float64)
    h = np.ascontiguousarray(h, np.float64)
    n = len(r)
    x = np.zeros((n, 3))
    q = np.zeros((n, 4))
    placed = C.c_int()
    st = lib().rdk_gen_rsa(n, dptr(r), dptr(h), C.byref(desc), float(gap), C.c_uint64(seed), int(max_attempts),
                           dptr(x), dptr(q), C.byref(placed))
    if st != RDK_OK:
        raise RdkError(st, "rdk_gen_rsa failed")
    return x, q, placed.value


def gen_thermal(q, r, h, mass, kT, seed):
    q = np.ascontiguousarray(q, np.float64)
    n = len(r)
    v = np.zeros((n, 3))
    w = np.zeros((n, 3))
    st = lib().rdk_gen_thermal(n, dptr(q), dptr(np.ascontiguousarray(r, np.float64)),
                               dptr(np.ascontiguousarray(h, np.float64)), dptr(np.ascontiguousarray(mass, np.float64)),
                               float(kT), C.c_uint64(seed), dptr(v), dptr(w))
    if st != RDK_OK:
        raise RdkError(st, "rdk_gen_thermal failed")
    return v, w


RANDOM_INIT_MODES = {"gaussian": 0, "thermal": 1, "uniform": 2}  # RDK_RANDOM_INIT_*


def random_init(q32, mode, v_sigma=0.0, w_sigma=0.0, w_speed=0.0, kbt_trans=1.0, kbt_rot=1.0, seed=0, project=True,
                mass32=None, i_perp32=None, skip=None):
    """The reference's randomInit stream (rdk_random_init): q32 = FP32
    quaternions (w, x, y, z) as the reference's rods hold them; mode 0
    gaussian, 1 thermal, 2 uniform.  Returns (v, w) float32 n x 3; rods with
    skip set keep zeros."""
    q32 = np.ascontiguousarray(q32, np.float32)
    n = len(q32)
    m = np.ascontiguousarray(mass32 if mass32 is not None else np.ones(n), np.float32)
    ip = np.ascontiguousarray(i_perp32 if i_perp32 is not None else np.ones(n), np.float32)
    sk = None if skip is None else np.ascontiguousarray(skip, np.uint8)
    v = np.zeros((n, 3), np.float32)
    w = np.zeros((n, 3), np.float32)
    st = lib().rdk_random_init(n, q32.ctypes.data, m.ctypes.data, ip.ctypes.data, int(mode), float(v_sigma),
                               float(w_sigma), float(w_speed), float(kbt_trans), float(kbt_rot), int(seed) & 0xFFFFFFFF,
                               int(bool(project)), None if sk is None else sk.ctypes.data, v.ctypes.data,
                               w.ctypes.data)
    if st != RDK_OK:
        raise RdkError(st, "rdk_random_init failed")
    return v, w


def glm_normalize_quat_f32(q):
    """glm::normalize of FP32 quaternions (w, x, y, z), as RigidBody::makeCapsule
    applies it (rigid_body.cpp:17): length = sqrt((ww + xx) + (yy + zz)), then
    every component times 1 / length (identity for length <= 0)."""
    q = np.ascontiguousarray(q, np.float32)
    sq = q * q
    ln = np.sqrt((sq[:, 0] + sq[:, 1]) + (sq[:, 2] + sq[:, 3]))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.float32(1.0) / ln
        out = q * inv[:, None]
    bad = ~(ln > 0)
    out[bad] = np.array([1, 0, 0, 0], np.float32)
    return out.astype(np.float32)
