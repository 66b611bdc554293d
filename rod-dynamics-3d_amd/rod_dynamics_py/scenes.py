"""Synthetic packings for the benchmark configurations.

Each builder returns a ``Scene``: body arrays (interleaved per body, the
layout of include/rdk.h), the boxes, the physics constants and the contact
model parameters.  Seeds default to ``0x5EED0000 + config id``.

  C1  soft/HM thermal gas: N=1000, d=1, L=10, periodic cube at phi=0.10
  C2  hard-contact PSOR settling: N=1e5, L/d=20, periodic x,y, floor and
      ceiling walls, phi=0.05, gravity -z, mu=0.5, e=0, omega=1.3, dt=1e-3
  C3  dense polydisperse Hertz-Mindlin: N=1e6, L ~ lognormal(median 16,
      sigma_log 0.3) truncated to [4, 64], d=1, periodic cube generated at
      phi=0.05 and compressed affinely to phi=0.30 (build protocol), mu=0.3,
      E=1e7, nu=0.25, e=0.5, thermal T=1, dt=5e-5
  C5  learned-warm-start bed: as C2 with N=2e5 (settled by running C2)

Masses follow the reference's capsule factory (density * pi r^2 * 2h, the
cylinder part only; rigid_body.cpp:19-24) unless a config asks for unit mass.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

SEED0 = 0x5EED0000


def capsule_volume(r, h):
    return math.pi * r * r * (2.0 * h) + 4.0 / 3.0 * math.pi * r ** 3


@dataclass
class Scene:
    name: str
    x: np.ndarray
    q: np.ndarray
    v: np.ndarray
    w: np.ndarray
    r: np.ndarray
    h: np.ndarray
    mass: np.ndarray
    rest: np.ndarray
    offsets: np.ndarray
    descs: list
    phys: object
    model: str  # "nsc" | "soft" | "hm"
    params: object = None
    detect: object = None
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.r)

    def load(self, ctx):
        ctx.set_bodies(self.x, self.q, self.v, self.w, self.r, self.h, self.mass, self.rest)
        ctx.set_systems(self.offsets, self.descs)
        ctx.set_physics(self.phys)

    def box(self, s=0):
        d = self.descs[s]
        return (np.array(d.lo[:]), np.array(d.hi[:]), [int(p) for p in d.periodic],
                int(d.floor_wall), int(d.ceil_wall))


def box_side(n, r, h, phi):
    return (n * capsule_volume(r, h) / phi) ** (1.0 / 3.0)


def c2_settling(n=100_000, phi=0.05, seed=SEED0 + 2, mu=0.5, aspect=20.0, gap=0.0):
    """Hard-contact PSOR settling, scaled to n rods."""
    r0, h0 = 0.5, 0.5 * aspect
    side = box_side(n, r0, h0, phi)
    desc = L.system_desc((0, 0, 0), (side, side, side), (1, 1, 0), floor_wall=True, ceil_wall=True, mu=mu)
    r = np.full(n, r0)
    h = np.full(n, h0)
    x, q, placed = L.gen_rsa(r, h, desc, gap=gap, seed=seed, max_attempts=20000)
    if placed != n:
        raise RuntimeError(f"RSA placed {placed} of {n} rods")
    mass = 1.0 * math.pi * r * r * (2.0 * h)
    # Euclidean disc projection: the reference's sequential clamp does not
    # converge to a 1e-8 fixed-point residual on sliding beds.  Tangential
    # update friction_proj 3: a sticking contact's slip is solved with the 2x2
    # tangential block in one step (the scalar step omega/g_tb of
    # friction_proj 1 is short along the block's soft direction), a sliding
    # one takes the scalar step; same fixed points (measured on the GPU at 1e5
    # rods: 3.8x fewer sweeps per step, the whole protocol 1.5x faster).
    # Plane-wall contacts also start from the previous step's impulses
    # (the reference has no plane walls and starts its cylinder-wall contacts cold).
    # Islands whose residual no longer drops by 10 % between 1000, 2000, 4000,
    # ... sweeps stop and count as not converged (wedged rods: impulses drift
    # along a null direction at any omega) instead of running to max_sweeps.
    # No omega = 1 continuation here: on this bed it lets the stalled islands
    # creep on for thousands of sweeps (measured: 20 % lower throughput for
    # 60 instead of about 100 non-converged steps out of 4980).
    # Hot sets: once a large island has few manifolds above tol, only the
    # manifolds around them are swept (the slow stick/slip modes of a few wall
    # contacts that hold the last decades of the residual are local); before
    # that, selective rounds sweep the manifolds around those above tol/2
    # whenever they are at most half the island (hot_set 2; measured at 1e5
    # rods: 35.5M against 33.2M rod-steps/s, the LDS island kernel 9 % faster).
    nscp = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=4,
                        friction_proj=3, position_stabilization=0, warm_mode=1, stop_rule=1, warm_walls=1,
                        stall_sweeps=1000, stall_omega=0.0, hot_set=2)
    return Scene(name=f"C2_settling_N{n}", x=x, q=q, v=np.zeros((n, 3)), w=np.zeros((n, 3)), r=r, h=h, mass=mass,
                 rest=np.zeros(n), offsets=np.array([0, n], np.int32), descs=[desc],
                 phys=L.physics((0.0, 0.0, -1.0), 1e-3, 0.08, 0.12, 80.0), model="nsc", params=nscp,
                 detect=L.detect_params(0.0, 4.0, 1.0), extra=dict(phi=phi, side=side, mu=mu, seed=seed))


def c1_thermal(n=1000, phi=0.10, seed=SEED0 + 1, model="hm", aspect=10.0, kT=1.0):
    """Soft / Hertz-Mindlin thermal gas, unit mass."""
    r0, h0 = 0.5, 0.5 * aspect
    side = box_side(n, r0, h0, phi)
    desc = L.system_desc((0, 0, 0), (side, side, side), (1, 1, 1), mu=0.5)
    r = np.full(n, r0)
    h = np.full(n, h0)
    x, q, placed = L.gen_rsa(r, h, desc, gap=0.0, seed=seed, max_attempts=20000)
    if placed != n:
        raise RuntimeError(f"RSA placed {placed} of {n} rods")
    mass = np.ones(n)
    v, w = L.gen_thermal(q, r, h, mass, kT, seed + 17)
    if model == "hm":
        params = L.hm_params(E=1e7, nu_poisson=0.25, restitution=0.5, mu_k=0.5, mu_s=0.5, v_c=1e-3, mu_roll=0.0,
                             enable_tangential=1, enable_rolling=0, law=0)
        det = L.detect_params(0.0, 0.0, 1.0)
    else:
        params = L.soft_params(delta=0.005, k_scaler=1e4, damping=0.0, mu=0.5, mu_static=0.0, nu=1e-3,
                               enable_friction=1, friction_model=0)
        det = L.detect_params(0.005, 0.0, 1.0)
    return Scene(name=f"C1_{model}_N{n}", x=x, q=q, v=v, w=w, r=r, h=h, mass=mass, rest=np.full(n, 0.5),
                 offsets=np.array([0, n], np.int32), descs=[desc], phys=L.physics((0, 0, 0), 1e-4, 0.0, 0.0, 1e30),
                 model=model, params=params, detect=det, extra=dict(phi=phi, side=side, seed=seed))


def lognormal_lengths(n, rng, median=16.0, sigma=0.3, lo=4.0, hi=64.0):
    """Lengths L ~ lognormal(median, sigma_log), truncated to [lo, hi] by redrawing."""
    L = np.exp(np.log(median) + sigma * rng.standard_normal(n))
    bad = (L < lo) | (L > hi)
    while bad.any():
        L[bad] = np.exp(np.log(median) + sigma * rng.standard_normal(int(bad.sum())))
        bad = (L < lo) | (L > hi)
    return L


def c3_dense(n=1_000_000, phi0=0.05, phi1=0.30, seed=SEED0 + 3, mu=0.3, kT=1.0):
    """Dense polydisperse Hertz-Mindlin packing, generated
    by RSA at phi0; ``extra['side1']`` is the side at the target phi1 that the
    compression protocol (rod_dynamics_py.protocols.compress) reaches."""
    rng = np.random.default_rng(seed)
    Ls = lognormal_lengths(n, rng)
    r = np.full(n, 0.5)
    h = 0.5 * Ls
    vol = float(np.sum(math.pi * r * r * Ls + 4.0 / 3.0 * math.pi * r ** 3))
    side0 = (vol / phi0) ** (1.0 / 3.0)
    side1 = (vol / phi1) ** (1.0 / 3.0)
    desc = L.system_desc((0, 0, 0), (side0, side0, side0), (1, 1, 1), mu=mu)
    x, q, placed = L.gen_rsa(r, h, desc, gap=0.0, seed=seed, max_attempts=20000)
    if placed != n:
        raise RuntimeError(f"RSA placed {placed} of {n} rods")
    mass = math.pi * r * r * Ls  # density 1, cylinder part (rigid_body.cpp:19-24)
    v, w = L.gen_thermal(q, r, h, mass, kT, seed + 17)
    params = L.hm_params(E=1e7, nu_poisson=0.25, restitution=0.5, mu_k=mu, mu_s=mu, v_c=1e-3, mu_roll=0.0,
                         enable_tangential=1, enable_rolling=0, law=0)
    return Scene(name=f"C3_hm_N{n}", x=x, q=q, v=v, w=w, r=r, h=h, mass=mass, rest=np.full(n, 0.5),
                 offsets=np.array([0, n], np.int32), descs=[desc], phys=L.physics((0, 0, 0), 5e-5, 0.0, 0.0, 1e30),
                 model="hm", params=params, detect=L.detect_params(0.0, 4.0, 1.0),
                 extra=dict(phi0=phi0, phi1=phi1, side=side0, side1=side1, seed=seed, mean_length=float(Ls.mean()),
                            truncation=[4.0, 64.0]))


RSA_SATURATION = {4: 0.256, 8: 0.199, 16: 0.142, 32: 0.090, 64: 0.054}


def c4_batched(nsys=256, seed=SEED0 + 4, kT=1.0, restitution=0.5, threads=8, n_range=(500, 20000),
               aspects=(4, 8, 16, 32, 64)):
    """Many independent periodic packings in one context.

    Per system: N = floor(exp(U[ln 500, ln 20000])), L/d from {4, 8, 16, 32,
    64}, mu ~ U[0, 1], phi ~ U[min(0.05, cap), cap] with cap = min(0.25,
    0.85 phi_sat(L/d)); phi_sat is where random sequential addition of
    isotropic rods stops placing them (measured with this generator at 20 000
    attempts per rod: 0.256, 0.199, 0.142, 0.090, 0.054).  A draw whose box
    is shorter than twice the rod reach (L + d) is redrawn, so the minimum
    image stays unique; the number of redraws is recorded.  Gravity 0, thermal
    velocities (T = kT), restitution as given (the configuration does not fix
    one; a thermal gas needs some to stay a gas).  n_range and aspects narrow
    the draws for reduced test versions.
    """
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    draws, redraws = [], 0
    while len(draws) < nsys:
        n = int(math.floor(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        aspect = float(rng.choice(list(aspects)))
        mu = float(rng.uniform(0.0, 1.0))
        cap = min(0.25, 0.85 * RSA_SATURATION[int(aspect)])
        phi = float(rng.uniform(min(0.05, cap), cap))
        side = box_side(n, 0.5, 0.5 * aspect, phi)
        if side < 2.0 * (aspect + 1.0):
            redraws += 1
            continue
        draws.append((n, aspect, mu, phi, side))

    def make(k):
        n, aspect, mu, phi, side = draws[k]
        r = np.full(n, 0.5)
        h = np.full(n, 0.5 * aspect)
        desc = L.system_desc((0, 0, 0), (side, side, side), (1, 1, 1), mu=mu)
        x, q, placed = L.gen_rsa(r, h, desc, gap=0.0, seed=seed + 1000 + k, max_attempts=100000)
        if placed != n:
            raise RuntimeError(f"RSA placed {placed} of {n} rods in system {k} (L/d {aspect}, phi {phi:.3f})")
        mass = math.pi * 0.25 * aspect * np.ones(n)
        v, w = L.gen_thermal(q, r, h, mass, kT, seed + 2000 + k)
        return x, q, v, w, r, h, mass, desc

    with ThreadPoolExecutor(max_workers=threads) as ex:
        parts = list(ex.map(make, range(nsys)))
    cat = lambda i: np.concatenate([p[i] for p in parts])
    offsets = np.zeros(nsys + 1, np.int32)
    offsets[1:] = np.cumsum([len(p[4]) for p in parts])
    n_tot = int(offsets[-1])
    nscp = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=10000, check_every=4, friction_proj=1,
                        position_stabilization=0, warm_mode=1, stop_rule=1, warm_walls=0, stall_sweeps=250,
                        stall_omega=1.0, hot_set=1)
    return Scene(name=f"C4_batched_{nsys}x", x=cat(0), q=cat(1), v=cat(2), w=cat(3), r=cat(4), h=cat(5),
                 mass=cat(6), rest=np.full(n_tot, restitution), offsets=offsets, descs=[p[7] for p in parts],
                 phys=L.physics((0.0, 0.0, 0.0), 1e-3, 0.0, 0.0, 1e30), model="nsc", params=nscp,
                 detect=L.detect_params(0.0, 4.0, 1.0),
                 extra=dict(seed=seed, draws=draws, redraws=redraws, restitution=restitution, kT=kT,
                            side=max(d[4] for d in draws)))
