"""GPU NSC parity: manifolds against the oracle, colouring validity, a
coloured sweep equal bit for bit to the serial reference sweep taken in the
same (colour) order, and the converged solve checked by the oracle's own
residual."""
import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu


def overlapping_bed(n=2000, side=22.0, seed=21, walls=True, mu=0.5, rest=0.3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    if walls:
        x[: n // 10, 2] = rng.uniform(0.0, 1.0, n // 10)  # some rods through the floor
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.5, n)
    h = rng.uniform(0.5, 2.5, n)
    mass = rng.uniform(0.5, 2.0, n)
    v = rng.normal(size=(n, 3))
    w = rng.normal(size=(n, 3))
    periodic = (1, 1, 0) if walls else (1, 1, 1)
    return dict(x=x, q=q, v=v, w=w, r=r, h=h, mass=mass, rest=np.full(n, rest), side=side, periodic=periodic,
                walls=walls, mu=mu)


def gpu_load(ctx, B, dt=1e-3):
    ctx.set_bodies(B["x"], B["q"], B["v"], B["w"], B["r"], B["h"], B["mass"], B["rest"])
    ctx.set_systems([0, len(B["r"])], [L.system_desc([0] * 3, [B["side"]] * 3, B["periodic"], B["walls"], B["walls"],
                                                     B["mu"])])
    ctx.set_physics(L.physics((0, 0, -1), dt, 0, 0, 1e30))


def oracle_load(B, dt=1e-3):
    W = O.World(B["x"], B["q"], B["v"], B["w"], B["r"], B["h"], B["mass"], B["rest"])
    W.set_box([0] * 3, [B["side"]] * 3, B["periodic"], B["walls"], B["walls"])
    W.set_physics((0, 0, -1), dt)
    return W


def test_manifolds_match_oracle(gpu_ctx_factory):
    B = overlapping_bed()
    print("manifolds n=2000 walls", flush=True)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.2, cfm=1e-3, omega=1.0, warm_mode=0)
    nc = ctx.detect(L.detect_params(0.0, 0.0, 1.0))
    nm = ctx.nsc_build(p)
    W = oracle_load(B)
    assert W.detect(0, 0.0) == nc
    op = O.nsc_params(mu=B["mu"], beta=0.2, cfm=1e-3, warm_start=0)
    assert W.nsc_build(op) == nm
    rec_g, ab_g, lam_g = ctx.nsc_manifolds(nm)
    rec_o, ab_o, lam_o = W.manifolds()
    assert nm > nc > 100  # walls present
    assert np.array_equal(ab_g, ab_o)
    # geometry fields are computed with identical operation order: bitwise
    assert np.array_equal(rec_g[:, 0:15], rec_o[:, 0:15])
    assert np.array_equal(rec_g[:, 38], rec_o[:, 38])
    # inertia-weighted fields: closed-form vs matrix inertia, FP64 rounding only
    scale = np.abs(rec_o).max(axis=0) + 1e-300
    err = np.abs(rec_g - rec_o) / scale
    assert err.max() < 1e-13, err.max(axis=0)


def test_coloring_is_proper(gpu_ctx_factory):
    B = overlapping_bed(seed=5)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    col, order, ncol = ctx.nsc_coloring(nm)
    _, ab, _ = ctx.nsc_manifolds(nm)
    assert sorted(order.tolist()) == list(range(nm))
    assert np.all(np.diff(col[order]) >= 0)
    for c in range(ncol):
        idx = np.nonzero(col == c)[0]
        bodies = np.concatenate([ab[idx, 0], ab[idx, 1][ab[idx, 1] >= 0]])
        assert len(np.unique(bodies)) == len(bodies), f"colour {c} shares a body"
    # within a colour the natural order is kept
    for c in range(ncol):
        seg = order[col[order] == c]
        assert np.all(np.diff(seg) > 0)
    # degree bound: colours <= 2 * max body degree
    deg = np.bincount(np.concatenate([ab[:, 0], ab[:, 1][ab[:, 1] >= 0]]))
    assert ncol <= 2 * deg.max()


@pytest.mark.parametrize("proj", [0, 1, 2, 3])
def test_colored_sweeps_bitwise_equal_serial_sweeps_in_color_order(gpu_ctx_factory, proj):
    B = overlapping_bed(seed=8)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.2, cfm=1e-4, omega=1.3, friction_proj=proj, warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    col, order, ncol = ctx.nsc_coloring(nm)
    s0 = ctx.download_state()
    ctx.nsc_sweeps(p, 7)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam0)
    op = O.nsc_params(mu=B["mu"], beta=0.2, cfm=1e-4, omega=1.3, friction_proj=proj)
    W.sweeps(op, 7, order)
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g, lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])
    assert np.abs(lam_g).max() > 0


@pytest.mark.parametrize("proj", [1, 2, 3])
def test_solve_reaches_residual_checked_by_oracle(gpu_ctx_factory, proj):
    B = overlapping_bed(seed=31, mu=0.4)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=20000, check_every=4, friction_proj=proj,
                     warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, _ = ctx.nsc_manifolds(nm)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p)
    assert st["residual"] <= 1e-8 and st["sweeps"] > 0
    _, _, lam = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    op = O.nsc_params(mu=0.4, beta=0.0, cfm=0.0, omega=1.3, friction_proj=proj)
    W = oracle_load(B)
    # velocities after the solve are the pre-solve velocities plus M^-1 J^T lambda
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam)
    W.apply_warm()
    so = W.state()
    assert np.allclose(so["v"], s1["v"], rtol=0, atol=1e-11)
    assert np.allclose(so["w"], s1["w"], rtol=0, atol=1e-11)
    # the oracle's residual of the GPU solution
    W.set_state(v=s1["v"], w=s1["w"])
    assert W.residual(op) <= 1e-8 * (1 + 1e-6)
    # admissibility: lambda_n >= 0 and the friction bound
    assert lam[:, 0].min() >= 0.0
    tang = np.hypot(lam[:, 1], lam[:, 2]) if proj >= 1 else np.maximum(np.abs(lam[:, 1]), np.abs(lam[:, 2]))
    assert np.all(tang <= 0.4 * lam[:, 0] * (1 + 1e-12) + 1e-15)


def test_frictionless_unique_solution_matches_serial_oracle(gpu_ctx_factory):
    # mu = 0 with positive-definite Delassus: the LCP solution is unique, so the
    # coloured and the serial sorted-order Gauss-Seidel must agree
    B = overlapping_bed(n=600, side=16.0, seed=44, walls=False, mu=0.0)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=1e-3, omega=1.0, tol=1e-12, max_sweeps=100000, check_every=8, warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    assert nm > 50
    st = ctx.nsc_solve(p)
    assert st["residual"] <= 1e-12
    _, _, lam_g = ctx.nsc_manifolds(nm)
    W = oracle_load(B)
    W.detect(0, 0.0)
    op = O.nsc_params(mu=0.0, beta=0.0, cfm=1e-3, omega=1.0, tol=1e-12, max_sweeps=100000, warm_start=0)
    W.nsc_build(op)
    r, sw = W.solve(op)
    assert r <= 1e-12
    _, _, lam_o = W.manifolds()
    assert np.allclose(lam_g, lam_o, rtol=1e-7, atol=1e-10)


def clusters(nclust=40, seed=77, mu=0.5):
    """Separate clusters of overlapping rods (islands of 2..40 rods) above a floor."""
    rng = np.random.default_rng(seed)
    xs, qs = [], []
    side = 12.0 * int(np.ceil(np.sqrt(nclust)))
    for c in range(nclust):
        k = int(rng.integers(2, 41))
        cx, cy = 12.0 * (c % int(np.sqrt(nclust) + 1)) + 6, 12.0 * (c // int(np.sqrt(nclust) + 1)) + 6
        x = np.stack([rng.uniform(cx - 2, cx + 2, k), rng.uniform(cy - 2, cy + 2, k), rng.uniform(0.2, 3.0, k)], 1)
        q = rng.normal(size=(k, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        xs.append(x)
        qs.append(q)
    x = np.concatenate(xs)
    q = np.concatenate(qs)
    n = len(x)
    side = max(side, x[:, 0].max() + 10, x[:, 1].max() + 10)
    return dict(x=x, q=q, v=rng.normal(size=(n, 3)), w=rng.normal(size=(n, 3)), r=np.full(n, 0.4),
                h=rng.uniform(0.8, 2.0, n), mass=rng.uniform(0.5, 2.0, n), rest=np.full(n, 0.0), side=side,
                periodic=(1, 1, 0), walls=True, mu=mu)


def big_clusters(nclust=3, k=160, seed=91, mu=0.5):
    """A few dense clusters of many rods: islands of several hundred manifolds."""
    rng = np.random.default_rng(seed)
    xs, qs = [], []
    for c in range(nclust):
        cx = 30.0 * c + 15.0
        xs.append(np.stack([rng.uniform(cx - 4, cx + 4, k), rng.uniform(11, 19, k), rng.uniform(0.2, 4.0, k)], 1))
        q = rng.normal(size=(k, 4))
        qs.append(q / np.linalg.norm(q, axis=1, keepdims=True))
    x, q = np.concatenate(xs), np.concatenate(qs)
    n = len(x)
    return dict(x=x, q=q, v=rng.normal(size=(n, 3)), w=rng.normal(size=(n, 3)), r=np.full(n, 0.4),
                h=rng.uniform(1.0, 3.0, n), mass=rng.uniform(0.5, 2.0, n), rest=np.full(n, 0.0),
                side=30.0 * nclust, periodic=(1, 1, 0), walls=True, mu=mu)


def log_pile(ny=120, nz=30, seed=97, mu=0.5):
    """Rods along x stacked on a (y, z) lattice with spacing 0.78 < 2r = 0.8
    and small random tilts: every rod overlaps its four lattice neighbours, so
    the pile is one island of ny * nz bodies (more than the LDS island class
    holds) and about 2 ny nz manifolds, plus floor contacts of the bottom layer."""
    rng = np.random.default_rng(seed)
    jy, jz = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    n = ny * nz
    x = np.stack([30.0 + rng.uniform(-0.5, 0.5, n), 5.0 + 0.78 * jy.ravel(), 0.38 + 0.78 * jz.ravel()], 1)
    # axis +Y rotated onto +X (quaternion about z by -90 degrees), then tilted slightly
    base = np.array([np.sqrt(0.5), 0.0, 0.0, -np.sqrt(0.5)])
    tilt = rng.normal(size=(n, 3)) * 0.02
    ang = np.linalg.norm(tilt, axis=1)
    dq = np.column_stack([np.cos(ang / 2), tilt * (np.sin(ang / 2) / ang)[:, None]])
    w0, v0 = dq[:, :1], dq[:, 1:]
    bw, bv = base[0], base[1:]
    q = np.column_stack([w0[:, 0] * bw - v0 @ bv, w0 * bv + bw * v0 + np.cross(v0, bv)])
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    side = 0.78 * ny + 12.0
    return dict(x=x, q=q, v=0.1 * rng.normal(size=(n, 3)), w=0.1 * rng.normal(size=(n, 3)), r=np.full(n, 0.4),
                h=np.full(n, 2.0), mass=np.full(n, 1.0), rest=np.full(n, 0.0), side=side, periodic=(1, 1, 0),
                walls=True, mu=mu)


def pile_and_clusters():
    """A huge island (log pile) next to clusters of every smaller class, so the
    huge-island solve runs beside the island kernels on the side stream."""
    P = log_pile(ny=90, nz=40, seed=99)
    C = clusters(nclust=30, seed=78)
    C["x"] = C["x"] + np.array([40.0, 0.0, 0.0])
    out = {k: np.concatenate([P[k], C[k]]) for k in ("x", "q", "v", "w", "r", "h", "mass", "rest")}
    out.update(side=max(P["side"], C["x"][:, 0].max() + 10.0, C["x"][:, 1].max() + 10.0), periodic=(1, 1, 0),
               walls=True, mu=0.5)
    return out


def largest_island(ab, n):
    """(manifolds, bodies) of the largest rod-rod island, by union-find on the host."""
    par = np.arange(n)

    def find(i):
        while par[i] != i:
            par[i] = par[par[i]]
            i = par[i]
        return i

    rr = ab[ab[:, 1] >= 0]
    for a, b in rr:
        ra, rb = find(a), find(b)
        if ra != rb:
            par[ra] = rb
    lab = np.array([find(a) for a in rr[:, 0]])
    u, cnt = np.unique(lab, return_counts=True)
    big = u[np.argmax(cnt)]
    bodies = np.unique(rr[lab == big].ravel())
    return int(cnt.max()), int(len(bodies))


@pytest.mark.parametrize("proj", [0, 1, 2, 3])
@pytest.mark.parametrize("scene", ["clusters", "big_clusters", "log_pile", "pile_and_clusters"])
def test_island_sweeps_bitwise_equal_serial_sweeps_in_solve_order(gpu_ctx_factory, proj, scene):
    B = {"clusters": clusters, "big_clusters": big_clusters, "log_pile": log_pile,
         "pile_and_clusters": pile_and_clusters}[scene]()
    print(f"island sweeps {scene} proj={proj}", flush=True)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.1, cfm=1e-4, omega=1.3, tol=0.0, max_sweeps=9, check_every=1, friction_proj=proj,
                     warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p, allow_noconv=True)
    nisl, counts = ctx.island_info()
    if scene == "clusters":
        assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0, counts  # tiny, small and medium islands present
    elif scene == "big_clusters":
        assert counts[3] > 0, counts  # large islands (bodies in LDS)
    else:
        m, nb = largest_island(ab, len(B["r"]))
        assert nb > 3360 and m <= 8192 and counts[3] >= 1, (m, nb, counts)  # a huge island (colour-by-colour path)
        if scene == "pile_and_clusters":
            assert counts[0] > 0 and counts[1] > 0, counts
    order = ctx.solve_order(nm)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam0)
    op = O.nsc_params(mu=B["mu"], beta=0.1, cfm=1e-4, omega=1.3, friction_proj=proj)
    W.sweeps(op, 9, order)
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g, lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])


def test_island_and_global_solves_agree(gpu_ctx_factory):
    B = clusters(seed=5, mu=0.0)  # frictionless: unique solution per island
    out = []
    for mode in (0, 1):
        ctx = gpu_ctx_factory()
        gpu_load(ctx, B)
        ctx.set_solver_mode(mode)
        p = L.nsc_params(beta=0.0, cfm=1e-3, omega=1.0, tol=1e-12, max_sweeps=200000, check_every=4, friction_proj=1,
                         warm_mode=0)
        ctx.detect(L.detect_params())
        nm = ctx.nsc_build(p)
        st = ctx.nsc_solve(p)
        assert st["residual"] <= 1e-12
        out.append(ctx.nsc_manifolds(nm)[2])
    assert np.allclose(out[0], out[1], rtol=1e-7, atol=1e-10)


def test_stall_rule(gpu_ctx_factory):
    """tol = 0 never converges: the residual reaches the rounding floor, stops
    dropping between milestones and the solve stops long before max_sweeps.  A
    solve that converges is not affected by the rule (bitwise the same)."""
    B = clusters(seed=5, mu=0.0)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=1e-3, omega=1.0, tol=0.0, max_sweeps=400000, check_every=4, friction_proj=1,
                     warm_mode=0, stall_sweeps=64)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    st = ctx.nsc_solve(p, allow_noconv=True)
    assert 64 <= st["sweeps"] < 400000 and st["residual"] < 1e-13, st
    lams = []
    for stall in (0, 64):
        gpu_load(ctx, B)
        q = L.nsc_params(beta=0.0, cfm=1e-3, omega=1.0, tol=1e-11, max_sweeps=400000, check_every=4,
                         friction_proj=1, warm_mode=0, stall_sweeps=stall)
        ctx.detect(L.detect_params())
        ctx.nsc_build(q)
        st = ctx.nsc_solve(q)
        assert st["residual"] <= 1e-11
        lams.append(ctx.nsc_manifolds(nm)[2])
    assert np.array_equal(lams[0], lams[1])


def test_stall_rule_omega_fallback_breaks_limit_cycle(gpu_ctx_factory):
    """The captured limit cycle: without the stall rule the over-relaxed solve
    ends at max_sweeps on the cycle; with it, omega drops to stall_omega = 1
    at the first stall and the solve converges (the residual is the oracle's);
    with stall_omega = 0 the solve stops at the first stall instead."""
    import os

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "c4_limit_cycle.npz"))
    out = {}
    for stall, stall_omega in ((0, 1.0), (64, 1.0), (64, 0.0)):
        ctx = gpu_ctx_factory()
        ctx.set_bodies(d["x"], d["q"], d["v"], d["w"], d["r"], d["h"], d["mass"], d["rest"])
        ctx.set_systems([0, len(d["r"])], [L.system_desc(d["lo"], d["hi"], (1, 1, 1), 0, 0, float(d["mu"]))])
        ctx.set_physics(L.physics((0, 0, 0), 1e-3, 0, 0, 1e30))
        p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-8, max_sweeps=3000, check_every=4, friction_proj=1,
                         warm_mode=1, stall_sweeps=stall, stall_omega=stall_omega)
        assert ctx.detect(L.detect_params()) == 11
        q0 = L.nsc_params(warm_mode=0)
        ctx.nsc_build(q0)
        rec, ab, _ = ctx.nsc_manifolds(11)
        assert np.array_equal(ab, d["ab"])
        ctx.nsc_set_lambda(d["lam0"])
        st = ctx.nsc_solve(p, allow_noconv=True)
        out[(stall, stall_omega)] = st
        if stall and stall_omega > 0:
            _, _, lam = ctx.nsc_manifolds(11)
            s1 = ctx.download_state()
            W = O.World(d["x"], d["q"], s1["v"], s1["w"], d["r"], d["h"], d["mass"], d["rest"])
            W.set_box(list(d["lo"]), list(d["hi"]), [1, 1, 1], 0, 0)
            W.set_physics((0, 0, 0), 1e-3)
            W.set_manifolds(rec, ab, lam)
            assert W.residual(O.nsc_params(mu=float(d["mu"]), beta=0.0, cfm=0.0, omega=1.3, friction_proj=1)) <= 1e-8
    cyc, fb, stop = out[(0, 1.0)], out[(64, 1.0)], out[(64, 0.0)]
    assert cyc["residual"] > 0.1 and cyc["sweeps"] >= 3000, cyc
    assert fb["residual"] <= 1e-8 and fb["sweeps"] < 3000, fb
    # the same trajectory up to the milestone (64 * 2^k) at which the
    # fallback solve switched omega; this one stops there
    assert stop["residual"] > 0.1 and stop["sweeps"] in (128, 256, 512, 1024, 2048), stop
    assert stop["sweeps"] < fb["sweeps"]


def test_huge_island_solve_reaches_residual_checked_by_oracle(gpu_ctx_factory):
    """The host-driven colour-by-colour solve of a huge island (residual
    checks, stall rule) stops at the tolerance; the oracle's residual of the
    GPU impulses confirms it, and the impulses are admissible."""
    B = log_pile(ny=90, nz=40, seed=98)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-6, max_sweeps=20000, check_every=4, friction_proj=1,
                     warm_mode=0, stall_sweeps=1000, stall_omega=1.0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    m, nb = largest_island(ab, len(B["r"]))
    assert nb > 3360 and m <= 8192, (m, nb)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p)
    assert st["residual"] <= 1e-6 and st["sweeps"] > 4, st
    _, _, lam = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam)
    W.apply_warm()
    so = W.state()
    assert np.allclose(so["v"], s1["v"], rtol=0, atol=1e-10)
    W.set_state(v=s1["v"], w=s1["w"])
    op = O.nsc_params(mu=B["mu"], beta=0.0, cfm=0.0, omega=1.3, friction_proj=1)
    assert W.residual(op) <= 1e-6 * (1 + 1e-6)
    assert lam[:, 0].min() >= 0.0
    assert np.all(np.hypot(lam[:, 1], lam[:, 2]) <= B["mu"] * lam[:, 0] * (1 + 1e-12) + 1e-15)


def test_huge_island_graph_replays_bitwise_equal_serial_sweeps(gpu_ctx_factory):
    """Check blocks of four sweeps replayed from captured graphs (one block,
    one block, then a replay of four blocks whose last one starts after the
    sweep budget is spent and must store nothing) equal twelve serial sweeps
    of the oracle in the same order, bit for bit."""
    B = log_pile(ny=90, nz=40, seed=100)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.1, cfm=1e-4, omega=1.3, tol=0.0, max_sweeps=12, check_every=4, friction_proj=1,
                     warm_mode=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p, allow_noconv=True)
    assert st["sweeps"] == 12, st
    order = ctx.solve_order(nm)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam0)
    W.sweeps(O.nsc_params(mu=B["mu"], beta=0.1, cfm=1e-4, omega=1.3, friction_proj=1), 12, order)
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g, lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])


def test_huge_island_stops_mid_replay_at_the_tolerance(gpu_ctx_factory):
    """The device-side check ends the solve at the first block whose residual
    meets the tolerance, inside a multi-block replay: the impulses and
    velocities equal the oracle's serial sweeps for exactly the reported
    sweep count, so the sweeps launched after the stop wrote nothing."""
    B = log_pile(ny=90, nz=40, seed=101)
    ctx = gpu_ctx_factory()
    gpu_load(ctx, B)
    p = L.nsc_params(beta=0.0, cfm=0.0, omega=1.3, tol=1e-4, max_sweeps=20000, check_every=4, friction_proj=1,
                     warm_mode=0, stall_sweeps=0)
    ctx.detect(L.detect_params())
    nm = ctx.nsc_build(p)
    rec, ab, lam0 = ctx.nsc_manifolds(nm)
    m, nb = largest_island(ab, len(B["r"]))
    assert nb > 3360 and m <= 8192, (m, nb)
    s0 = ctx.download_state()
    st = ctx.nsc_solve(p, allow_noconv=True)
    k = st["sweeps"]
    print(f"huge island converged after {k} sweeps, residual {st['residual']}", flush=True)
    assert k % 4 == 0 and k > 8 and st["residual"] <= 1e-4, st
    order = ctx.solve_order(nm)
    _, _, lam_g = ctx.nsc_manifolds(nm)
    s1 = ctx.download_state()
    W = oracle_load(B)
    W.set_state(v=s0["v"], w=s0["w"])
    W.set_manifolds(rec, ab, lam0)
    W.sweeps(O.nsc_params(mu=B["mu"], beta=0.0, cfm=0.0, omega=1.3, friction_proj=1), k, order)
    _, _, lam_o = W.manifolds()
    so = W.state()
    assert np.array_equal(lam_g, lam_o)
    assert np.array_equal(s1["v"], so["v"]) and np.array_equal(s1["w"], so["w"])
