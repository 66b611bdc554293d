"""GPU detection parity: the librdk contact list must equal the oracle's
bit for bit (pairs, order, and every FP64 field) on the same FP64 state."""
import numpy as np
import pytest

import oracle_py as O
from rod_dynamics_py import _lib as L
from rod_dynamics_py import scenes

pytestmark = pytest.mark.gpu


def load(ctx, x, q, r, h, mass, lo, hi, periodic, fw=0, cw=0):
    n = len(r)
    ctx.set_bodies(x, q, np.zeros((n, 3)), np.zeros((n, 3)), r, h, mass, np.zeros(n))
    ctx.set_systems([0, n], [L.system_desc(lo, hi, periodic, fw, cw, 0.5)])
    ctx.set_physics(L.physics())


def compare(gc, oc, nc):
    assert np.array_equal(gc["a"], oc["a"]) and np.array_equal(gc["b"], oc["b"])
    for k in ("pa", "pb", "n", "dist", "shift", "s", "t", "type"):
        assert np.array_equal(gc[k], oc[k]), k


def rods(n, side, seed, rlo=0.2, rhi=0.5, hlo=0.5, hhi=3.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, side, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q, rng.uniform(rlo, rhi, n), rng.uniform(hlo, hhi, n)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 0, 0), (1, 1, 0)])
@pytest.mark.parametrize("chunk_len", [0.0, 1.0, 2.5])
def test_detect_bit_identical_polydisperse(gpu_ctx_factory, periodic, chunk_len):
    n, side = 3000, 30.0
    x, q, r, h = rods(n, side, 5)
    print(f"detect polydisperse periodic={periodic} chunk_len={chunk_len}", flush=True)
    ctx = gpu_ctx_factory()
    load(ctx, x, q, r, h, np.ones(n), [0] * 3, [side] * 3, periodic)
    W = O.World(x, q, None, None, r, h, np.ones(n))
    W.set_box([0] * 3, [side] * 3, periodic)
    for delta in (0.0, 0.05):
        nc = ctx.detect(L.detect_params(delta, chunk_len, 1.0))
        no = W.detect(0, delta)
        assert nc == no and nc > 100
        compare(ctx.contacts(nc), W.contacts(no), nc)


@pytest.mark.parametrize("block", [64, 128, 512])
def test_detect_bit_identical_for_every_broad_phase_workgroup_size(gpu_ctx_factory, block):
    """The autotuned broad-phase workgroup size changes the launch shape only:
    the contact list equals the oracle's for each size."""
    n, side = 3000, 30.0
    x, q, r, h = rods(n, side, 6)
    print(f"detect with broad-phase workgroup {block}", flush=True)
    ctx = gpu_ctx_factory()
    load(ctx, x, q, r, h, np.ones(n), [0] * 3, [side] * 3, (1, 1, 1))
    ctx.set_tuning(L.TUNE_BROAD_BLOCK, block)
    W = O.World(x, q, None, None, r, h, np.ones(n))
    W.set_box([0] * 3, [side] * 3, (1, 1, 1))
    for chunk_len in (0.0, 1.0):
        nc = ctx.detect(L.detect_params(0.0, chunk_len, 1.0))
        no = W.detect(0, 0.0)
        assert nc == no and nc > 100
        compare(ctx.contacts(nc), W.contacts(no), nc)


def staged_scene(scene):
    """Rods and systems of the broad-phase layout tests: (n, X, Q, R, H, loader)."""
    if scene == "multi":
        rng = np.random.default_rng(21)
        parts, descs, offs = [], [], [0]
        for k, (n, side, per, hl) in enumerate([(700, 18.0, (1, 1, 1), 2.0), (1500, 25.0, (1, 1, 0), 4.0),
                                                (300, 9.0, (1, 1, 1), 1.0), (900, 30.0, (0, 0, 0), 8.0)]):
            x, q, r, h = rods(n, side, 200 + k, hlo=0.5, hhi=hl)
            x += 40.0 * k
            parts.append((x, q, r, h, side, per, 40.0 * k))
            descs.append(L.system_desc([40.0 * k] * 3, [40.0 * k + side] * 3, per, 0, 0, 0.5))
            offs.append(offs[-1] + n)
        X, Q, R, H = (np.concatenate([p_[i] for p_ in parts]) for i in range(4))
        n = len(R)

        def put(ctx):
            ctx.set_bodies(X, Q, np.zeros((n, 3)), np.zeros((n, 3)), R, H, np.ones(n), np.zeros(n))
            ctx.set_systems(offs, descs)
            ctx.set_physics(L.physics())
        return n, X, Q, R, H, None, None, put
    n, side, per = {"periodic": (3000, 30.0, (1, 1, 1)), "open": (3000, 30.0, (0, 0, 0)),
                    "dense": (6000, 16.0, (1, 1, 0))}[scene]
    X, Q, R, H = rods(n, side, 31)
    return n, X, Q, R, H, side, per, lambda ctx: load(ctx, X, Q, R, H, np.ones(n), [0] * 3, [side] * 3, per)


@pytest.mark.parametrize("band", [1, 2, 5])
@pytest.mark.parametrize("stage", [0, 1])
@pytest.mark.parametrize("scene", ["periodic", "open", "dense", "multi"])
def test_banded_cell_order_bit_identical(gpu_ctx_factory, band, stage, scene):
    """The cell order in bands of y-rows (RDK_TUNE_CELL_BAND) changes which
    lane owns a chunk pair and the order of the sorted chunks, not which pairs
    are tested: contact list and candidate counts equal to the unbanded
    order's, and the contact list equal to the oracle's, for the plain and
    the LDS-staged walk (bands of 1, 2 and 5 rows, so the last band of a grid
    is short and periodic rows wrap across bands)."""
    print(f"banded cell order: {scene}, band {band}, staged {stage}", flush=True)
    n, X, Q, R, H, side, per, put = staged_scene(scene)
    res = {}
    for b in (0, band):
        ctx = gpu_ctx_factory()
        put(ctx)
        ctx.set_tuning(L.TUNE_BROAD_BLOCK, 256)
        ctx.set_tuning(L.TUNE_BROAD_STAGE, stage)
        ctx.set_tuning(L.TUNE_CELL_BAND, b)
        for chunk_len in (0.0, 1.0):
            nc = ctx.detect(L.detect_params(0.02, chunk_len, 1.0))
            res[b, chunk_len] = (ctx.contacts(nc), ctx.detect_info())
    for chunk_len in (0.0, 1.0):
        (c0, i0), (c1, i1) = res[0, chunk_len], res[band, chunk_len]
        assert len(c0["a"]) > 100
        compare(c1, c0, len(c0["a"]))
        assert i1 == i0, (i0, i1)
    if scene != "multi":
        W = O.World(X, Q, None, None, R, H, np.ones(n))
        W.set_box([0] * 3, [side] * 3, per)
        no = W.detect(0, 0.02)
        c1 = res[band, 1.0][0]
        assert len(c1["a"]) == no
        compare(c1, W.contacts(no), no)


@pytest.mark.parametrize("block", [64, 256, 512])
@pytest.mark.parametrize("scene", ["periodic", "open", "dense", "multi"])
def test_lds_staged_broad_phase_bit_identical(gpu_ctx_factory, block, scene):
    """Stencil rows staged in LDS (RDK_TUNE_BROAD_STAGE 1) change where the
    candidate records are read, not which pairs are tested: contact list equal
    to the oracle's and candidate counts equal to the unstaged walk's, for
    dilute, dense (many chunks per cell, entries that overflow the stage) and
    multi-system layouts."""
    print(f"staged broad phase: {scene}, workgroup {block}", flush=True)
    if scene == "multi":
        rng = np.random.default_rng(21)
        parts, descs, offs = [], [], [0]
        for k, (n, side, per, hl) in enumerate([(700, 18.0, (1, 1, 1), 2.0), (1500, 25.0, (1, 1, 0), 4.0),
                                                (300, 9.0, (1, 1, 1), 1.0), (900, 30.0, (0, 0, 0), 8.0)]):
            x, q, r, h = rods(n, side, 200 + k, hlo=0.5, hhi=hl)
            x += 40.0 * k
            parts.append((x, q, r, h, side, per, 40.0 * k))
            descs.append(L.system_desc([40.0 * k] * 3, [40.0 * k + side] * 3, per, 0, 0, 0.5))
            offs.append(offs[-1] + n)
        X, Q, R, H = (np.concatenate([p_[i] for p_ in parts]) for i in range(4))
        n = len(R)
    else:
        n, side, per = {"periodic": (3000, 30.0, (1, 1, 1)), "open": (3000, 30.0, (0, 0, 0)),
                        "dense": (6000, 16.0, (1, 1, 0))}[scene]
        X, Q, R, H = rods(n, side, 31)
    res = {}
    for stage in (0, 1):
        ctx = gpu_ctx_factory()
        if scene == "multi":
            ctx.set_bodies(X, Q, np.zeros((n, 3)), np.zeros((n, 3)), R, H, np.ones(n), np.zeros(n))
            ctx.set_systems(offs, descs)
            ctx.set_physics(L.physics())
        else:
            load(ctx, X, Q, R, H, np.ones(n), [0] * 3, [side] * 3, per)
        ctx.set_tuning(L.TUNE_BROAD_BLOCK, block)
        ctx.set_tuning(L.TUNE_BROAD_STAGE, stage)
        for chunk_len in (0.0, 1.0):
            nc = ctx.detect(L.detect_params(0.02, chunk_len, 1.0))
            res[stage, chunk_len] = (ctx.contacts(nc), ctx.detect_info())
    for chunk_len in (0.0, 1.0):
        (c0, i0), (c1, i1) = res[0, chunk_len], res[1, chunk_len]
        assert len(c0["a"]) > 100
        compare(c1, c0, len(c0["a"]))
        assert i1 == i0, (i0, i1)
    if scene != "multi":
        W = O.World(X, Q, None, None, R, H, np.ones(n))
        W.set_box([0] * 3, [side] * 3, per)
        no = W.detect(0, 0.02)
        c1 = res[1, 1.0][0]
        assert len(c1["a"]) == no
        compare(c1, W.contacts(no), no)


def test_detect_positions_outside_box_and_static_pairs(gpu_ctx_factory):
    n, side = 1500, 20.0
    x, q, r, h = rods(n, side, 9)
    x[:50] += side * 1.3  # unwrapped positions on periodic axes / outside a walled axis
    mass = np.ones(n)
    mass[::7] = 0.0  # static bodies: static-static pairs are skipped
    for periodic in [(1, 1, 1), (1, 1, 0)]:
        ctx = gpu_ctx_factory()
        load(ctx, x, q, r, h, mass, [0] * 3, [side] * 3, periodic)
        W = O.World(x, q, None, None, r, h, mass)
        W.set_box([0] * 3, [side] * 3, periodic)
        nc = ctx.detect(L.detect_params(0.02, 1.5, 1.0))
        no = W.detect(0, 0.02)
        assert nc == no
        compare(ctx.contacts(nc), W.contacts(no), nc)


def test_detect_c2_like_and_c1_like(gpu_ctx_factory):
    sc = scenes.c2_settling(n=4000)
    ctx = gpu_ctx_factory()
    sc.load(ctx)
    W = O.world_from_scene(sc)
    # the RSA start has no overlaps; a contact margin of 1.0 gives many near pairs
    for delta, chunk in [(1.0, 0.0), (1.0, 4.0), (0.0, 0.0)]:
        nc = ctx.detect(L.detect_params(delta, chunk, 1.0))
        no = W.detect(0, delta)
        assert nc == no
        if nc:
            compare(ctx.contacts(nc), W.contacts(no), nc)
    sc1 = scenes.c1_thermal(n=1000)
    ctx1 = gpu_ctx_factory()
    sc1.load(ctx1)
    W1 = O.world_from_scene(sc1)
    nc = ctx1.detect(L.detect_params(0.5, 2.0, 1.0))
    no = W1.detect(1, 0.5)
    assert nc == no and nc > 500
    compare(ctx1.contacts(nc), W1.contacts(no), nc)


def test_detect_capacity_regrow(gpu_ctx_factory):
    n, side = 2000, 18.0
    x, q, r, h = rods(n, side, 13)
    ctx = gpu_ctx_factory()
    load(ctx, x, q, r, h, np.ones(n), [0] * 3, [side] * 3, (1, 1, 1))
    ctx.set_pair_capacity(16)  # force the overflow path from the host
    nc = ctx.detect(L.detect_params(0.05, 0.0, 1.0))
    info = ctx.detect_info()
    assert info["regrows"] >= 1
    W = O.World(x, q, None, None, r, h, np.ones(n))
    W.set_box([0] * 3, [side] * 3, (1, 1, 1))
    no = W.detect(0, 0.05)
    assert nc == no and nc > 16
    compare(ctx.contacts(nc), W.contacts(no), nc)


def test_multi_system_detection_equals_per_system_oracle(gpu_ctx_factory):
    """Several independent boxes in one context (the batched workload's
    layout): the contact list is the concatenation of each system's own list,
    with body indices offset by the system's first body."""
    rng = np.random.default_rng(12)
    parts, descs, offs = [], [], [0]
    for k, (n, side, per, hl) in enumerate([(700, 18.0, (1, 1, 1), 2.0), (1500, 25.0, (1, 1, 1), 4.0),
                                            (400, 12.0, (1, 1, 0), 1.0), (900, 30.0, (1, 1, 1), 8.0)]):
        x, q, r, h = rods(n, side, 100 + k, hlo=0.5, hhi=hl)
        x += 7.0 * k  # boxes at different origins
        parts.append((x, q, r, h, side, per, 7.0 * k))
        descs.append(L.system_desc([7.0 * k] * 3, [7.0 * k + side] * 3, per, 0, 0, 0.5))
        offs.append(offs[-1] + n)
    X = np.concatenate([p[0] for p in parts])
    Q = np.concatenate([p[1] for p in parts])
    R = np.concatenate([p[2] for p in parts])
    H = np.concatenate([p[3] for p in parts])
    n = len(R)
    ctx = gpu_ctx_factory()
    ctx.set_bodies(X, Q, np.zeros((n, 3)), np.zeros((n, 3)), R, H, np.ones(n), np.zeros(n))
    ctx.set_systems(offs, descs)
    ctx.set_physics(L.physics())
    for chunk_len in (0.0, 2.0):
        nc = ctx.detect(L.detect_params(0.0, chunk_len, 1.0))
        gc = ctx.contacts(nc)
        expect = {k: [] for k in ("a", "b", "pa", "pb", "n", "dist", "shift", "s", "t", "type")}
        for k, (x, q, r, h, side, per, o) in enumerate(parts):
            W = O.World(x, q, None, None, r, h, np.ones(len(r)))
            W.set_box([o] * 3, [o + side] * 3, per)
            no = W.detect(0, 0.0)
            oc = W.contacts(no)
            for key in expect:
                v = oc[key] + offs[k] if key in ("a", "b") else oc[key]
                expect[key].append(v)
        expect = {k: np.concatenate(v) for k, v in expect.items()}
        assert nc == len(expect["a"]) and nc > 200
        compare(gc, expect, nc)


def test_affine_scale_then_detect(gpu_ctx_factory):
    """rdk_affine_scale maps x -> lo + s (x - lo) and the box likewise; the
    next detection (grid resized) equals the oracle's on the scaled state."""
    n, side = 2500, 30.0
    x, q, r, h = rods(n, side, 31)
    lo = np.array([-3.0, 1.0, 2.0])
    x = x + lo
    ctx = gpu_ctx_factory()
    load(ctx, x, q, r, h, np.ones(n), lo, lo + side, (1, 1, 1))
    dp = L.detect_params(0.0, 2.0, 1.0)
    ctx.detect(dp)  # grid sized for the first box
    s = 0.75  # exact in binary: the box spans stay exact on both sides
    ctx.affine_scale(s)
    st = ctx.download_state()
    assert np.array_equal(st["x"], lo + s * (x - lo))
    nc = ctx.detect(dp)
    W = O.World(lo + s * (x - lo), q, None, None, r, h, np.ones(n))
    W.set_box(lo, lo + s * side, (1, 1, 1))
    no = W.detect(0, 0.0)
    assert nc == no and nc > 200
    compare(ctx.contacts(nc), W.contacts(no), nc)


def test_reference_broadphase_benchmark_geometry(gpu_ctx_factory):
    """The reference's broad-phase benchmark geometry (benchmark_broadphase.cpp)
    at N = 512 and 4000: the GPU list equals the oracle's bit for bit, and the
    counts are the ones the reference publishes (1730, 13190)."""
    from bench_geometry import bench_rods

    for n, ref in ((512, 1730), (4000, 13190)):
        x, q, r, h, lo, hi = bench_rods(n)
        ctx = gpu_ctx_factory()
        load(ctx, x, q, r, h, np.ones(n), [lo] * 3, [hi] * 3, (1, 1, 1))
        for chunk_len in (0.0, 0.2):
            nc = ctx.detect(L.detect_params(0.005, chunk_len, 1.0))
            assert nc == ref
            W = O.World(x, q, None, None, r, h, np.ones(n))
            W.set_box([lo] * 3, [hi] * 3, (1, 1, 1))
            no = W.detect(0, 0.005)
            compare(ctx.contacts(nc), W.contacts(no), nc)


def test_reference_packings(gpu_ctx_factory):
    """Frozen relaxed packings of the reference (aspect ratios 10 to 1000, in
    free space) at the contact criterion of its metadata (segment distance
    < 1.01 d): the GPU list equals the oracle's bit for bit and its count
    equals assets/packings_metadata.csv's n_contacts."""
    from packings import contact_delta, free_box, golden_packings

    for pid, x, q, r, h, ref in golden_packings():
        print(f"packing {pid}", flush=True)
        n = len(r)
        lo, hi = free_box(x)
        ctx = gpu_ctx_factory()
        load(ctx, x, q, r, h, np.ones(n), lo, hi, (0, 0, 0))
        delta = contact_delta(r)
        W = O.World(x, q, None, None, r, h, np.ones(n))
        W.set_box(lo, hi, (0, 0, 0))
        no = W.detect(0, delta)
        oc = W.contacts(no)
        for chunk_len in (0.0, 0.1):
            nc = ctx.detect(L.detect_params(delta, chunk_len, 1.0))
            assert nc == ref, (pid, chunk_len, nc, ref)
            compare(ctx.contacts(nc), oc, nc)


@pytest.mark.parametrize("pbc", [False, True])
def test_plugin_raw_records(gpu_ctx_factory, pbc):
    """rdk_detect_capsule_pairs_raw, the one-call replacement of the reference's
    cudaDetectCapsulePairsTwoPass: FP32 capsules in, GpuContactRaw-layout
    records out.  Pairs equal the oracle's FP64 detection of the same
    (FP32-valued) capsules; every float field is the oracle's FP64 value rounded
    to FP32; surface_limit = r_a + r_b; a too-small capacity is reported."""
    n, side = 2000, 25.0
    x, q, r, h = rods(n, side, 12)
    if pbc:
        x = x - side / 2  # the reference keeps positions in [min, max); any origin works
    x32, q32, r32, h32 = (np.asarray(v, np.float32) for v in (x, q, r, h))
    print(f"raw plug-in records pbc={pbc}", flush=True)
    ctx = gpu_ctx_factory()
    margin = 0.05
    raw = ctx.detect_capsule_pairs_raw(x32, q32, r32, h32, margin, pbc=pbc, box=(side, side, side))
    W = O.World(x32.astype(np.float64), q32.astype(np.float64), None, None, r32.astype(np.float64),
                h32.astype(np.float64), np.ones(n))
    if pbc:
        W.set_box(np.zeros(3), np.full(3, side), (1, 1, 1))
    else:
        W.set_box(x32.min(0) - 10.0, x32.max(0) + 10.0, (0, 0, 0))
    no = W.detect(0, float(np.float32(margin)))
    oc = W.contacts(no)
    assert len(raw) == no > 0
    assert np.array_equal(raw["a"], oc["a"]) and np.array_equal(raw["b"], oc["b"])
    f32 = lambda v: np.asarray(v, np.float64).astype(np.float32)
    for k, name in enumerate("xyz"):
        assert np.array_equal(raw[f"p{name}_a"], f32(oc["pa"][:, k]))
        assert np.array_equal(raw[f"p{name}_b"], f32(oc["pb"][:, k]))
        assert np.array_equal(raw[f"n{name}"], f32(oc["n"][:, k]))
        assert np.array_equal(raw[f"shift_b{name}"], f32(oc["shift"][:, k]))
    assert np.array_equal(raw["dist"], f32(oc["dist"]))
    assert np.array_equal(raw["s"], f32(oc["s"])) and np.array_equal(raw["t"], f32(oc["t"]))
    rr = r32.astype(np.float64)
    assert np.array_equal(raw["surface_limit"], f32(rr[raw["a"]] + rr[raw["b"]]))
    assert np.all(raw["dist"] < raw["surface_limit"] + np.float32(margin) + 1e-6)
    if pbc:
        assert np.any(raw["shift_bx"] != 0) or np.any(raw["shift_by"] != 0)
    from rod_dynamics_py import RdkError

    with pytest.raises(RdkError):
        ctx.detect_capsule_pairs_raw(x32, q32, r32, h32, margin, pbc=pbc, box=(side, side, side), cap=no - 1)
    again = ctx.contacts_raw()
    assert np.array_equal(again, raw)


def test_plugin_raw_records_edge_cases(gpu_ctx_factory):
    """No capsules, and capsules far apart: no records, no error."""
    ctx = gpu_ctx_factory()
    z = np.zeros((0, 3), np.float32)
    assert len(ctx.detect_capsule_pairs_raw(z, np.zeros((0, 4)), np.zeros(0), np.zeros(0), 0.1)) == 0
    x = np.array([[0, 0, 0], [50, 0, 0]], np.float32)
    q = np.array([[1, 0, 0, 0], [1, 0, 0, 0]], np.float32)
    out = ctx.detect_capsule_pairs_raw(x, q, [0.5, 0.5], [1.0, 1.0], 0.1)
    assert len(out) == 0
    x[1, 0] = 0.9  # parallel rods 0.9 apart, radii 0.5: one contact, normal along +x
    out = ctx.detect_capsule_pairs_raw(x, q, [0.5, 0.5], [1.0, 1.0], 0.0)
    assert len(out) == 1 and out["a"][0] == 0 and out["b"][0] == 1
    assert out["dist"][0] == np.float32(0.9) and out["nx"][0] == np.float32(1.0)
    assert out["surface_limit"][0] == np.float32(1.0)
