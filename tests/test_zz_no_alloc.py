"""rdk_entanglement_noalloc (no device allocation; scratch in a static device
array per GPU) against rdk_entanglement, bit for bit, on two calls in a row
through the same context, the second with fewer rods and more systems, with
and without a cutoff; and its capacity error."""
import ctypes as C
import math

import numpy as np
import pytest

from rod_dynamics_py import _lib as L

pytestmark = pytest.mark.gpu


class DevArray:
    """Device buffer from the HIP runtime librdk itself links (already loaded
    with it), so the test does not bring a second runtime into the process."""

    def __init__(self, host):
        L.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes), 1) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def packing(sizes, seed):
    rng = np.random.default_rng(seed)
    xs, qs, offs, descs = [], [], [0], []
    for k, n in enumerate(sizes):
        side = 6.0 * n ** (1 / 3) + 4.0
        periodic = (1, 1, 1) if k % 2 == 0 else (0, 0, 0)
        xs.append(rng.uniform(0.0, side, size=(n, 3)))
        q = rng.normal(size=(n, 4))
        qs.append(q / np.linalg.norm(q, axis=1, keepdims=True))
        offs.append(offs[-1] + n)
        descs.append(L.system_desc((0, 0, 0), (side, side, side), periodic, mu=0.5))
    n = offs[-1]
    return dict(x=np.concatenate(xs), q=np.concatenate(qs), r=np.full(n, 0.5), h=rng.uniform(1.0, 4.0, n),
                offs=offs, descs=descs)


def load(ctx, B):
    n = len(B["r"])
    ctx.set_bodies(B["x"], B["q"], np.zeros((n, 3)), np.zeros((n, 3)), B["r"], B["h"], np.ones(n))
    ctx.set_systems(B["offs"], B["descs"])
    ctx.set_physics(L.physics())


def test_noalloc_equals_entanglement_on_two_calls(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    for call, (sizes, seed) in enumerate((([1900, 1200], 5), ([300, 1, 260, 139], 6))):
        B = packing(sizes, seed)
        load(ctx, B)
        nsys = len(sizes)
        for cutoff in (-1.0, 3.0):
            print(f"noalloc call {call} sizes {sizes} cutoff {cutoff}", flush=True)
            ref_s, ref_p = ctx.entanglement(cutoff)
            d_s = DevArray(np.full(nsys, math.nan))
            d_p = DevArray(np.full(nsys, -1, np.int64))
            ctx.entanglement_noalloc(d_s.ptr.value, d_p.ptr.value, cutoff)
            ctx.sync()
            got_s, got_p = d_s.get(), d_p.get()
            d_s.free()
            d_p.free()
            assert np.array_equal(got_s, ref_s), (got_s, ref_s)
            assert np.array_equal(got_p, ref_p), (got_p, ref_p)
            if cutoff < 0:
                assert np.array_equal(ref_p, [k * (k - 1) // 2 for k in sizes])


def test_noalloc_rejects_more_rods_than_its_scratch(gpu_ctx_factory):
    ctx = gpu_ctx_factory()
    n = 32769
    B = packing([n], 7)
    load(ctx, B)
    d_s = DevArray(np.zeros(1))
    d_p = DevArray(np.zeros(1, np.int64))
    with pytest.raises(L.RdkError) as e:
        ctx.entanglement_noalloc(d_s.ptr.value, d_p.ptr.value, -1.0)
    d_s.free()
    d_p.free()
    assert e.value.status == L.RDK_ERR_CAPACITY
    assert "RDK_NOALLOC_MAX_RODS" in str(e.value)
