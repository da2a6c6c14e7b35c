"""phx_propagate_noalloc (no device allocation in the call) against
phx_propagate with per-lane (grid) epoch records, force_grid = 1: the same
kernels and roundings, so the final epochs and states agree bit for bit; two
calls in a row on the same device scratch array, the second with fewer objects
and a shorter duration; and its limits. The caller's device buffers come from
hipMalloc of the HIP runtime libphx links (the test allocates, not the library)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime the library uses."""

    def __init__(self):
        import parahpop_amd
        parahpop_amd.library()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0   # host to device
        return p.value

    def download(self, ptr, shape):
        out = np.zeros(shape)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(ctx, hip, t, duration):
    n = t["n"]
    ptrs = [hip.upload(t["epoch"]), hip.upload(t["state"])] + [hip.upload(t[k]) for k in ("mass", "drag_area", "srp_area", "cd", "cr")]
    st = ctx.propagate_noalloc(n, *ptrs, duration)
    return hip.download(ptrs[0], (n,)), hip.download(ptrs[1], (6, n)), st


def _subset(t, idx):
    out = {k: (v[idx] if hasattr(v, "shape") and v.ndim == 1 else v) for k, v in t.items()}
    out["state"] = np.ascontiguousarray(t["state"][:, idx])
    out["n"] = len(idx)
    return out


def test_noalloc_equals_propagate_twice(full_workload):
    import parahpop_amd
    cfg, t = full_workload
    t = dict(t)
    t["epoch"] = t["epoch"] + np.arange(t["n"]) * 0.37 / 86400.0   # per-object epochs
    ctx = parahpop_amd.Context(cfg, force_grid=1)
    hip = _Hip()
    try:
        e_ref, s_ref, st_ref = ctx.propagate(t, 1800.0)
        e, s, st = _noalloc(ctx, hip, t, 1800.0)
        # per-object epochs: some spans round above 1800 s and take a shortened 31st step
        assert st["object_steps"] == st_ref["object_steps"] >= 30 * t["n"]
        np.testing.assert_array_equal(e, e_ref)
        np.testing.assert_array_equal(s, s_ref)
        # second call: fewer objects, shorter run, on the array the first left
        small = _subset(t, np.arange(40))
        e_ref2, s_ref2, st_ref2 = ctx.propagate(small, 600.0)
        e2, s2, st2 = _noalloc(ctx, hip, small, 600.0)
        assert st2["object_steps"] == st_ref2["object_steps"] >= 10 * 40
        np.testing.assert_array_equal(e2, e_ref2)
        np.testing.assert_array_equal(s2, s_ref2)
    finally:
        hip.free()


def test_noalloc_limits(full_workload):
    import parahpop_amd
    cfg, t = full_workload
    ctx = parahpop_amd.Context(cfg)
    hip = _Hip()
    try:
        p = [hip.upload(np.zeros(8)) for _ in range(7)]
        with pytest.raises(parahpop_amd.PhxError) as ex:
            ctx.propagate_noalloc(16385, *p, 600.0)   # rejected before any device work
        assert "at most 16384" in str(ex.value)
        bad = _subset(t, np.arange(8))
        bad["mass"] = bad["mass"].copy()
        bad["mass"][3] = 0.0
        with pytest.raises(parahpop_amd.PhxError) as ex:
            _noalloc(ctx, hip, bad, 600.0)
        assert "Invalid target epoch or physical parameter" in str(ex.value)
        with pytest.raises(parahpop_amd.PhxError) as ex:
            _noalloc(ctx, hip, _subset(t, np.arange(8)), 40 * 86400.0)   # span beyond the record rows
        assert "record rows" in str(ex.value)
    finally:
        hip.free()
