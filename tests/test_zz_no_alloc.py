"""sl_ripple_frame_noalloc (the ripple frame on caller buffers, no device allocation) against the
existing entry point sl_run_stage(sim, SL_S_RIPPLE) with the mask off: state, texture and the cleared
impulse grid bit for bit, on two calls in a row, the second with a smaller layer (M = 512, then 128).
The caller's device buffers come from the HIP runtime the library itself is linked against."""
import ctypes
import importlib

import numpy as np
import pytest

import twin

suiri = importlib.import_module("hahenn-_amd")

pytestmark = pytest.mark.gpu


def reference_frame(R, seed, dt=1 / 60, depth=0.04):
    sc = twin.small_scene(R=R, extra=dict(ripple_depth=depth))
    o, g = twin.make_twin(sc)
    o.close()
    rng = np.random.default_rng(seed)
    M = 2 * R
    src = np.zeros((R + 1) * (R + 1), np.int32)
    idx = rng.choice(src.size, 50, replace=False)
    src[idx] = rng.integers(1000, 2_000_000, 50)
    st = rng.normal(0, 1e-6, (M * M, 2)).astype(np.float32)
    g.upload(suiri.F_RSRC, src)
    g.upload(suiri.F_RSTATE, st)
    g.set_ripple_mask(0)
    g.run_stage(suiri.S_RIPPLE, 0, dt)
    out = {"state": g.readback(suiri.F_RSTATE).copy(), "tex": g.readback(suiri.F_RTEX).copy(),
           "src": g.readback(suiri.F_RSRC).copy(), "Lx": g.nx * g.h, "Lz": g.nz * g.h}
    cfg = g.cfg
    g.close()
    return cfg, src, st, out


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the process's HIP runtime (the caller's own buffers)."""

    def __init__(self):
        suiri.load()
        self.L = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        self.ptrs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        self.L.hipDeviceSynchronize()
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


@pytest.mark.parametrize("order", [(256, 32), (64, 16)])
def test_no_alloc_ripple_frame_equals_the_stage(order):
    for R, seed in zip(order, (3, 4)):
        print(f"no-alloc ripple frame R={R}", flush=True)
        cfg, src, st, want = reference_frame(R, seed)
        hip = Hip()
        tex0 = np.zeros(((R + 1) * (R + 1), 4), np.float32)
        ps, pr, pt = hip.upload(st), hip.upload(src), hip.upload(tex0)
        suiri.ripple_frame_noalloc(cfg, R, want["Lx"], want["Lz"], 1 / 60, 0.04, ps, pr, pt, None)
        assert np.array_equal(hip.download(ps, st).view(np.uint32), want["state"].view(np.uint32))
        assert np.array_equal(hip.download(pt, tex0).view(np.uint32), want["tex"].view(np.uint32))
        assert np.array_equal(hip.download(pr, src), want["src"]) and not want["src"].any()
        hip.close()


def test_no_alloc_rejects_sizes_beyond_the_maximum():
    cfg = suiri.config("standard")
    hip = Hip()
    b = hip.upload(np.zeros(16, np.float32))
    with pytest.raises(suiri.SuiriError, match="M = 2R <= 512"):
        suiri.ripple_frame_noalloc(cfg, 512, 0.5, 0.5, 1 / 60, 0.05, b, b, b)
    with pytest.raises(suiri.SuiriError, match="power of two"):
        suiri.ripple_frame_noalloc(cfg, 48, 0.5, 0.5, 1 / 60, 0.05, b, b, b)
    hip.close()
