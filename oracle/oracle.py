"""Oracle binding (test infrastructure only).

ctypes wrapper of oracle/_build/liboracle.so, the plain-C fp32 restatement of the reference liquid
step (oracle/suiri_oracle.c). Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker; never by the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


class OrQuality(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_double), ("cap", ctypes.c_uint32), ("ww_cap", ctypes.c_uint32),
        ("iters", ctypes.c_uint32), ("ripple_R", ctypes.c_uint32), ("rc_factor", ctypes.c_double),
        ("gravity", ctypes.c_float), ("flip", ctypes.c_float), ("ppc", ctypes.c_float),
        ("drift", ctypes.c_float), ("omega", ctypes.c_float), ("friction", ctypes.c_float),
        ("ww_rate", ctypes.c_float), ("r_sigma_rho", ctypes.c_float), ("r_nu", ctypes.c_float),
        ("r_film", ctypes.c_float), ("r_hp_k", ctypes.c_float), ("r_rho_kick", ctypes.c_float),
        ("wmap", ctypes.c_uint32), ("reorder_every", ctypes.c_uint32), ("reorder_phase", ctypes.c_uint32),
        ("top_height", ctypes.c_uint32),
        ("splash_rho", ctypes.c_double), ("splash_sigma", ctypes.c_double), ("splash_mu", ctypes.c_double),
        ("splash_k_jet", ctypes.c_float), ("splash_k_crown", ctypes.c_float), ("ripple_depth", ctypes.c_double),
        ("droplets", ctypes.c_uint32), ("drop_seed", ctypes.c_uint32), ("drop_d50", ctypes.c_float),
        ("drop_sigma_ln", ctypes.c_float), ("drop_vmin", ctypes.c_float), ("drop_vmax", ctypes.c_float),
    ]


class OrEmit(ctypes.Structure):
    _fields_ = [("a", ctypes.c_float * 4), ("b", ctypes.c_float * 4), ("c", ctypes.c_float * 4),
                ("d", ctypes.c_float * 4), ("info", ctypes.c_uint32 * 4)]


INFO_NAMES = ["nx", "ny", "nz", "h", "ox", "oy", "oz", "rim", "vp", "nodes", "cells", "cap", "ww_cap",
              "frame", "obj_count", "count", "ww_count", "impacts", "time", "vmax", "full", "cur",
              "ww_cur", "R", "M", "iters"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.or_create.restype = P
        L.or_create.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(OrQuality)]
        L.or_destroy.argtypes = [P]
        L.or_quality_default.argtypes = [ctypes.POINTER(OrQuality), ctypes.c_char_p]
        L.or_field.restype = P
        L.or_field.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
        L.or_info.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.or_set_scalar.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
        L.or_set_threads.argtypes = [P, ctypes.c_int]
        L.or_write_params.argtypes = [P, ctypes.c_double]
        L.or_set_objects.argtypes = [P, P, ctypes.c_uint32]
        L.or_set_objects.restype = ctypes.c_int
        for name in ["or_geom", "or_p2g", "or_normalize", "or_forces", "or_divergence", "or_project",
                     "or_g2p", "or_compact", "or_substep", "or_step_begin", "or_poll_stats",
                     "or_clear_water", "or_revolve", "or_rmaskbuild", "or_rscale", "or_reorder", "or_toph",
                     "or_toph_splat", "or_toph_blur"]:
            getattr(L, name).argtypes = [P]
        L.or_pressure.argtypes = [P, ctypes.c_uint32]
        L.or_count.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
        L.or_ww.argtypes = [P, ctypes.c_float]
        L.or_emit_kernel.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
        L.or_write_emits.restype = ctypes.c_uint32
        L.or_write_emits.argtypes = [P, P, P, ctypes.c_uint32, P]
        L.or_emit_now.argtypes = [P, P, P, ctypes.c_uint32, P, ctypes.c_int]
        L.or_step_substep.argtypes = [P, ctypes.c_double, P, P, ctypes.c_uint32, P]
        L.or_frame_tail.argtypes = [P, ctypes.c_double]
        L.or_ripple_frame.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.or_substep_count.argtypes = [P, ctypes.c_double]
        L.or_substep_count.restype = ctypes.c_int
        L.or_rfft_pass.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]
        L.or_rfft2d.restype = P
        L.or_rfft2d.argtypes = [P, P, ctypes.c_float, P]
        L.or_rsrc.argtypes = [P, P]
        L.or_rmask.argtypes = [P, P, P]
        L.or_rextract.argtypes = [P, P]
        L.or_crown.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.c_float, ctypes.c_uint32]
        L.or_droplets.argtypes = [P, ctypes.c_double]
        L.or_droplet_draws.argtypes = [P, P]
        L.or_hashU.restype = ctypes.c_uint32
        L.or_hashU.argtypes = [ctypes.c_uint32]
        L.or_scene_sdf.restype = ctypes.c_float
        L.or_scene_sdf.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
        L.or_scene_normal.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.or_face_open.restype = ctypes.c_float
        L.or_face_open.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]
        _lib = L
    return _lib


def quality(name="standard", **kw):
    q = OrQuality()
    lib().or_quality_default(ctypes.byref(q), name.encode())
    for k, v in kw.items():
        setattr(q, k, v)
    return q


_DT = {"parts": (np.float32, 8), "parts_other": (np.float32, 8), "ww": (np.float32, 8),
       "ww_other": (np.float32, 8), "counters": (np.uint32, 1), "args": (np.uint32, 1),
       "acc": (np.int32, 6), "cellCount": (np.uint32, 1), "vel": (np.float32, 4),
       "velOld": (np.float32, 4), "frac": (np.float32, 4), "svel": (np.float32, 4),
       "cellType": (np.uint32, 1), "divg": (np.float32, 1), "pres": (np.float32, 1),
       "objs": (np.float32, 28), "emits": (np.uint32, 20), "params": (np.uint32, 1),
       "rsrc": (np.int32, 1), "rstate": (np.float32, 2), "ra": (np.float32, 2), "rb": (np.float32, 2),
       "rhat": (np.float32, 2), "tw": (np.float32, 2), "mask": (np.float32, 1), "tex": (np.float32, 4),
       "topH": (np.float32, 1)}


def emit_array(records):
    """records: list of dicts {shape, a, b, c, d, count, seed} -> (OrEmit array, counts, seeds)."""
    n = len(records)
    arr = (OrEmit * max(n, 1))()
    counts = (ctypes.c_uint32 * max(n, 1))()
    seeds = (ctypes.c_uint32 * max(n, 1))()
    for i, r in enumerate(records):
        for key in "abcd":
            v = list(r.get(key) or [0, 0, 0, 0]) + [0, 0, 0, 0]
            getattr(arr[i], key)[:] = [float(x) for x in v[:4]]
        arr[i].info[3] = int(r["shape"])
        counts[i] = int(r["count"])
        seeds[i] = int(r.get("seed", 0)) & 0xFFFFFFFF
    return arr, counts, seeds


class Oracle:
    """One restated FluidSim in host memory (reference buffer layouts)."""

    def __init__(self, tank, q=None, threads=1):
        self.L = lib()
        self.q = q if q is not None else quality()
        self.ptr = self.L.or_create(float(tank[0]), float(tank[1]), float(tank[2]), ctypes.byref(self.q))
        self.L.or_set_threads(self.ptr, int(threads))

    def close(self):
        if self.ptr:
            self.L.or_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = (ctypes.c_double * 26)()
        self.L.or_info(self.ptr, out)
        return dict(zip(INFO_NAMES, list(out)))

    def set(self, name, value):
        self.L.or_set_scalar(self.ptr, name.encode(), float(value))

    def field(self, name):
        """numpy view (no copy) of a named buffer, shaped (n, width)."""
        nb = ctypes.c_size_t(0)
        p = self.L.or_field(self.ptr, name.encode(), ctypes.byref(nb))
        dt, w = _DT[name]
        if not p or nb.value == 0:
            return np.zeros((0, w), dtype=dt)
        isz = np.dtype(dt).itemsize
        buf = (ctypes.c_char * nb.value).from_address(p)
        a = np.frombuffer(buf, dtype=dt, count=nb.value // isz)
        return a.reshape(-1, w) if w > 1 else a

    # driver
    def write_params(self, dt):
        self.L.or_write_params(self.ptr, float(dt))

    def set_objects(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, 28)
        return self.L.or_set_objects(self.ptr, rows.ctypes.data, rows.shape[0])

    def emit_now(self, records, fill=False):
        arr, counts, seeds = emit_array(records)
        self.L.or_emit_now(self.ptr, ctypes.byref(arr), counts, len(records), seeds, int(fill))

    def step_substep(self, dt, records=()):
        arr, counts, seeds = emit_array(list(records))
        self.L.or_step_substep(self.ptr, float(dt), ctypes.byref(arr), counts, len(records), seeds)

    def step(self, frame_dt, nsub=None, emit_fn=None):
        """FluidSim.step: emit_fn(dt, t) -> records; nsub pins the substep count."""
        n = nsub if nsub else self.L.or_substep_count(self.ptr, float(frame_dt))
        dt = frame_dt / n
        self.L.or_step_begin(self.ptr)
        for _ in range(n):
            recs = emit_fn(dt, self.info()["time"]) if emit_fn else []
            self.step_substep(dt, recs)
        self.L.or_frame_tail(self.ptr, float(frame_dt))
        self.L.or_poll_stats(self.ptr)

    def kernel(self, name, *args):
        getattr(self.L, "or_" + name)(self.ptr, *args)
