"""suiri (MI355X) — Python host binding of the liquid-step C ABI (include/suiri.h).

The product is the HIP library libsuiri_amd.so (built in-tree from csrc/ for gfx950). This module
binds it with ctypes and mirrors the reference's JavaScript driver interface (class FluidSim,
src/js/50-sim.js) for Python callers (tests, bench, the batch and serving front-ends). The Node host
(js/fluid_sim.js through the N-API addon) is the drop-in for the reference's own callers.

There is no fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsuiri_amd.so")

SL_OK = 0
SL_EBUDGET = -5
F_PARTICLES, F_WW, F_COUNTERS, F_ACC, F_CELLCOUNT, F_VEL, F_VELOLD, F_FRAC, F_SVEL, F_CELLTYPE, \
    F_DIVG, F_PRES, F_OBJS, F_RSRC, F_RSTATE, F_RHAT, F_RFIELD, F_RMASK, F_RTEX, F_TOPH, F_RTEX16, F_ARGS = range(22)
S_GEOM, S_P2G, S_NORMFORCE, S_DIVERGENCE, S_PRESSURE, S_PROJECT, S_G2P, S_COMPACT, S_WW, S_RIPPLE, \
    S_SOR, S_DROPLETS = range(12)
SOLVER_SOR, SOLVER_MGPCG = 0, 1
T_NAMES = ["emit", "clear", "geom", "p2g", "normforce", "divergence", "sor", "project", "g2p", "compact",
           "ww", "ripple", "droplets"]

FIELD_DTYPE = {F_PARTICLES: (np.float32, 8), F_WW: (np.float32, 8), F_COUNTERS: (np.uint32, 1),
               F_ACC: (np.int32, 6), F_CELLCOUNT: (np.uint32, 1), F_VEL: (np.float32, 4),
               F_VELOLD: (np.float32, 4), F_FRAC: (np.float32, 4), F_SVEL: (np.float32, 4),
               F_CELLTYPE: (np.uint32, 1), F_DIVG: (np.float32, 1), F_PRES: (np.float32, 1),
               F_OBJS: (np.float32, 28), F_RSRC: (np.int32, 1), F_RSTATE: (np.float32, 2),
               F_RHAT: (np.float32, 2), F_RFIELD: (np.float32, 2), F_RMASK: (np.float32, 1),
               F_RTEX: (np.float32, 4), F_TOPH: (np.float32, 1), F_RTEX16: (np.float16, 4),
               F_ARGS: (np.uint32, 1)}


class SlConfig(ctypes.Structure):
    _fields_ = [
        ("tank_w", ctypes.c_double), ("tank_h", ctypes.c_double), ("tank_d", ctypes.c_double),
        ("h", ctypes.c_double), ("cap", ctypes.c_uint32), ("ww_cap", ctypes.c_uint32),
        ("iters", ctypes.c_uint32), ("ripple_R", ctypes.c_uint32), ("rc_factor", ctypes.c_double),
        ("gravity", ctypes.c_float), ("flip", ctypes.c_float), ("ppc", ctypes.c_float),
        ("drift", ctypes.c_float), ("omega", ctypes.c_float), ("friction", ctypes.c_float),
        ("ww_rate", ctypes.c_float), ("r_sigma_rho", ctypes.c_float), ("r_nu", ctypes.c_float),
        ("r_film", ctypes.c_float), ("r_hp_k", ctypes.c_float), ("r_rho_kick", ctypes.c_float),
        ("wmap", ctypes.c_uint32), ("whitewater", ctypes.c_uint32), ("device", ctypes.c_int32),
        ("reorder_every", ctypes.c_uint32), ("reorder_phase", ctypes.c_uint32),
        ("top_height", ctypes.c_uint32), ("seed", ctypes.c_uint32),
        ("splash_rho", ctypes.c_double), ("splash_sigma", ctypes.c_double), ("splash_mu", ctypes.c_double),
        ("splash_k_jet", ctypes.c_float), ("splash_k_crown", ctypes.c_float), ("ripple_depth", ctypes.c_double),
        ("droplets", ctypes.c_uint32), ("drop_seed", ctypes.c_uint32), ("drop_d50", ctypes.c_float),
        ("drop_sigma_ln", ctypes.c_float), ("drop_vmin", ctypes.c_float), ("drop_vmax", ctypes.c_float),
        ("reserved", ctypes.c_uint32 * 2),
    ]


class SlEmit(ctypes.Structure):
    _fields_ = [("a", ctypes.c_float * 4), ("b", ctypes.c_float * 4), ("c", ctypes.c_float * 4),
                ("d", ctypes.c_float * 4), ("shape", ctypes.c_uint32), ("count", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class SlGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nodes", ctypes.c_uint32), ("cells", ctypes.c_uint32), ("cap", ctypes.c_uint32),
                ("ww_cap", ctypes.c_uint32), ("R", ctypes.c_uint32), ("M", ctypes.c_uint32),
                ("iters", ctypes.c_uint32), ("h", ctypes.c_double), ("origin", ctypes.c_double * 3),
                ("rim", ctypes.c_double), ("vp", ctypes.c_double), ("radius", ctypes.c_double)]


class SlStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("ww_count", ctypes.c_uint32), ("impacts", ctypes.c_uint32),
                ("frames", ctypes.c_uint32), ("full", ctypes.c_uint32), ("gpu_errors", ctypes.c_uint32),
                ("substeps_last", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("vmax", ctypes.c_double), ("time", ctypes.c_double), ("liters", ctypes.c_double),
                ("grid", ctypes.c_int32 * 3)]


# every entry point declared in include/suiri.h
EXPORTS = ["sl_config_default", "sl_create", "sl_destroy", "sl_last_error", "sl_grid_info", "sl_set_objects",
           "sl_emit_now", "sl_step_begin", "sl_substep", "sl_step_end", "sl_step", "sl_set_substeps",
           "sl_clear_water", "sl_save_water", "sl_has_water_copy", "sl_restore_water", "sl_poll_stats",
           "sl_stats", "sl_sync", "sl_set_ripple_mask", "sl_field_bytes", "sl_upload", "sl_readback",
           "sl_set_scalar", "sl_get_scalar", "sl_run_stage", "sl_timing_enable", "sl_timing_read",
           "sl_version", "sl_build_info", "sl_reorder", "sl_autotune", "sl_set_pressure_solver",
           "sl_pressure_solver_stats",
           "sl_batch_create", "sl_batch_destroy", "sl_batch_last_error", "sl_batch_size", "sl_batch_sim",
           "sl_batch_step", "sl_batch_capture", "sl_batch_sync", "sl_batch_reorder", "sl_batch_autotune",
           "sl_batch_set_variants", "sl_batch_timing_enable", "sl_batch_timing_read", "sl_batch_open",
           "sl_batch_add", "sl_batch_remove", "sl_batch_set_stepped", "sl_batch_emit", "sl_batch_captured",
           "sl_batch_set_dt", "sl_job_bytes", "sl_server_config_default", "sl_server_create", "sl_server_destroy",
           "sl_server_last_error", "sl_server_submit", "sl_server_cancel", "sl_server_tick", "sl_server_poll",
           "sl_server_info", "sl_server_job_sim", "sl_server_release", "sl_server_stats", "sl_bake_frame",
           "sl_write_exr_rgb", "sl_scene_create", "sl_scene_destroy", "sl_scene_last_error", "sl_sim_scene",
           "sl_scene_pack_objects", "sl_set_items", "sl_scene_source_emits", "sl_scene_block_emits",
           "sl_block_liters", "sl_step_items", "sl_ripple_frame_noalloc", "sl_wait_idle"]

EMIT_FN = ctypes.CFUNCTYPE(ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                           ctypes.POINTER(SlEmit), ctypes.c_uint32)


class SlServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_jobs", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("autotune_after", ctypes.c_uint32), ("graph_after", ctypes.c_uint32),
                ("frames_per_tick", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class SlJob(ctypes.Structure):
    _fields_ = [("cfg", SlConfig), ("objects", ctypes.POINTER(ctypes.c_float)), ("n_objects", ctypes.c_uint32),
                ("fill", ctypes.POINTER(SlEmit)), ("n_fill", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("dt", ctypes.c_double), ("source", EMIT_FN), ("user", ctypes.c_void_p), ("keep", ctypes.c_uint32),
                ("bake_every", ctypes.c_uint32), ("bake_dir", ctypes.c_char_p), ("items", ctypes.c_void_p),
                ("n_items", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class SlJobInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("frames", ctypes.c_uint32),
                ("steps", ctypes.c_uint32), ("slot", ctypes.c_uint32), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("particle_steps", ctypes.c_double),
                ("bytes", ctypes.c_uint64), ("count", ctypes.c_uint32), ("ww_count", ctypes.c_uint32)]


class SlServerStats(ctypes.Structure):
    _fields_ = [("budget", ctypes.c_uint64), ("used", ctypes.c_uint64), ("queued", ctypes.c_uint32),
                ("running", ctypes.c_uint32), ("resident", ctypes.c_uint32), ("done", ctypes.c_uint32),
                ("ticks", ctypes.c_uint64), ("frames", ctypes.c_uint64), ("particle_steps", ctypes.c_double),
                ("uptime", ctypes.c_double), ("step_ms_p50", ctypes.c_double), ("step_ms_p99", ctypes.c_double),
                ("step_ms_sum", ctypes.c_double), ("latency_p50", ctypes.c_double), ("latency_p99", ctypes.c_double)]


JOB_STATUS = {0: "queued", 1: "running", 2: "done", 3: "cancelled", 4: "refused"}

_lib = None


class SuiriError(RuntimeError):
    pass


def load():
    """Load libsuiri_amd.so (raises when it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SuiriError(f"suiri HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P, I, U, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_double
    sig = {
        "sl_config_default": (I, [ctypes.POINTER(SlConfig), ctypes.c_char_p]),
        "sl_create": (I, [ctypes.POINTER(SlConfig), ctypes.POINTER(P)]),
        "sl_destroy": (None, [P]),
        "sl_last_error": (ctypes.c_char_p, [P]),
        "sl_grid_info": (I, [P, ctypes.POINTER(SlGrid)]),
        "sl_set_objects": (I, [P, P, U, ctypes.POINTER(I)]),
        "sl_emit_now": (I, [P, P, U, I, ctypes.POINTER(U)]),
        "sl_step_begin": (I, [P, D, ctypes.POINTER(U), ctypes.POINTER(D)]),
        "sl_substep": (I, [P, D, P, U, ctypes.POINTER(U)]),
        "sl_step_end": (I, [P, D]),
        "sl_step": (I, [P, D, P, P]),
        "sl_set_substeps": (I, [P, U]),
        "sl_set_pressure_solver": (I, [P, I, D, U]),
        "sl_pressure_solver_stats": (I, [P, ctypes.POINTER(U), ctypes.POINTER(I), ctypes.POINTER(D),
                                         ctypes.POINTER(D)]),
        "sl_clear_water": (I, [P]),
        "sl_reorder": (I, [P]),
        "sl_autotune": (I, [P, U, ctypes.POINTER(U)]),
        "sl_save_water": (I, [P, ctypes.POINTER(U)]),
        "sl_has_water_copy": (I, [P, U]),
        "sl_restore_water": (I, [P, U]),
        "sl_poll_stats": (I, [P, ctypes.POINTER(I)]),
        "sl_stats": (I, [P, ctypes.POINTER(SlStats)]),
        "sl_sync": (I, [P]),
        "sl_wait_idle": (I, [P]),
        "sl_set_ripple_mask": (I, [P, I]),
        "sl_field_bytes": (ctypes.c_size_t, [P, I]),
        "sl_upload": (I, [P, I, P, ctypes.c_size_t]),
        "sl_readback": (I, [P, I, P, ctypes.c_size_t]),
        "sl_set_scalar": (I, [P, ctypes.c_char_p, D]),
        "sl_get_scalar": (I, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
        "sl_run_stage": (I, [P, I, U, D]),
        "sl_timing_enable": (I, [P, I]),
        "sl_timing_read": (I, [P, ctypes.POINTER(D), ctypes.POINTER(U), I, I]),
        "sl_batch_create": (I, [ctypes.POINTER(SlConfig), U, ctypes.POINTER(P)]),
        "sl_batch_destroy": (None, [P]),
        "sl_batch_last_error": (ctypes.c_char_p, [P]),
        "sl_batch_size": (U, [P]),
        "sl_batch_sim": (P, [P, U]),
        "sl_batch_step": (I, [P, D]),
        "sl_batch_capture": (I, [P, D]),
        "sl_batch_sync": (I, [P]),
        "sl_batch_reorder": (I, [P]),
        "sl_batch_autotune": (I, [P, U, ctypes.POINTER(U)]),
        "sl_batch_set_variants": (I, [P, ctypes.POINTER(U)]),
        "sl_batch_timing_enable": (I, [P, I]),
        "sl_batch_open": (I, [U, I, ctypes.POINTER(P)]),
        "sl_batch_add": (I, [P, ctypes.POINTER(SlConfig), ctypes.POINTER(U)]),
        "sl_batch_remove": (I, [P, U]),
        "sl_batch_set_stepped": (I, [P, U, I]),
        "sl_batch_emit": (I, [P, P, ctypes.POINTER(U)]),
        "sl_batch_captured": (I, [P]),
        "sl_batch_set_dt": (I, [P, U, D]),
        "sl_job_bytes": (ctypes.c_size_t, [ctypes.POINTER(SlConfig)]),
        "sl_server_config_default": (I, [ctypes.POINTER(SlServerConfig)]),
        "sl_server_create": (I, [ctypes.POINTER(SlServerConfig), ctypes.POINTER(P)]),
        "sl_server_destroy": (None, [P]),
        "sl_server_last_error": (ctypes.c_char_p, [P]),
        "sl_server_submit": (I, [P, ctypes.POINTER(SlJob), ctypes.POINTER(ctypes.c_uint64)]),
        "sl_server_cancel": (I, [P, ctypes.c_uint64]),
        "sl_server_tick": (I, [P, ctypes.POINTER(U)]),
        "sl_server_poll": (I, [P, ctypes.POINTER(SlJobInfo), U, ctypes.POINTER(U)]),
        "sl_server_info": (I, [P, ctypes.c_uint64, ctypes.POINTER(SlJobInfo)]),
        "sl_server_job_sim": (P, [P, ctypes.c_uint64]),
        "sl_server_release": (I, [P, ctypes.c_uint64]),
        "sl_server_stats": (I, [P, ctypes.POINTER(SlServerStats)]),
        "sl_bake_frame": (I, [P, ctypes.c_char_p, U]),
        "sl_write_exr_rgb": (I, [ctypes.c_char_p, U, U, P, P, P]),
        "sl_batch_timing_read": (I, [P, ctypes.POINTER(D), U, ctypes.POINTER(U)]),
        "sl_scene_create": (I, [ctypes.POINTER(SlConfig), ctypes.POINTER(P)]),
        "sl_scene_destroy": (None, [P]),
        "sl_scene_last_error": (ctypes.c_char_p, [P]),
        "sl_sim_scene": (P, [P]),
        "sl_scene_pack_objects": (I, [P, P, U, D, P, ctypes.POINTER(U), ctypes.POINTER(I)]),
        "sl_set_items": (I, [P, P, U, D, ctypes.POINTER(I)]),
        "sl_scene_source_emits": (I, [P, P, U, D, P, U, ctypes.POINTER(U)]),
        "sl_scene_block_emits": (I, [P, P, U, I, P, U, ctypes.POINTER(U)]),
        "sl_block_liters": (D, [P, D, D, D]),
        "sl_step_items": (I, [P, D, P, U]),
        "sl_ripple_frame_noalloc": (I, [ctypes.POINTER(SlConfig), U, D, D, D, D, P, P, P, P]),
        "sl_version": (I, []),
        "sl_build_info": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def config(quality="standard", **kw):
    c = SlConfig()
    L = load()
    r = L.sl_config_default(ctypes.byref(c), quality.encode())
    if r != SL_OK:
        raise SuiriError(L.sl_last_error(None).decode())
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def emit_records(records):
    n = len(records)
    arr = (SlEmit * max(n, 1))()
    for i, r in enumerate(records):
        for key in "abcd":
            v = list(r.get(key) or [0, 0, 0, 0]) + [0, 0, 0, 0]
            getattr(arr[i], key)[:] = [float(x) for x in v[:4]]
        arr[i].shape = int(r["shape"])
        arr[i].count = int(r["count"])
        arr[i].seed = int(r.get("seed", 0)) & 0xFFFFFFFF
    return arr


class Sim:
    """One simulation on the GPU through the C ABI (mirrors FluidSim, src/js/50-sim.js:28-376)."""

    def __init__(self, cfg, _handle=None, _owner=None):
        self.L = load()
        self.ptr = ctypes.c_void_p()
        self.cfg = cfg
        self._owner = _owner  # a Batch owns its tanks: they are destroyed with it
        if _handle is not None:
            self.ptr = ctypes.c_void_p(_handle)
        else:
            self._check(self.L.sl_create(ctypes.byref(cfg), ctypes.byref(self.ptr)), None)
        g = SlGrid()
        self._check(self.L.sl_grid_info(self.ptr, ctypes.byref(g)))
        self.grid = g
        self.nx, self.ny, self.nz, self.h = g.nx, g.ny, g.nz, g.h
        self.origin = list(g.origin)
        self.vp, self.rim, self.cap = g.vp, g.rim, g.cap

    def _check(self, r, ptr=0):
        if r != SL_OK:
            msg = self.L.sl_last_error(self.ptr if ptr is not None else None)
            raise SuiriError(f"suiri error {r}: {msg.decode() if msg else ''}")

    def close(self):
        if self.ptr and self._owner is None:
            self.L.sl_destroy(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FluidSim interface
    def set_items(self, items, dt_move=0.0):
        """FluidSim.setObjects(items, dtMove): packs the obstacles and drains among the scene items
        (velocities from the previous call's poses) and uploads them; returns whether one moved."""
        from . import scene as _scene
        arr, n = _scene.items_array(items)
        mv = ctypes.c_int(0)
        self._check(self.L.sl_set_items(self.ptr, arr, n, float(dt_move), ctypes.byref(mv)))
        return bool(mv.value)

    def scene(self):
        """The simulation's own scene state (sources' progress, object poses, seed stream)."""
        from . import scene as _scene
        return _scene.Scene.of(self)

    def fill_blocks(self, items, fill=False):
        """emitNow(blockEmits(items, fill), fill): the water blocks among the scene items."""
        recs = self.scene().block_emits(items, fill)
        return self.emit_now(recs, fill) if recs else 0

    def step_items(self, frame_dt, items):
        """FluidSim.step with the sources among the scene items as the emitter (sl_step_items)."""
        from . import scene as _scene
        arr, n = _scene.items_array(items)
        self._check(self.L.sl_step_items(self.ptr, float(frame_dt), arr, n))

    def set_objects_rows(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, 28)
        ch = ctypes.c_int(0)
        self._check(self.L.sl_set_objects(self.ptr, rows.ctypes.data, rows.shape[0], ctypes.byref(ch)))
        return bool(ch.value)

    def emit_now(self, records, fill=False):
        arr = emit_records(records)
        out = ctypes.c_uint32(0)
        self._check(self.L.sl_emit_now(self.ptr, ctypes.byref(arr), len(records), int(fill), ctypes.byref(out)))
        return out.value

    def set_substeps(self, n):
        self._check(self.L.sl_set_substeps(self.ptr, int(n)))

    def set_pressure_solver(self, solver, tol_per_s=1e-5, max_iters=40):
        """SOLVER_SOR: the reference's red-black SOR sweeps (default). SOLVER_MGPCG: multigrid-
        preconditioned CG on the same system until the largest residual is at most tol_per_s (1/s)."""
        self._check(self.L.sl_set_pressure_solver(self.ptr, int(solver), float(tol_per_s), int(max_iters)))

    def pressure_solver_stats(self):
        it, br = ctypes.c_uint32(0), ctypes.c_int(0)
        r, r0 = ctypes.c_double(0), ctypes.c_double(0)
        self._check(self.L.sl_pressure_solver_stats(self.ptr, ctypes.byref(it), ctypes.byref(br), ctypes.byref(r),
                                                    ctypes.byref(r0)))
        return {"iters": it.value, "breakdown": bool(br.value), "resid_per_s": r.value, "resid0_per_s": r0.value}

    def step(self, frame_dt, emit_fn=None):
        n = ctypes.c_uint32(0)
        dt = ctypes.c_double(0)
        self._check(self.L.sl_step_begin(self.ptr, float(frame_dt), ctypes.byref(n), ctypes.byref(dt)))
        for _ in range(n.value):
            recs = emit_fn(dt.value, self.scalar("time")) if emit_fn else []
            arr = emit_records(recs)
            self._check(self.L.sl_substep(self.ptr, dt.value, ctypes.byref(arr), len(recs), None))
        self._check(self.L.sl_step_end(self.ptr, float(frame_dt)))
        return n.value

    def autotune(self, reps=3):
        ch = (ctypes.c_uint32 * 2)()
        self._check(self.L.sl_autotune(self.ptr, int(reps), ch))
        return {"p2g_variant": ch[0], "g2p_variant": ch[1]}

    def reorder(self):
        self._check(self.L.sl_reorder(self.ptr))

    def bake(self, directory, frame):
        """frame_%05d.npz (+ surface_%05d.exr with a ripple layer) of the current state into `directory`."""
        self._check(self.L.sl_bake_frame(self.ptr, os.fsencode(directory), int(frame)))

    def clear_water(self):
        self._check(self.L.sl_clear_water(self.ptr))

    def save_water(self):
        i = ctypes.c_uint32(0)
        self._check(self.L.sl_save_water(self.ptr, ctypes.byref(i)))
        return i.value

    def has_water_copy(self, i):
        return bool(self.L.sl_has_water_copy(self.ptr, int(i)))

    def restore_water(self, i):
        return bool(self.L.sl_restore_water(self.ptr, int(i)))

    def poll_stats(self):
        u = ctypes.c_int(0)
        self._check(self.L.sl_poll_stats(self.ptr, ctypes.byref(u)))
        return bool(u.value)

    def stats(self):
        st = SlStats()
        self._check(self.L.sl_stats(self.ptr, ctypes.byref(st)))
        return {"count": st.count, "ww": st.ww_count, "impacts": st.impacts, "frames": st.frames,
                "full": bool(st.full), "gpuErrors": st.gpu_errors, "vmax": st.vmax, "time": st.time,
                "liters": st.liters, "grid": list(st.grid), "substeps": st.substeps_last}

    def sync(self):
        self._check(self.L.sl_sync(self.ptr))

    def set_ripple_mask(self, bound):
        self._check(self.L.sl_set_ripple_mask(self.ptr, int(bound)))

    def scalar(self, name):
        v = ctypes.c_double(0)
        self._check(self.L.sl_get_scalar(self.ptr, name.encode(), ctypes.byref(v)))
        return v.value

    def set_scalar(self, name, value):
        self._check(self.L.sl_set_scalar(self.ptr, name.encode(), float(value)))

    # state transfer (reference layouts)
    def readback(self, field):
        nb = self.L.sl_field_bytes(self.ptr, field)
        dt, w = FIELD_DTYPE[field]
        out = np.empty(nb // np.dtype(dt).itemsize, dtype=dt)
        self._check(self.L.sl_readback(self.ptr, field, out.ctypes.data, nb))
        return out.reshape(-1, w) if w > 1 else out

    def upload(self, field, arr):
        nb = self.L.sl_field_bytes(self.ptr, field)
        a = np.ascontiguousarray(arr)
        if a.nbytes < nb:
            raise SuiriError(f"upload of field {field}: {a.nbytes} bytes < {nb}")
        self._check(self.L.sl_upload(self.ptr, field, a.ctypes.data, nb))

    def run_stage(self, stage, arg=0, dt=1.0 / 60.0):
        self._check(self.L.sl_run_stage(self.ptr, int(stage), int(arg), float(dt)))

    # timing
    def timing(self, enable=True):
        self._check(self.L.sl_timing_enable(self.ptr, int(enable)))

    def timing_read(self, reset=True):
        ms = (ctypes.c_double * len(T_NAMES))()
        ln = (ctypes.c_uint32 * len(T_NAMES))()
        self._check(self.L.sl_timing_read(self.ptr, ms, ln, len(T_NAMES), int(reset)))
        return {n: (ms[i], ln[i]) for i, n in enumerate(T_NAMES)}


class Batch:
    """A table of tank slots stepped together, one kernel launch per stage for all stepped tanks
    (sl_batch_*). `Batch(cfgs)` fills one slot per configuration; `Batch(capacity=n)` starts empty.
    Per-tank set-up, single emissions and readback go through `tanks[slot]`, a Sim on the batch's
    stream (None for an empty slot)."""

    def __init__(self, cfgs=None, capacity=None, device=0):
        self.L = load()
        self.ptr = ctypes.c_void_p()
        if cfgs is not None:
            arr = (SlConfig * max(len(cfgs), 1))(*cfgs)
            r = self.L.sl_batch_create(arr, len(cfgs), ctypes.byref(self.ptr))
        else:
            r = self.L.sl_batch_open(int(capacity), int(device), ctypes.byref(self.ptr))
        if r != SL_OK:
            msg = self.L.sl_batch_last_error(None)
            raise SuiriError(f"suiri error {r}: {msg.decode() if msg else ''}")
        self.capacity = self.L.sl_batch_size(self.ptr)
        self.tanks = [None] * self.capacity
        for i in range(self.capacity):
            h = self.L.sl_batch_sim(self.ptr, i)
            if h:
                self.tanks[i] = Sim(cfgs[i] if cfgs is not None else None, _handle=h, _owner=self)

    def _check(self, r):
        if r != SL_OK:
            msg = self.L.sl_batch_last_error(self.ptr)
            raise SuiriError(f"suiri error {r}: {msg.decode() if msg else ''}")

    def __len__(self):
        return self.capacity

    def add(self, cfg):
        slot = ctypes.c_uint32(0)
        self._check(self.L.sl_batch_add(self.ptr, ctypes.byref(cfg), ctypes.byref(slot)))
        self.tanks[slot.value] = Sim(cfg, _handle=self.L.sl_batch_sim(self.ptr, slot.value), _owner=self)
        return slot.value

    def remove(self, slot):
        t = self.tanks[slot]
        if t is not None:
            t.ptr = ctypes.c_void_p()
        self.tanks[slot] = None
        self._check(self.L.sl_batch_remove(self.ptr, int(slot)))

    def set_stepped(self, slot, on):
        self._check(self.L.sl_batch_set_stepped(self.ptr, int(slot), int(bool(on))))

    def emit(self, per_slot):
        """per_slot: {slot: [emission records]} for one emission into many tanks."""
        counts = (ctypes.c_uint32 * self.capacity)()
        recs = []
        for slot in sorted(per_slot):
            counts[slot] = len(per_slot[slot])
            recs.extend(per_slot[slot])
        arr = emit_records(recs)
        self._check(self.L.sl_batch_emit(self.ptr, ctypes.byref(arr), counts))

    def step(self, dt):
        self._check(self.L.sl_batch_step(self.ptr, float(dt)))

    def capture(self, dt):
        self._check(self.L.sl_batch_capture(self.ptr, float(dt)))

    @property
    def captured(self):
        return bool(self.L.sl_batch_captured(self.ptr))

    def sync(self):
        self._check(self.L.sl_batch_sync(self.ptr))

    def reorder(self):
        self._check(self.L.sl_batch_reorder(self.ptr))

    def autotune(self, reps=2):
        ch = (ctypes.c_uint32 * 6)()
        self._check(self.L.sl_batch_autotune(self.ptr, int(reps), ch))
        return [(ch[2 * k], ch[2 * k + 1]) for k in range(3)]

    def set_variants(self, pairs):
        ch = (ctypes.c_uint32 * 6)(*[int(v) for pr in pairs for v in pr])
        self._check(self.L.sl_batch_set_variants(self.ptr, ch))

    def timing(self, enable=True):
        self._check(self.L.sl_batch_timing_enable(self.ptr, int(enable)))

    def step_times(self, max_steps=100000):
        """Device time [ms] of each step since timing() or the last call, oldest first."""
        buf = (ctypes.c_double * max_steps)()
        n = ctypes.c_uint32(0)
        self._check(self.L.sl_batch_timing_read(self.ptr, buf, max_steps, ctypes.byref(n)))
        return np.array(buf[:n.value])

    def close(self):
        for t in getattr(self, "tanks", []):
            if t is not None:
                t.ptr = ctypes.c_void_p()
        if self.ptr:
            self.L.sl_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ripple_frame_noalloc(cfg, R, Lx, Lz, dt, depth, state_ptr, src_ptr, tex_ptr, stream=None):
    """One unmasked ripple frame on caller-owned device buffers (addresses), allocating nothing
    (sl_ripple_frame_noalloc)."""
    L = load()
    r = L.sl_ripple_frame_noalloc(ctypes.byref(cfg), int(R), float(Lx), float(Lz), float(dt), float(depth),
                                  ctypes.c_void_p(state_ptr), ctypes.c_void_p(src_ptr), ctypes.c_void_p(tex_ptr),
                                  ctypes.c_void_p(stream) if stream else None)
    if r != SL_OK:
        raise SuiriError(f"suiri error {r}: {L.sl_last_error(None).decode()}")


def job_bytes(cfg):
    """Device bytes a simulation of this configuration holds (host-side computation)."""
    return int(load().sl_job_bytes(ctypes.byref(cfg)))


def _info_dict(i):
    return {"id": i.id, "status": JOB_STATUS.get(i.status, i.status), "frames": i.frames, "steps": i.steps,
            "t_submit": i.t_submit, "t_admit": i.t_admit, "t_done": i.t_done, "particle_steps": i.particle_steps,
            "bytes": i.bytes, "count": i.count, "ww_count": i.ww_count}


class Server:
    """The serving loop (sl_server_*): jobs admitted against a device-memory budget, stepped together
    as a continuous batch, one frame per tick (up to `frames_per_tick` back to back while no job
    finishes or bakes). `source(dt, t)` returns a job's per-frame emission records (like the emitter
    callback of Sim.step)."""

    def __init__(self, budget=200 << 30, max_jobs=128, device=0, autotune_after=0, graph_after=0,
                 frames_per_tick=4):
        self.L = load()
        c = SlServerConfig()
        self.L.sl_server_config_default(ctypes.byref(c))
        c.hbm_budget, c.max_jobs, c.device = int(budget), int(max_jobs), int(device)
        c.autotune_after, c.graph_after = int(autotune_after), int(graph_after)
        c.frames_per_tick = int(frames_per_tick)
        self.ptr = ctypes.c_void_p()
        r = self.L.sl_server_create(ctypes.byref(c), ctypes.byref(self.ptr))
        if r != SL_OK:
            raise SuiriError(f"suiri error {r}: sl_server_create")
        self._keep = {}   # per job: ctypes objects the library points at (callbacks)

    def _check(self, r):
        if r != SL_OK:
            msg = self.L.sl_server_last_error(self.ptr)
            raise SuiriError(f"suiri error {r}: {msg.decode() if msg else ''}")

    def submit(self, cfg, rows=None, fill=(), steps=100, dt=1.0 / 120, source=None, keep=False, bake_dir=None,
               bake_every=0, items=None):
        """`items`: scene items whose obstacles and drains are packed and whose blocks fill the tank at
        admission, and whose sources emit every frame (instead of `rows`)."""
        j = SlJob()
        j.cfg = cfg
        items_keep = None
        if items:
            from . import scene as _scene
            items_keep, n_items = _scene.items_array(items)
            j.items = ctypes.cast(items_keep, ctypes.c_void_p)
            j.n_items = n_items
        rows_a = np.ascontiguousarray(rows if rows is not None else np.zeros((0, 28)), dtype=np.float32).reshape(-1, 28)
        j.objects = rows_a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        j.n_objects = rows_a.shape[0]
        fill_a = emit_records(list(fill))
        j.fill = ctypes.cast(fill_a, ctypes.POINTER(SlEmit))
        j.n_fill = len(fill)
        j.steps, j.dt, j.keep = int(steps), float(dt), int(bool(keep))
        j.bake_every = int(bake_every)
        j.bake_dir = os.fsencode(bake_dir) if bake_dir else None
        cb = None
        if source is not None:
            def _cb(user, sim, d, t, out, mx):
                recs = source(d, t)[:mx]
                arr = emit_records(recs)
                for k in range(len(recs)):
                    out[k] = arr[k]
                return len(recs)
            cb = EMIT_FN(_cb)
            j.source = cb
        jid = ctypes.c_uint64(0)
        r = self.L.sl_server_submit(self.ptr, ctypes.byref(j), ctypes.byref(jid))
        if cb is not None:
            self._keep[jid.value] = cb
        if r == SL_EBUDGET:
            return jid.value   # refused: reported by poll() / info()
        self._check(r)
        return jid.value

    def cancel(self, jid):
        self._check(self.L.sl_server_cancel(self.ptr, ctypes.c_uint64(jid)))

    def tick(self):
        n = ctypes.c_uint32(0)
        self._check(self.L.sl_server_tick(self.ptr, ctypes.byref(n)))
        return n.value

    def poll(self):
        buf = (SlJobInfo * 256)()
        out = []
        while True:
            n = ctypes.c_uint32(0)
            self._check(self.L.sl_server_poll(self.ptr, buf, 256, ctypes.byref(n)))
            out.extend(_info_dict(buf[i]) for i in range(n.value))
            for i in range(n.value):
                self._keep.pop(buf[i].id, None)
            if n.value < 256:
                return out

    def info(self, jid):
        i = SlJobInfo()
        self._check(self.L.sl_server_info(self.ptr, ctypes.c_uint64(jid), ctypes.byref(i)))
        return _info_dict(i)

    def job_sim(self, jid):
        h = self.L.sl_server_job_sim(self.ptr, ctypes.c_uint64(jid))
        return Sim(None, _handle=h, _owner=self) if h else None

    def release(self, jid):
        self._check(self.L.sl_server_release(self.ptr, ctypes.c_uint64(jid)))

    def stats(self):
        st = SlServerStats()
        self._check(self.L.sl_server_stats(self.ptr, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in SlServerStats._fields_}

    def close(self):
        if self.ptr:
            self.L.sl_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_exr_rgb(path, r, g, b):
    """The bake maps' EXR layout: three H x W float32 planes, row 0 at the top of the picture."""
    r, g, b = (np.ascontiguousarray(x, dtype=np.float32) for x in (r, g, b))
    H, W = r.shape
    rc = load().sl_write_exr_rgb(os.fsencode(path), W, H, r.ctypes.data, g.ctypes.data, b.ctypes.data)
    if rc != SL_OK:
        raise SuiriError(f"write_exr_rgb({path}) failed")
