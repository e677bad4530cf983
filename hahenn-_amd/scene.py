"""Scene items for Python callers, over the library's scene functions.

Items are dicts in the reference's scene model (src/js/10-scene.js:17-87): obstacles, drains, water
sources and water blocks. Object packing (FluidSim.setObjects, src/js/50-sim.js:165-211) and the
emitters (src/js/55-emit.js:13-113) run in the library (csrc/scene.cpp, `sl_scene_*`), the same code
that the Node host and the serving loop call; this module only converts items to `sl_item` records.
"""
import ctypes
import importlib

import numpy as np

suiri = importlib.import_module("hahenn-_amd")

MAX_OBJ = 64

KINDS = {"obstacle": 0, "drain": 1, "source": 2, "block": 3}
# obstacle shapes: id and the order of their dimensions in sl_item.dims (include/suiri.h)
SHAPES = {
    "box": (0, ("w", "h", "d")), "plate": (1, ("w", "d", "t")), "sphere": (2, ("r",)),
    "cylinder": (3, ("r", "h")), "cone": (4, ("r", "h")), "torus": (5, ("R", "r")),
    "cup": (6, ("r", "h", "t")), "container": (7, ("w", "h", "d", "t")), "gutter": (8, ("len", "r", "t")),
}
# default dimensions of a new obstacle of each shape (10-scene.js:5-15)
SHAPE_DIMS = {
    "box": {"w": 0.10, "h": 0.06, "d": 0.08}, "plate": {"w": 0.22, "d": 0.10, "t": 0.012},
    "sphere": {"r": 0.035}, "cylinder": {"r": 0.025, "h": 0.12}, "cone": {"r": 0.04, "h": 0.08},
    "torus": {"R": 0.05, "r": 0.014}, "cup": {"r": 0.04, "h": 0.10, "t": 0.006},
    "container": {"w": 0.14, "h": 0.06, "d": 0.10, "t": 0.008}, "gutter": {"len": 0.28, "r": 0.03, "t": 0.006},
}
MODES = {"stream": 0, "drops": 1, "shower": 2, "burst": 3, "rain": 4}
# source fields of the scene schema -> sl_item fields
SOURCE_FIELDS = {"pitch": "pitch", "yaw": "yaw", "flow": "flow", "speed": "speed", "dropD": "drop_d",
                 "dropRate": "drop_rate", "spread": "spread", "burst": "burst", "burstEvery": "burst_every",
                 "rainRate": "rain_rate", "rainSize": "rain_size", "total": "total"}


class SlItem(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("kind", ctypes.c_uint32), ("shape", ctypes.c_uint32),
                ("enabled", ctypes.c_uint32), ("pos", ctypes.c_double * 3), ("rot", ctypes.c_double * 3),
                ("dims", ctypes.c_double * 4)] + [(f, ctypes.c_double) for f in SOURCE_FIELDS.values()]


# ---- item factories (the scene model, 10-scene.js:56-87) ----
_uid = [1]


def _next_id():
    _uid[0] += 1
    return _uid[0] - 1


def make_obstacle(shape, pos=None, dims=None, rot=None):
    return {"id": _next_id(), "kind": "obstacle", "shape": shape, "pos": list(pos or [0, 0.05, 0]),
            "rot": list(rot or [0, 0, 0]), "dims": dict(dims or SHAPE_DIMS[shape])}


def make_source(pos=None, **kw):
    s = {"id": _next_id(), "kind": "source", "pos": list(pos or [0, 0.4, 0]), "pitch": -90, "yaw": 0,
         "mode": "stream", "enabled": True, "flow": 3.0, "speed": 0.6, "dropD": 5, "dropRate": 3,
         "spread": 18, "burst": 0.5, "burstEvery": 0, "rainRate": 60, "rainSize": 0.3, "total": 0}
    s.update(kw)
    return s


def make_block(pos=None, dims=None):
    return {"id": _next_id(), "kind": "block", "pos": list(pos or [0, 0.05, 0]),
            "dims": dict(dims or {"w": 0.16, "h": 0.10, "d": 0.16})}


def make_drain(pos=None, r=0.025, rot=None):
    return {"id": _next_id(), "kind": "drain", "pos": list(pos or [0, 0, 0]), "rot": list(rot or [0, 0, 0]), "r": r}


def item_record(it):
    """One scene item as an sl_item."""
    r = SlItem()
    r.id = int(it["id"]) & 0xFFFFFFFF
    kind = it["kind"]
    r.kind = KINDS[kind]
    r.pos[:] = [float(x) for x in it["pos"]]
    if kind in ("obstacle", "drain"):
        r.rot[:] = [float(x) for x in it.get("rot", (0, 0, 0))]
    if kind == "obstacle":
        sid, order = SHAPES[it["shape"]]
        r.shape = sid
        d = [float(it["dims"][k]) for k in order] + [0.0] * 4
        r.dims[:] = d[:4]
    elif kind == "block":
        r.dims[:] = [float(it["dims"][k]) for k in ("w", "h", "d")] + [0.0]
    elif kind == "drain":
        r.dims[:] = [float(it["r"]), 0.0, 0.0, 0.0]
    else:
        r.shape = MODES[it["mode"]]
        r.enabled = 1 if it.get("enabled", True) else 0
        for k, f in SOURCE_FIELDS.items():
            setattr(r, f, float(it.get(k, 0)))
    return r


def items_array(items):
    items = list(items)
    arr = (SlItem * max(len(items), 1))()
    for i, it in enumerate(items):
        arr[i] = item_record(it)
    return arr, len(items)


def _record_dict(e):
    return {"shape": int(e.shape), "a": list(e.a), "b": list(e.b), "c": list(e.c), "d": list(e.d),
            "count": int(e.count), "seed": int(e.seed)}


class Scene:
    """Host-side scene state (sl_scene): source progress, previous object poses and the seed stream
    of a tank of configuration `cfg` (default: the standard preset). `Scene.of(sim)` is the state a
    simulation keeps itself (saved and restored with its water)."""

    def __init__(self, cfg=None, _ptr=None, _owner=None):
        self.L = suiri.load()
        self._owner = _owner
        if _ptr is not None:
            self.ptr = ctypes.c_void_p(_ptr)
            return
        cfg = cfg if cfg is not None else suiri.config("standard")
        self.ptr = ctypes.c_void_p()
        r = self.L.sl_scene_create(ctypes.byref(cfg), ctypes.byref(self.ptr))
        if r != suiri.SL_OK:
            raise suiri.SuiriError(f"sl_scene_create: {r}")

    @classmethod
    def of(cls, sim):
        return cls(_ptr=sim.L.sl_sim_scene(sim.ptr), _owner=sim)

    def close(self):
        if self.ptr and self._owner is None:
            self.L.sl_scene_destroy(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, r):
        if r != suiri.SL_OK:
            raise suiri.SuiriError(f"scene error {r}: {self.L.sl_scene_last_error(self.ptr).decode()}")

    def pack(self, items, dt_move=0.0):
        """Obj records (n x 28 f32) of the obstacles and drains among `items`, velocities from the
        previous call's poses; returns (rows, moving)."""
        arr, n = items_array(items)
        rows = np.zeros((MAX_OBJ, 28), dtype=np.float32)
        k, mv = ctypes.c_uint32(0), ctypes.c_int(0)
        self._check(self.L.sl_scene_pack_objects(self.ptr, arr, n, float(dt_move), rows.ctypes.data,
                                                 ctypes.byref(k), ctypes.byref(mv)))
        return rows[: k.value].copy(), bool(mv.value)

    def source_emits(self, items, dt):
        """Emission records (dicts) of the sources among `items` for one substep of length dt."""
        arr, n = items_array(items)
        out = (suiri.SlEmit * 64)()
        k = ctypes.c_uint32(0)
        self._check(self.L.sl_scene_source_emits(self.ptr, arr, n, float(dt), out, 64, ctypes.byref(k)))
        return [_record_dict(out[i]) for i in range(k.value)]

    def block_emits(self, items, fill=False):
        """Lattice records (dicts) of the blocks among `items`."""
        arr, n = items_array(items)
        out = (suiri.SlEmit * max(n, 1))()
        k = ctypes.c_uint32(0)
        self._check(self.L.sl_scene_block_emits(self.ptr, arr, n, int(bool(fill)), out, max(n, 1), ctypes.byref(k)))
        return [_record_dict(out[i]) for i in range(k.value)]


def block_liters(block, tank):
    """Litres of a block inside a tank (w, h, d)."""
    return float(suiri.load().sl_block_liters(ctypes.byref(item_record(block)), *[float(x) for x in tank]))


class Emitters:
    """The emitter callback of a step: `Emitters(items, scene)(dt, t)` -> the records of the sources
    among `items` for one substep (sl_scene_source_emits with the given scene state)."""

    def __init__(self, items, scene):
        self.items, self.scene = list(items), scene

    def __call__(self, dt, t):
        return self.scene.source_emits(self.items, dt)


def cell_block_emit(h, origin, i0, i1, j0, j1, k0, k1, seed=1):
    """8 particles per cell (2x2x2 jittered lattice) over the cell range [i0,i1) x [j0,j1) x [k0,k1):
    one shape-2 record of the reference's lattice emitter (31-wgsl-sim2.js:295-306)."""
    sp = h / 2
    start = [origin[0] + i0 * h, origin[1] + j0 * h, origin[2] + k0 * h]
    n = [2 * (i1 - i0), 2 * (j1 - j0), 2 * (k1 - k0)]
    return {"shape": 2, "a": start + [sp], "b": [n[0], n[1], n[2], 0], "c": [0, 0, 0, 0],
            "count": n[0] * n[1] * n[2], "seed": seed}
