"""Host-side scene logic for Python callers: obstacle packing and emission lists.

Mirrors the reference's scene model and emitters (src/js/10-scene.js, src/js/55-emit.js) and the
object packing of FluidSim.setObjects (src/js/50-sim.js:165-211). All arithmetic is in double
precision like the JavaScript, and rounded to f32 where the reference writes a Float32Array.
"""
import math

import numpy as np

DEG = math.pi / 180.0
RAIN_SPEED = 6.5  # 55-emit.js:10
MAX_OBJ = 64

# SHAPES, 10-scene.js:5-15 (type id and default dims)
SHAPES = {
    "box": (0, {"w": 0.10, "h": 0.06, "d": 0.08}),
    "plate": (0, {"w": 0.22, "d": 0.10, "t": 0.012}),
    "sphere": (1, {"r": 0.035}),
    "cylinder": (2, {"r": 0.025, "h": 0.12}),
    "cone": (3, {"r": 0.04, "h": 0.08}),
    "torus": (4, {"R": 0.05, "r": 0.014}),
    "cup": (5, {"r": 0.04, "h": 0.10, "t": 0.006}),
    "container": (6, {"w": 0.14, "h": 0.06, "d": 0.10, "t": 0.008}),
    "gutter": (7, {"len": 0.28, "r": 0.03, "t": 0.006}),
}


# ---- m3 / v3, 00-util.js:10-61 ----
def euler(rx, ry, rz):
    """R = Rz * Ry * Rx from degrees, row-major (00-util.js:22-30)."""
    cx, sx, cy, sy = math.cos(rx * DEG), math.sin(rx * DEG), math.cos(ry * DEG), math.sin(ry * DEG)
    cz, sz = math.cos(rz * DEG), math.sin(rz * DEG)
    return [cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
            sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
            -sy, cy * sx, cy * cx]


def transpose(R):
    return [R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]]


def mmul(A, B):
    return [A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c] + A[r * 3 + 2] * B[6 + c] for r in range(3) for c in range(3)]


def mlog(R):
    """rotation vector of R (00-util.js:53-60)."""
    c = min(1.0, max(-1.0, (R[0] + R[4] + R[8] - 1) / 2))
    ang = math.acos(c)
    if ang < 1e-6:
        return [0.0, 0.0, 0.0]
    s = 2 * math.sin(ang)
    if abs(s) < 1e-6:
        return [0.0, 0.0, 0.0]
    return [(R[7] - R[5]) / s * ang, (R[2] - R[6]) / s * ang, (R[3] - R[1]) / s * ang]


def vlen(a):
    return math.hypot(a[0], a[1], a[2])


def vnorm(a):
    n = vlen(a) or 1.0
    return [a[0] / n, a[1] / n, a[2] / n]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


# ---- items, 10-scene.js:56-87 ----
_uid = [1]


def _next_id():
    _uid[0] += 1
    return _uid[0] - 1


def make_obstacle(shape, pos=None, dims=None, rot=None):
    t, d = SHAPES[shape]
    return {"id": _next_id(), "kind": "obstacle", "shape": shape, "pos": list(pos or [0, 0.05, 0]),
            "rot": list(rot or [0, 0, 0]), "dims": dict(dims or d)}


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


# ---- shape parameters, 10-scene.js:143-188 ----
def shape_params(o):
    d, s = o["dims"], o["shape"]
    if s == "box":
        return [0, d["w"] / 2, d["h"] / 2, d["d"] / 2, min(0.003, d["w"] / 4, d["h"] / 4, d["d"] / 4)]
    if s == "plate":
        return [0, d["w"] / 2, d["t"] / 2, d["d"] / 2, min(0.002, d["t"] / 4)]
    if s == "sphere":
        return [1, d["r"], 0, 0, 0]
    if s == "cylinder":
        return [2, d["r"], d["h"] / 2, 0, 0]
    if s == "cone":
        return [3, d["r"], d["h"] / 2, 0, 0]
    if s == "torus":
        return [4, d["R"], d["r"], 0, 0]
    if s == "cup":
        return [5, d["r"], d["h"] / 2, min(d["t"], d["r"] * 0.8), 0]
    if s == "container":
        return [6, d["w"] / 2, d["h"] / 2, d["d"] / 2, min(d["t"], d["w"] / 3, d["d"] / 3, d["h"] * 0.8)]
    if s == "gutter":
        return [7, d["len"] / 2, d["r"], d["t"], 0]
    return [0, 0.05, 0.05, 0.05, 0]


def bound_radius(o):
    t, a, b, c, e = shape_params(o)
    if t == 0 or t == 6:
        return math.hypot(a, b, c)
    if t == 1:
        return a
    if t in (2, 3, 5):
        return math.hypot(a, b)
    if t == 4:
        return a + b
    if t == 7:
        return math.hypot(a, b + c)
    return 0.1


class ObjectPacker:
    """FluidSim.setObjects packing with moving-object velocities, 50-sim.js:165-202."""

    def __init__(self):
        self.prev = {}

    def pack(self, items, dt_move):
        rows = []
        nxt = {}
        moving = False
        for it in items:
            if len(rows) >= MAX_OBJ:
                break
            if it["kind"] not in ("obstacle", "drain"):
                continue
            R = euler(*it["rot"])
            RT = transpose(R)
            if it["kind"] == "obstacle":
                sp = shape_params(it)
                typ, d, flag, bound = sp[0], sp[1:], 1, bound_radius(it) + 0.002
            else:
                typ, d, flag, bound = 9, [it["r"], 0.012, 0, 0], 2, it["r"] + 0.02
            lin, ang = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
            prev = self.prev.get(it["id"])
            if prev and dt_move > 0:
                lin = [(it["pos"][k] - prev[0][k]) * (1 / dt_move) for k in range(3)]
                lg = mlog(mmul(R, transpose(prev[1])))
                ang = [x * (1 / dt_move) for x in lg]
                if vlen(lin) > 1e-5 or vlen(ang) > 1e-4:
                    moving = True
                if vlen(lin) > 4:
                    lin = [x * 4 for x in vnorm(lin)]
                if vlen(ang) > 40:
                    ang = [x * 40 for x in vnorm(ang)]
            nxt[it["id"]] = (list(it["pos"]), R)
            rows.append([RT[0], RT[1], RT[2], 0, RT[3], RT[4], RT[5], 0, RT[6], RT[7], RT[8], 0,
                         it["pos"][0], it["pos"][1], it["pos"][2], typ, d[0], d[1], d[2], d[3],
                         lin[0], lin[1], lin[2], bound, ang[0], ang[1], ang[2], flag])
        self.prev = nxt
        return np.array(rows, dtype=np.float32).reshape(-1, 28), moving


# ---- emitters, 10-scene.js:91-99, 55-emit.js ----
def source_nozzle(src):
    Q = src["flow"] / 60000
    v = max(src["speed"], 0.05)
    return math.sqrt(4 * Q / (math.pi * v))


def source_dir(src):
    p, y = src["pitch"] * DEG, src["yaw"] * DEG
    return [math.cos(p) * math.sin(y), math.sin(p), math.cos(p) * math.cos(y)]


def basis_of(d):
    ref = [0, 1, 0] if abs(d[1]) < 0.95 else [1, 0, 0]
    u = vnorm(cross(d, ref))
    return u, cross(d, u)


class Seeds:
    """Deterministic replacement of Math.random() seeds (50-sim.js:234): a 32-bit LCG stream."""

    def __init__(self, seed=1):
        self.s = seed & 0xFFFFFFFF

    def __call__(self):
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return self.s


class Emitters:
    """sourceEmits / burstEmit state of one simulation (55-emit.js:13-79)."""

    def __init__(self, items, vp, seed=1):
        self.items, self.vp = items, vp
        self.acc = {}
        self.seeds = Seeds(seed)

    def reset(self):
        self.acc = {}

    def __call__(self, dt, t):
        out = []
        Vp = self.vp
        for src in self.items:
            if src["kind"] != "source" or not src.get("enabled", True):
                continue
            st = self.acc.setdefault(src["id"], {"acc": 0.0, "liters": 0.0, "timer": 0.0, "fired": False})
            if src.get("total", 0) > 0 and st["liters"] >= src["total"]:
                continue
            d = source_dir(src)
            u, w = basis_of(d)
            mode = src["mode"]
            if mode in ("stream", "shower"):
                Q = src["flow"] / 60000
                st["acc"] += (Q * dt) / Vp
                n = math.floor(st["acc"])
                st["acc"] -= n
                if n <= 0:
                    continue
                shower = mode == "shower"
                r = max(source_nozzle(src) / 2, 0.022) if shower else max(source_nozzle(src) / 2, 0.0005)
                spread = min(max(src["spread"], 0), 80) * DEG if shower else 0
                p = [src["pos"][k] + d[k] * 0.002 for k in range(3)]
                out.append({"shape": 0, "a": p + [r], "b": d + [src["speed"]], "c": u + [spread],
                            "d": w + [src["speed"] * dt], "count": n, "seed": self.seeds()})
                st["liters"] += n * Vp * 1000
            elif mode == "drops":
                st["timer"] -= dt
                if st["timer"] > 0:
                    continue
                st["timer"] += 1 / min(max(src["dropRate"], 0.05), 60)
                if st["timer"] < 0:
                    st["timer"] = 0
                dd = min(max(src["dropD"], 1), 40) / 1000
                V = math.pi / 6 * dd * dd * dd
                n = max(1, round(V / Vp))
                R = max(dd / 2, ((3 * n * Vp) / (4 * math.pi)) ** (1 / 3) * 0.9)
                p = [src["pos"][k] + d[k] * (R + 0.002) for k in range(3)]
                out.append({"shape": 1, "a": p + [0 if n == 1 else R],
                            "b": [d[k] * src["speed"] for k in range(3)] + [1 if n <= 8 else 0],
                            "count": n, "seed": self.seeds()})
                st["liters"] += n * Vp * 1000
            elif mode == "burst":
                if st["fired"]:
                    if not (src["burstEvery"] > 0):
                        continue
                    st["timer"] -= dt
                    if st["timer"] > 0:
                        continue
                st["fired"] = True
                st["timer"] = src["burstEvery"] if src["burstEvery"] > 0 else 0
                e = burst_emit(Vp, src)
                e["seed"] = self.seeds()
                out.append(e)
                st["liters"] += e["count"] * Vp * 1000
            elif mode == "rain":
                side = min(max(src["rainSize"], 0.05), 2)
                flux = (min(max(src["rainRate"], 1), 500) / 3.6e6) * side * side
                st["acc"] += (flux * dt) / Vp
                n = math.floor(st["acc"])
                st["acc"] -= n
                if n <= 0:
                    continue
                out.append({"shape": 3, "a": list(src["pos"]) + [side / 2], "b": [0, -RAIN_SPEED, 0, 0],
                            "d": [0, 0, 0, RAIN_SPEED * dt], "count": n, "seed": self.seeds()})
                st["liters"] += n * Vp * 1000
        return out


def burst_emit(vp, src):
    d = source_dir(src)
    V = min(max(src["burst"], 0.01), 50) / 1000
    n = max(1, round(V / vp))
    R = ((3 * V) / (4 * math.pi)) ** (1 / 3)
    p = [src["pos"][k] + d[k] * (R + 0.004) for k in range(3)]
    return {"shape": 1, "a": p + [R], "b": [d[k] * src["speed"] for k in range(3)] + [0], "count": n}


def block_emits(h, origin, items, fill=False, seeds=None):
    """Lattice fills of water blocks, 55-emit.js:83-106."""
    out = []
    sp = h / 2
    seeds = seeds or Seeds(7)
    for b in items:
        if b["kind"] != "block":
            continue
        d = b["dims"]
        lo = [b["pos"][0] - d["w"] / 2, b["pos"][1] - d["h"] / 2, b["pos"][2] - d["d"] / 2]
        hi = [b["pos"][0] + d["w"] / 2, b["pos"][1] + d["h"] / 2, b["pos"][2] + d["d"] / 2]
        tlo = [origin[0] + sp * 0.5, 0.0005, origin[2] + sp * 0.5]
        thi = [-origin[0] - sp * 0.5, math.inf, -origin[2] - sp * 0.5]
        inside = lo[0] < thi[0] and hi[0] > tlo[0] and lo[2] < thi[2] and hi[2] > tlo[2]
        if inside:
            for k in (0, 2):
                lo[k] = max(lo[k], tlo[k])
                hi[k] = min(hi[k], thi[k])
            lo[1] = max(lo[1], tlo[1])
        n = [max(0, math.floor((hi[k] - lo[k]) / sp)) for k in range(3)]
        total = n[0] * n[1] * n[2]
        if total <= 0:
            continue
        start = [lo[k] + ((hi[k] - lo[k]) - n[k] * sp) / 2 for k in range(3)]
        out.append({"shape": 2, "a": start + [sp], "b": [n[0], n[1], n[2], 1 if fill else 0], "c": [0, 0, 0, 0],
                    "count": total, "seed": seeds()})
    return out


def cell_block_emit(h, origin, i0, i1, j0, j1, k0, k1, seed=1):
    """8 particles per cell (2x2x2 jittered lattice) over the cell range [i0,i1) x [j0,j1) x [k0,k1):
    one shape-2 record of the reference's lattice emitter (31-wgsl-sim2.js:295-306)."""
    sp = h / 2
    start = [origin[0] + i0 * h, origin[1] + j0 * h, origin[2] + k0 * h]
    n = [2 * (i1 - i0), 2 * (j1 - j0), 2 * (k1 - k0)]
    return {"shape": 2, "a": start + [sp], "b": [n[0], n[1], n[2], 0], "c": [0, 0, 0, 0],
            "count": n[0] * n[1] * n[2], "seed": seed}
