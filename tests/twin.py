"""Twin set-up for parity tests: one oracle (CPU restatement) and one GPU simulation built from the
same scene description, plus comparison helpers."""
import importlib

import numpy as np

import oracle as O

suiri = importlib.import_module("hahenn-_amd")
scene = importlib.import_module("hahenn-_amd.scene")


def small_scene(tank=(0.5, 0.25, 0.25), h=1 / 64, iters=30, R=64, cap=60000, ww_cap=20000, flip=0.95,
                omega=1.7, whitewater=True, obstacle=True, fill_cells=(12, 10, 16), reorder_every=0, top_height=0,
                extra=None):
    """A dam break with a box obstacle (the shape of the 64x32x32 dam-break workload, smaller)."""
    items = []
    if obstacle:
        items.append(scene.make_obstacle("box", pos=[0.05, 0.06, 0.0], dims={"w": 0.06, "h": 0.08, "d": 0.1}))
        items.append(scene.make_obstacle("sphere", pos=[0.15, 0.04, 0.05], dims={"r": 0.03}))
    return dict(tank=tank, h=h, iters=iters, R=R, cap=cap, ww_cap=ww_cap, flip=flip, omega=omega,
                whitewater=whitewater, items=items, fill_cells=fill_cells, reorder_every=reorder_every,
                top_height=top_height, extra=dict(extra or {}))


def make_twin(sc, threads=4):
    q = O.quality("standard", h=sc["h"], iters=sc["iters"], ripple_R=sc["R"], cap=sc["cap"],
                  ww_cap=sc["ww_cap"], flip=sc["flip"], omega=sc["omega"], reorder_every=sc["reorder_every"],
                  top_height=sc.get("top_height", 0))
    if not sc["whitewater"]:
        q.ww_rate = 0.0
    for k, v in sc.get("extra", {}).items():   # the same settings on both sides (same field names)
        setattr(q, k, v)
    o = O.Oracle(sc["tank"], q, threads=threads)
    cfg = suiri.config("standard", tank_w=sc["tank"][0], tank_h=sc["tank"][1], tank_d=sc["tank"][2], h=sc["h"],
                       iters=sc["iters"], ripple_R=sc["R"], cap=sc["cap"], ww_cap=sc["ww_cap"], flip=sc["flip"],
                       omega=sc["omega"], whitewater=1 if sc["whitewater"] else 0,
                       reorder_every=sc["reorder_every"], top_height=sc.get("top_height", 0),
                       **sc.get("extra", {}))
    g = suiri.Sim(cfg)
    rows, _ = scene.Scene(cfg).pack(sc["items"], 0)
    o.set_objects(rows)
    g.set_objects_rows(rows)
    inf = o.info()
    fc = sc["fill_cells"]
    rec = scene.cell_block_emit(inf["h"], [inf["ox"], inf["oy"], inf["oz"]], 0, fc[0], 0, fc[1], 0, fc[2], seed=11)
    o.emit_now([rec])
    g.emit_now([rec])
    g.set_substeps(1)
    return o, g


def oracle_step(o, frame_dt, records=()):
    o.L.or_step_begin(o.ptr)
    o.step_substep(frame_dt, list(records))
    o.L.or_frame_tail(o.ptr, float(frame_dt))
    o.L.or_poll_stats(o.ptr)


def gpu_step(g, frame_dt, records=()):
    g.step(frame_dt, (lambda dt, t: list(records)) if records else None)


def ww_multiset(a):
    """Whitewater records sorted by their bit patterns (a multiset compared bit for bit: survivors are
    appended with atomics, so their order is not fixed)."""
    a = np.ascontiguousarray(a, np.float32)
    return a[np.lexsort(a.view(np.uint32).T[::-1])].view(np.uint32)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


def assert_close(name, got, want, tol):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    e = rel_err(got, want) if got.size else 0.0
    assert e <= tol, f"{name}: max relative error {e:.3e} > {tol:.1e}"
    return e


def make_twin_workload(wl, quality="standard", threads=8):
    """Oracle and GPU simulation of a bench workload (hahenn-_amd/workloads.py): the same
    configuration, objects (packed by the library's host code), lattice fill and, when the bench
    re-orders after the fill, the same re-ordering."""
    c = dict(wl["cfg"])
    qkeys = {f for f, _ in O.OrQuality._fields_}
    q = O.quality(quality, **{k: v for k, v in c.items() if k in qkeys})
    if not c.get("whitewater", 1):
        q.ww_rate = 0.0
    o = O.Oracle((c["tank_w"], c["tank_h"], c["tank_d"]), q, threads=threads)
    g = suiri.Sim(suiri.config(quality, **c))
    rows, _ = scene.Scene(g.cfg).pack(wl["items"], 0)
    o.set_objects(rows)
    g.set_objects_rows(rows)
    inf = o.info()
    recs = [scene.cell_block_emit(inf["h"], [inf["ox"], inf["oy"], inf["oz"]], *blk, seed=wl["seed"] + i)
            for i, blk in enumerate(wl["blocks"])]
    o.emit_now(recs)
    g.emit_now(recs)
    g.set_substeps(1)
    return o, g
