"""saveWater / clearWater / restoreWater (src/js/50-sim.js:110-141, :365-373) through the C ABI.

A run saves its water, clears the tank, refills and steps it, then restores the copy and steps on.
The oracle does the same sequence, restating the reference's copy (particles, whitewater, counters,
ripple state and the host's time / count / whitewater count / impacts / full flag), and is fed the
emission records the library's emitters produce: particles, counters and the whitewater multiset
agree bit for bit and the ripple state to 1e-5 after the restore. A second run that never clears
emits, after the save point, the same records as the restored run: the sources' progress and the
seed stream come back with the water. The restore happens with a counters readback in flight and
is followed by a poll, which must not bring back the count of before the restore."""
import importlib

import numpy as np
import pytest

import twin
from twin import assert_close, suiri, ww_multiset

scene = importlib.import_module("hahenn-_amd.scene")

pytestmark = pytest.mark.gpu

DT = 1 / 120


def source_items(o):
    inf = o.info()
    return [scene.make_source(pos=[inf["ox"] + 0.35, inf["oy"] + 0.2, inf["oz"] + 0.12], mode="stream", flow=40.0,
                              speed=1.0, pitch=-70, yaw=30),
            scene.make_source(pos=[inf["ox"] + 0.1, inf["oy"] + 0.22, inf["oz"] + 0.1], mode="drops", dropD=8,
                              dropRate=20, speed=0.5, pitch=-90)]


def step_both(o, g, items, log=None):
    recs = []

    def emit(dt, t):
        r = g.scene().source_emits(items, dt)
        recs.append(r)
        return r
    g.step(DT, emit)
    assert len(recs) == 1
    if log is not None:
        log.append(recs[0])
    twin.oracle_step(o, DT, recs[0])


def oracle_save(o):
    inf = o.info()
    return {"parts": o.field("parts").copy(), "ww": o.field("ww").copy(), "counters": o.field("counters").copy(),
            "rstate": o.field("rstate").copy(),
            "meta": {k: inf[k] for k in ("time", "count", "ww_count", "impacts", "full")}}


def oracle_restore(o, c):
    o.field("parts")[:] = c["parts"]
    o.field("ww")[:] = c["ww"]
    o.field("counters")[:] = c["counters"]
    o.field("rstate")[:] = c["rstate"]
    for k, v in c["meta"].items():
        o.set(k, v)


def test_save_clear_restore_matches_oracle_and_uninterrupted_emission():
    sc = twin.small_scene(whitewater=True, R=64)
    o, g = twin.make_twin(sc)
    items = source_items(o)
    for f in range(4):
        step_both(o, g, items)
    print("saving after 4 frames", flush=True)
    cid = g.save_water()
    assert g.has_water_copy(cid)
    copy = oracle_save(o)
    # clear, refill and run on (the reference's reset: clearWater + emitNow(blockEmits))
    g.clear_water()
    o.L.or_clear_water(o.ptr)
    inf = o.info()
    rec = scene.cell_block_emit(inf["h"], [inf["ox"], inf["oy"], inf["oz"]], 2, 14, 0, 6, 0, 16, seed=321)
    g.emit_now([rec])
    o.emit_now([rec])
    for f in range(2):
        step_both(o, g, items)
    assert g.stats()["count"] == int(o.field("counters")[0])
    # restore with the last frame's counters readback still pending, then poll
    assert g.restore_water(cid)
    oracle_restore(o, copy)
    g.poll_stats()
    assert not g.has_water_copy(cid) and not g.restore_water(cid)
    st = g.stats()
    assert st["count"] == int(copy["meta"]["count"]) and st["time"] == copy["meta"]["time"]
    after = []
    for f in range(4):
        print(f"frame {f} after the restore", flush=True)
        step_both(o, g, items, after)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
    gc, oc = g.readback(suiri.F_COUNTERS), o.field("counters")
    assert [gc[0], gc[2], gc[5]] == [oc[0], oc[2], oc[5]]
    wo = int(o.info()["ww_count"])
    assert g.stats()["ww"] == wo and wo > 0
    assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo]))
    assert_close("rstate", g.readback(suiri.F_RSTATE), o.field("rstate"), 1e-5)

    # an uninterrupted run emits the same records after the save point
    o2, g2 = twin.make_twin(sc)
    for f in range(4):
        step_both(o2, g2, items)
    plain = []
    for f in range(4):
        step_both(o2, g2, items, plain)
    assert sum(len(r) for r in plain) > 0
    assert plain == after
