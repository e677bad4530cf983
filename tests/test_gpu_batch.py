"""Batched tanks: every tank of a batch (one launch per stage for all tanks, eager or replayed from a
captured hipGraph) ends bit-identical to the same tank run alone through the single-simulation path,
which the oracle parity tests pin. Whitewater is compared as a multiset: its survivors are appended
with atomics, so their order is not fixed (the reference's WW pass appends the same way,
src/js/31-wgsl-sim2.js:361-409)."""
import importlib

import numpy as np
import pytest

suiri = importlib.import_module("hahenn-_amd")
scene = importlib.import_module("hahenn-_amd.scene")

pytestmark = pytest.mark.gpu

DT = 1 / 120

# grid sizes 16x10x12 .. 77x32x38 (two size buckets), different SOR iteration counts, ripple on/off,
# absorbing mask on/off, periodic re-ordering (one with a phase offset), one drawing its top-height map
SPECS = [
    dict(tank=(0.32, 0.2, 0.24), h=1 / 50, iters=20, R=32, reorder=4, phase=2, obst="sphere", fill=(6, 5, 8),
         mask=True),
    dict(tank=(0.5, 0.25, 0.25), h=1 / 64, iters=30, R=64, reorder=3, obst="box", fill=(12, 10, 16), mask=False),
    dict(tank=(0.3, 0.3, 0.3), h=1 / 40, iters=26, R=0, reorder=0, obst=None, fill=(5, 7, 5), mask=True),
    dict(tank=(0.6, 0.25, 0.3), h=1 / 128, iters=24, R=64, reorder=0, obst="box", fill=(20, 12, 30), mask=True,
         top=1),
]


def cfg_of(sp):
    return suiri.config("standard", tank_w=sp["tank"][0], tank_h=sp["tank"][1], tank_d=sp["tank"][2], h=sp["h"],
                        iters=sp["iters"], ripple_R=sp["R"], cap=40000, ww_cap=30000, flip=0.95, omega=1.7,
                        whitewater=1, reorder_every=sp["reorder"], reorder_phase=sp.get("phase", 0),
                        top_height=sp.get("top", 0))


def items_of(sp, shift=0.0):
    if sp["obst"] == "sphere":
        return [dict(scene.make_obstacle("sphere", pos=[0.02 + shift, 0.05, 0.0], dims={"r": 0.03}), id=1)]
    if sp["obst"] == "box":
        return [dict(scene.make_obstacle("box", pos=[0.05 + shift, 0.06, 0.0], dims={"w": 0.06, "h": 0.08, "d": 0.1}),
                     id=1)]
    return []


class Driver:
    """Applies the same scene events to a tank, whichever path steps it."""

    def __init__(self, sim, sp, seed):
        self.sim, self.sp, self.seed = sim, sp, seed
        self.packer = scene.Scene()
        rows, _ = self.packer.pack(items_of(sp), 0)
        sim.set_objects_rows(rows)
        if not sp["mask"]:
            sim.set_ripple_mask(False)
        fc = sp["fill"]
        self.emit(0, fc[0], 0, fc[1], 0, fc[2])

    def emit(self, i0, i1, j0, j1, k0, k1):
        s = self.sim
        rec = scene.cell_block_emit(s.h, s.origin, i0, i1, j0, j1, k0, k1, seed=self.seed)
        self.seed += 1
        s.emit_now([rec])
        # the single path's ripple depth reads the host count, a batch reads the device count at frame
        # start: refreshing the host count makes the two the same
        s.stats()

    def event(self, frame):
        if frame == 3 and self.sp["obst"]:   # the obstacle moves: new rows, velocities from the move
            rows, _ = self.packer.pack(items_of(self.sp, shift=0.01), DT)
            self.sim.set_objects_rows(rows)
        if frame == 5:                       # water poured in between frames
            nx = self.sim.nx
            self.emit(nx - 4, nx - 1, 3, 5, 1, 4)


def source_records(sim, frame, seed):
    """A falling ball of water (emitter shape 1) every other frame."""
    if frame % 2:
        return []
    x = sim.origin[0] + (0.3 + 0.05 * (frame % 5)) * sim.nx * sim.h
    z = sim.origin[2] + 0.5 * sim.nz * sim.h
    y = 0.8 * sim.ny * sim.h
    return [dict(a=[x, y, z, 2.5 * sim.h], b=[0.1, -0.8, 0.0, 1.0], shape=1, count=300, seed=seed * 1000 + frame)]


def single_run(sp, seed, frames, sources=False):
    s = suiri.Sim(cfg_of(sp))
    d = Driver(s, sp, seed)
    s.set_substeps(1)
    for f in range(frames):
        d.event(f)
        recs = source_records(s, f, seed) if sources else []
        s.step(DT, (lambda dt, t: recs) if recs else None)
        s.stats()   # waits and refreshes the counts (the next frame's ripple depth reads them)
    return s


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint16)


def assert_same_tank(got, want, label):
    sg, sw = got.stats(), want.stats()
    for k in ("count", "ww", "impacts", "frames", "time"):
        assert sg[k] == sw[k], f"{label}: {k} {sg[k]} != {sw[k]}"
    assert sg["vmax"] == sw["vmax"], label
    n = sg["count"]
    assert np.array_equal(bits(got.readback(suiri.F_PARTICLES)[:n]), bits(want.readback(suiri.F_PARTICLES)[:n])), \
        f"{label}: particles differ"
    for f in (suiri.F_VEL, suiri.F_PRES, suiri.F_CELLCOUNT, suiri.F_DIVG):
        assert np.array_equal(bits(got.readback(f)), bits(want.readback(f))), f"{label}: field {f} differs"
    w = sg["ww"]
    a = got.readback(suiri.F_WW)[:w]
    b = want.readback(suiri.F_WW)[:w]
    a = a[np.lexsort(a.T[::-1])]
    b = b[np.lexsort(b.T[::-1])]
    assert np.array_equal(bits(a), bits(b)), f"{label}: whitewater multiset differs"
    if got.grid.R:
        for f in (suiri.F_RSTATE, suiri.F_RTEX, suiri.F_RSRC, suiri.F_TOPH, suiri.F_RMASK):
            assert np.array_equal(bits(got.readback(f)), bits(want.readback(f))), f"{label}: ripple field {f} differs"


def run_batch(frames, capture_at=None, tune_at=None, variants=None):
    b = suiri.Batch([cfg_of(sp) for sp in SPECS])
    drivers = [Driver(t, sp, 100 * i + 7) for i, (t, sp) in enumerate(zip(b.tanks, SPECS))]
    if variants:
        b.set_variants(variants)
    for f in range(frames):
        for d in drivers:
            d.event(f)
        if f == capture_at:
            b.capture(DT)
        if f == tune_at:
            b.autotune(reps=1)
        b.step(DT)
        b.sync()
    return b


@pytest.fixture(scope="module")
def singles():
    return [single_run(sp, 100 * i + 7, 8) for i, sp in enumerate(SPECS)]


def test_batch_eager_equals_single(singles):
    print("batch eager, 4 tanks, 8 frames", flush=True)
    b = run_batch(8)
    for i, (t, s) in enumerate(zip(b.tanks, singles)):
        assert_same_tank(t, s, f"tank {i}")
    b.close()


def test_batch_graph_replay_equals_single(singles):
    """Frames 2..7 replay one captured graph (with an object move, an emission and re-orderings in
    between): each replay reads the tank state from the device table."""
    print("batch graph replay, 4 tanks, 8 frames", flush=True)
    b = run_batch(8, capture_at=2)
    for i, (t, s) in enumerate(zip(b.tanks, singles)):
        assert_same_tank(t, s, f"tank {i}")
    b.close()


def test_batch_autotune_keeps_results(singles):
    print("batch autotune at frame 4, 4 tanks, 8 frames", flush=True)
    b = run_batch(8, capture_at=1, tune_at=4)
    for i, (t, s) in enumerate(zip(b.tanks, singles)):
        assert_same_tank(t, s, f"tank {i}")
    b.set_variants([(7, 0), (1, 3), (4, 2)])
    b.step(DT)
    b.sync()
    assert b.tanks[0].stats()["frames"] == 9
    b.close()


def test_batch_binned_p2g_equals_single(singles):
    """The binned P2G in every size bucket (its cell lists and tiles over the batch's work maps),
    eager then replayed from a captured graph: each tank equals its single run with the default
    kernels."""
    print("batch binned P2G, 4 tanks, 8 frames", flush=True)
    b = run_batch(8, capture_at=3, variants=[(8, 0), (8, 3), (8, 1)])
    for i, (t, s) in enumerate(zip(b.tanks, singles)):
        assert_same_tank(t, s, f"tank {i}")
    with pytest.raises(suiri.SuiriError, match="out of range"):
        b.set_variants([(8, 4), (3, 2), (8, 0)])
    b.close()


def test_batch_errors():
    with pytest.raises(suiri.SuiriError):
        suiri.Batch([])
    bad = cfg_of(SPECS[0])
    bad.ripple_R = 48   # not a power of two
    with pytest.raises(suiri.SuiriError, match="tank 1"):
        suiri.Batch([cfg_of(SPECS[0]), bad])


def test_batch_membership_changes_and_batched_emission():
    """Continuous batching: tanks join at different frames, one pauses and resumes, one leaves, every
    stepped tank gets source emission through one sl_batch_emit per frame, and the last frames replay
    a captured graph. Each tank still equals its own single run."""
    print("batch membership: 5 slots, 8 frames, sources", flush=True)
    plan = [dict(spec=0, add=0, pause=(), remove=None), dict(spec=1, add=2, pause=(), remove=None),
            dict(spec=2, add=0, pause=(3, 4), remove=None), dict(spec=3, add=1, pause=(), remove=None),
            dict(spec=0, add=0, pause=(), remove=4)]
    b = suiri.Batch(capacity=5)
    live = {}   # plan index -> [slot, driver, frames stepped]
    for f in range(8):
        for i, p in enumerate(plan):
            if f == p["add"]:
                slot = b.add(cfg_of(SPECS[p["spec"]]))
                live[i] = [slot, Driver(b.tanks[slot], SPECS[p["spec"]], 50 + i), 0]
            if f == p["remove"] and i in live:
                b.remove(live.pop(i)[0])
        stepped = {}
        for i, (slot, d, n) in live.items():
            on = f not in plan[i]["pause"]
            b.set_stepped(slot, on)
            if on:
                d.event(n)
                recs = source_records(b.tanks[slot], n, 50 + i)
                if recs:
                    stepped[slot] = recs
        if stepped:
            b.emit(stepped)
        if f == 6:
            b.capture(DT)
            assert b.captured
        b.step(DT)
        b.sync()
        for i, (slot, d, n) in live.items():
            if f not in plan[i]["pause"]:
                live[i][2] += 1
    assert sorted(live) == [0, 1, 2, 3]
    for i, (slot, d, n) in live.items():
        want = single_run(SPECS[plan[i]["spec"]], 50 + i, n, sources=True)
        assert_same_tank(b.tanks[slot], want, f"plan {i} ({n} frames)")
        want.close()
    b.close()
