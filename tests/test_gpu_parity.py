"""GPU parity: the HIP path (through the C ABI) against the oracle on identical inputs.

Tolerances follow the stated bar: after one step grid velocities, pressure and particle positions
agree to <= 1e-5 relative error (relative to the field's largest magnitude); integer and index
buffers (counters, cell counts, fixed-point accumulators) must match exactly. Kernels that only use
IEEE arithmetic and sqrt reproduce the oracle bit for bit; the tests report that too.
"""
import numpy as np
import pytest

import twin
from twin import assert_close, scene, suiri, ww_multiset

pytestmark = pytest.mark.gpu

TOL = 1e-5


def live(parts, n):
    return parts[: int(n)]


def test_emission_is_bitwise():
    o, g = twin.make_twin(twin.small_scene())
    n = int(o.field("counters")[0])
    assert g.readback(suiri.F_COUNTERS)[0] == n > 0
    gp = g.readback(suiri.F_PARTICLES)[:n]
    op = o.field("parts")[:n]
    assert np.array_equal(gp, op)


def test_geom_face_fractions():
    o, g = twin.make_twin(twin.small_scene())
    o.write_params(1 / 120)
    o.kernel("geom")
    g.run_stage(suiri.S_GEOM, 0, 1 / 120)
    assert_close("frac", g.readback(suiri.F_FRAC), o.field("frac"), TOL)
    assert_close("svel", g.readback(suiri.F_SVEL), o.field("svel"), TOL)
    assert np.array_equal(g.readback(suiri.F_FRAC), o.field("frac"))


def test_one_step_fields():
    o, g = twin.make_twin(twin.small_scene())
    twin.oracle_step(o, 1 / 120)
    twin.gpu_step(g, 1 / 120)
    st = g.stats()
    n = int(o.field("counters")[0])
    assert st["count"] == n
    assert np.array_equal(g.readback(suiri.F_CELLCOUNT), o.field("cellCount"))
    assert_close("vel", g.readback(suiri.F_VEL)[:, :3], o.field("vel")[:, :3], TOL)
    assert_close("velOld", g.readback(suiri.F_VELOLD), o.field("velOld"), TOL)
    assert_close("pres", g.readback(suiri.F_PRES), o.field("pres"), TOL)
    assert_close("divg", g.readback(suiri.F_DIVG), o.field("divg"), TOL)
    gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
    assert_close("positions", gp[:, :3], op[:, :3], TOL)
    assert_close("velocities", gp[:, 4:7], op[:, 4:7], TOL)
    assert np.array_equal(gp[:, 3], op[:, 3]) and np.array_equal(gp[:, 7], op[:, 7])
    # renderer-facing dispatch / draw-indirect arguments of the last COUNT (31-wgsl-sim2.js:343-348)
    assert np.array_equal(g.readback(suiri.F_ARGS)[:14], o.field("args")[:14])


def test_p2g_variants_from_identical_state():
    """Every P2G variant (chunked LDS boxes 0-7, binned 8) reproduces the oracle's fixed-point
    accumulators and cell counts exactly, also when the particle order is shuffled (integer sums do
    not depend on the order; scattered chunks exercise the global-atomic fallback) and when two cells
    hold 600 and 100 particles (the binned splat's chunk queue)."""
    o, g = twin.make_twin(twin.small_scene())
    for _ in range(8):
        twin.oracle_step(o, 1 / 120)
    n = int(o.field("counters")[0])
    dt = 1 / 120
    o.L.or_step_begin(o.ptr)
    o.set("frame", o.info()["frame"] + 1)
    o.write_params(dt)
    o.field("acc")[:] = 0
    o.field("cellCount")[:] = 0
    o.kernel("p2g")
    want_acc, want_cc = o.field("acc").copy(), o.field("cellCount").copy()
    parts = o.field("parts").copy()
    rng = np.random.default_rng(5)
    shuffled = parts.copy()
    shuffled[:n] = parts[:n][rng.permutation(n)]
    # crowded: 700 particles packed into two cells (as under an emitter's nozzle), the rest unchanged
    crowded = parts.copy()
    inf = o.info()
    cell = np.array([inf["ox"], inf["oy"], inf["oz"]]) + inf["h"] * np.array([10.0, 5.0, 7.0])
    crowded[:600, :3] = cell + inf["h"] * rng.uniform(0.01, 0.99, (600, 3))
    crowded[600:700, :3] = cell + inf["h"] * (np.array([1.0, 0.0, 0.0]) + rng.uniform(0.01, 0.99, (100, 3)))
    crowded[:700, 3] = 1.0
    o.field("parts")[:] = crowded
    o.field("acc")[:] = 0
    o.field("cellCount")[:] = 0
    o.kernel("p2g")
    want_crowded = o.field("acc").copy(), o.field("cellCount").copy()
    cases = (("sorted", parts, (want_acc, want_cc)), ("shuffled", shuffled, (want_acc, want_cc)),
             ("crowded", crowded, want_crowded))
    for order, p, (wa, wc) in cases:
        g.upload(suiri.F_PARTICLES, p)
        g.upload(suiri.F_COUNTERS, o.field("counters"))
        g.set_scalar("count", n)
        for v in range(9):
            g.set_scalar("p2g_variant", v)
            g.upload(suiri.F_ACC, np.zeros_like(want_acc))
            g.run_stage(suiri.S_P2G, 0, dt)
            assert np.array_equal(g.readback(suiri.F_ACC), wa), (order, v)
            assert np.array_equal(g.readback(suiri.F_CELLCOUNT), wc), (order, v)


def test_stage_by_stage_from_identical_state():
    """Each stage started from the oracle's own state reproduces the oracle's output."""
    o, g = twin.make_twin(twin.small_scene())
    for _ in range(6):   # let the dam break get moving
        twin.oracle_step(o, 1 / 120)
    n = int(o.field("counters")[0])
    dt = 1 / 120
    o.L.or_step_begin(o.ptr)
    o.set("frame", o.info()["frame"] + 1)
    o.write_params(dt)
    # identical inputs on the GPU
    g.upload(suiri.F_PARTICLES, o.field("parts"))
    g.upload(suiri.F_COUNTERS, o.field("counters"))
    g.upload(suiri.F_PRES, o.field("pres"))
    g.set_scalar("frame", o.info()["frame"])
    g.set_scalar("count", n)
    g.run_stage(suiri.S_GEOM, 0, dt)
    # P2G
    o.field("acc")[:] = 0
    o.field("cellCount")[:] = 0
    o.kernel("p2g")
    g.run_stage(suiri.S_P2G, 0, dt)
    assert np.array_equal(g.readback(suiri.F_ACC), o.field("acc"))
    assert np.array_equal(g.readback(suiri.F_CELLCOUNT), o.field("cellCount"))
    # NORMALIZE + FORCES
    o.kernel("normalize")
    o.kernel("forces")
    g.run_stage(suiri.S_NORMFORCE, 0, dt)
    assert np.array_equal(g.readback(suiri.F_VELOLD), o.field("velOld"))
    assert np.array_equal(g.readback(suiri.F_VEL)[:, :3], o.field("vel")[:, :3])
    # DIVERGENCE
    o.kernel("divergence")
    g.run_stage(suiri.S_DIVERGENCE, 0, dt)
    assert np.array_equal(g.readback(suiri.F_DIVG), o.field("divg"))
    # SOR, pass by pass
    for it in range(int(o.info()["iters"])):
        for par in (0, 1):
            o.kernel("pressure", par)
            g.run_stage(suiri.S_PRESSURE, par, dt)
        if it in (0, 5):
            assert_close(f"pres it{it}", g.readback(suiri.F_PRES), o.field("pres"), TOL)
    assert_close("pres", g.readback(suiri.F_PRES), o.field("pres"), TOL)
    # PROJECT
    o.kernel("project")
    g.run_stage(suiri.S_PROJECT, 0, dt)
    assert_close("vel", g.readback(suiri.F_VEL)[:, :3], o.field("vel")[:, :3], TOL)
    # G2P
    o.kernel("g2p")
    g.run_stage(suiri.S_G2P, 0, dt)
    gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
    assert_close("g2p positions", gp[:, :3], op[:, :3], TOL)
    assert_close("g2p velocities", gp[:, 4:7], op[:, 4:7], TOL)
    assert np.array_equal(gp[:, 3], op[:, 3])


def test_g2p_from_uploaded_grid():
    """G2P from grid velocities uploaded through the C ABI (vel, and velOld with its mask in the 4th
    column). The GPU keeps the mask in velOld's place as a bit pattern, so a masked-out node's velOld
    is dropped (the reference's transfer1 never reads it, 31-wgsl-sim2.js:18-19; NORMALIZE leaves it
    0) and reads back as 0. Values planted there must not change the particles."""
    o, g = twin.make_twin(twin.small_scene())
    dt = 1 / 120
    for _ in range(5):
        twin.oracle_step(o, dt)
    o.L.or_step_begin(o.ptr)
    o.set("frame", o.info()["frame"] + 1)
    o.write_params(dt)
    o.field("acc")[:] = 0
    o.field("cellCount")[:] = 0
    for k in ("p2g", "normalize", "forces", "divergence"):
        o.kernel(k)
    for _ in range(int(o.info()["iters"])):
        o.kernel("pressure", 0)
        o.kernel("pressure", 1)
    o.kernel("project")
    vo = o.field("velOld")
    mask = vo[:, 3].astype(np.uint32)
    off = [(mask >> c) & 1 == 0 for c in range(3)]
    assert all(o_.any() for o_ in off) and all((~o_).any() for o_ in off)
    rng = np.random.default_rng(5)
    for c in range(3):   # planted at masked-out nodes: transfer1 skips them
        vo[off[c], c] = rng.uniform(-3, 3, int(off[c].sum())).astype(np.float32)
    n = int(o.field("counters")[0])
    g.upload(suiri.F_PARTICLES, o.field("parts"))
    g.upload(suiri.F_COUNTERS, o.field("counters"))
    g.upload(suiri.F_CELLCOUNT, o.field("cellCount"))
    g.upload(suiri.F_VELOLD, vo)
    g.upload(suiri.F_VEL, o.field("vel"))
    g.set_scalar("frame", o.info()["frame"])
    g.run_stage(suiri.S_GEOM, 0, dt)
    want_vo = vo.copy()
    for c in range(3):
        want_vo[off[c], c] = 0
    assert np.array_equal(g.readback(suiri.F_VELOLD), want_vo)
    assert np.array_equal(g.readback(suiri.F_VEL)[:, :3], o.field("vel")[:, :3])
    print(f"G2P from uploaded grid: {n} particles", flush=True)
    o.kernel("g2p")
    g.run_stage(suiri.S_G2P, 0, dt)
    gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))


def test_reorder_is_bitwise():
    """The spatial re-ordering (stable sort by cell index) produces the oracle's permutation."""
    o, g = twin.make_twin(twin.small_scene())
    for _ in range(3):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
    o.kernel("reorder")
    g.reorder()
    n = int(o.field("counters")[0])
    assert g.stats()["count"] == n
    assert np.array_equal(g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n])


def test_steps_with_periodic_reorder():
    o, g = twin.make_twin(twin.small_scene(reorder_every=3, whitewater=False))
    for _ in range(10):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
    n = int(o.field("counters")[0])
    assert g.stats()["count"] == n
    gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
    assert_close("positions", gp[:, :3], op[:, :3], TOL)
    assert np.array_equal(gp[:, 3], op[:, 3])


def test_explicit_reorder_keeps_the_cadence():
    """An explicit re-ordering between frames leaves the periodic cadence where it was (the oracle
    restarts its clock only when the period elapses), so the automatic re-orderings of later frames
    fall on the same frames on both sides."""
    o, g = twin.make_twin(twin.small_scene(reorder_every=3, whitewater=False))
    dt = 1 / 120
    twin.oracle_step(o, dt)
    twin.gpu_step(g, dt)
    o.kernel("reorder")
    g.reorder()
    for f in range(5):
        twin.oracle_step(o, dt)
        twin.gpu_step(g, dt)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f


def test_autotune_keeps_results_bitwise():
    """The tuner times every kernel variant on the live state and restores it; the variants compute
    identical results, so the run continues exactly like the oracle."""
    o, g = twin.make_twin(twin.small_scene(whitewater=False))
    for _ in range(4):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
    ch = g.autotune(reps=1)
    assert 0 <= ch["p2g_variant"] <= 8 and 0 <= ch["g2p_variant"] < 4
    for v in range(9):   # and every P2G variant on its own (8: binned)
        g.set_scalar("p2g_variant", v)
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
    n = int(o.field("counters")[0])
    assert g.stats()["count"] == n
    assert np.array_equal(g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n])


def test_g2p_variants_with_stream_source_bitwise():
    """Every G2P variant (occupancy targets) after the binned and a chunked P2G gives the oracle's
    particles bit for bit, frame after frame, with a stream source pouring in from above the grid
    (particles outside the grid and crowded cells under the stream), whitewater spawns and splash
    crowns (same multiset: their slot order follows atomics), ripple impulses and counters."""
    sc = twin.small_scene(whitewater=True)
    o, g = twin.make_twin(sc)
    inf = o.info()
    top = inf["oy"] + 0.25
    src = scene.make_source(pos=[inf["ox"] + 0.3, top + 0.03, inf["oz"] + 0.12], mode="stream", flow=120.0,
                            speed=1.2, pitch=-90, yaw=0)
    em = scene.Emitters([src], scene.Scene(g.cfg))
    plan = [(8, 0), (8, 1), (8, 2), (8, 3), (3, 2), (5, 3)] * 3
    outside = 0
    for f, (pv, gv) in enumerate(plan):
        g.set_scalar("p2g_variant", pv)
        g.set_scalar("g2p_variant", gv)
        recs = em(1 / 120, f / 120)
        twin.oracle_step(o, 1 / 120, recs)
        twin.gpu_step(g, 1 / 120, recs)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), (f, pv, gv)
        outside += int((op[:, 1] > top).sum())
        assert np.array_equal(g.readback(suiri.F_CELLCOUNT), o.field("cellCount")), f
        gc, oc = g.readback(suiri.F_COUNTERS), o.field("counters")
        assert [gc[0], gc[4], gc[5]] == [oc[0], oc[4], oc[5]], f   # count, vmax bits, impacts
    assert outside > 0, "the stream should leave particles above the grid"
    wo = int(o.info()["ww_count"])
    assert wo > 0 and g.stats()["ww"] == wo
    assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo]))


def test_stream_source_over_three_substeps_bitwise():
    """Three substeps per frame with a stream source emitting in every substep (each substep's emission
    is committed to the count by the substep's own cell-count clear): the oracle's particles, cell
    counts and counters bit for bit after every frame, the substeps' records drawn as FluidSim.step
    draws them (src/js/50-sim.js:300-317)."""
    sc = twin.small_scene(fill_cells=(12, 8, 16))
    o, g = twin.make_twin(sc)
    inf = o.info()
    top = inf["oy"] + 0.2
    src = scene.make_source(pos=[inf["ox"] + 0.28, top, inf["oz"] + 0.12], mode="stream", flow=90.0,
                            speed=1.0, pitch=-90, yaw=0)
    em = scene.Emitters([src], scene.Scene(g.cfg))
    g.set_substeps(3)
    frame_dt = 1 / 120
    sub = frame_dt / 3   # the library's substep: frame_dt / n in double
    emitted = 0
    for f in range(10):
        subs = [em(sub, (3 * f + k) * sub) for k in range(3)]
        emitted += sum(r["count"] for recs in subs for r in recs)
        o.L.or_step_begin(o.ptr)
        for recs in subs:
            o.step_substep(sub, recs)
        o.L.or_frame_tail(o.ptr, frame_dt)
        o.L.or_poll_stats(o.ptr)
        it = iter(subs)
        assert g.step(frame_dt, lambda dt, t: next(it)) == 3
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
        assert np.array_equal(g.readback(suiri.F_CELLCOUNT), o.field("cellCount")), f
        gc, oc = g.readback(suiri.F_COUNTERS), o.field("counters")
        assert [gc[0], gc[4], gc[5]] == [oc[0], oc[4], oc[5]], f   # count, vmax bits, impacts
    assert emitted > 0 and n > 12 * 8 * 16 * 8, "the source should have added particles"


def test_every_obstacle_shape_and_drain_bitwise():
    """G2P skips the signed distances of obstacles (and the drain test of drains) whose bounding
    sphere is beyond the contact radius; a dam break through one obstacle of every shape type (one
    rotated) and a drain must still give the oracle's particles bit for bit, frame after frame."""
    items = [scene.make_obstacle("box", pos=[-0.18, 0.04, -0.06]),
             scene.make_obstacle("plate", pos=[-0.1, 0.09, 0.06], rot=[0.3, 0.2, 0.1]),
             scene.make_obstacle("sphere", pos=[-0.05, 0.03, 0.0]),
             scene.make_obstacle("cylinder", pos=[0.02, 0.06, -0.07]),
             scene.make_obstacle("cone", pos=[0.08, 0.04, 0.06]),
             scene.make_obstacle("torus", pos=[0.14, 0.02, -0.04]),
             scene.make_obstacle("cup", pos=[0.2, 0.05, 0.05]),
             scene.make_obstacle("container", pos=[-0.02, 0.12, 0.05]),
             scene.make_obstacle("gutter", pos=[0.1, 0.1, -0.08]),
             scene.make_drain(pos=[0.21, 0.0, -0.08], r=0.02)]
    sc = twin.small_scene(fill_cells=(32, 12, 16))
    sc["items"] = items
    o, g = twin.make_twin(sc)
    n0 = int(o.field("counters")[0])
    for f in range(12):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
    assert n < n0, "the drain should take particles"


def test_moving_obstacle_and_drain_bitwise():
    """A sphere sweeping through the water and a drain moving along the floor, both moved every frame
    (G2P's object mask, which lets particles far from every object skip the object loops, is rebuilt
    after each move): the oracle's particles bit for bit, frame after frame, with collisions and drain
    kills happening."""
    sc = twin.small_scene(fill_cells=(24, 12, 16))
    o, g = twin.make_twin(sc)
    sc_g = scene.Scene(g.cfg)
    n0 = int(o.field("counters")[0])
    items = [scene.make_obstacle("sphere", dims={"r": 0.035}), scene.make_drain(r=0.025)]
    for f in range(14):
        items[0]["pos"] = [-0.2 + 0.025 * f, 0.05, 0.01 * f - 0.05]   # the same ids: moving, with velocities
        items[1]["pos"] = [0.2 - 0.02 * f, 0.0, -0.06 + 0.008 * f]
        rows, _ = sc_g.pack(items, 1 / 120)
        o.set_objects(rows)
        g.set_objects_rows(rows)
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
        n = int(o.field("counters")[0])
        assert g.stats()["count"] == n, f
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
    assert n < n0, "the drain should take particles"
    assert np.any(o.field("parts")[:n, 4:7] != 0)


def test_multi_step_dam_break():
    """500 steps of a dam break with two obstacles (60 SOR iterations, omega 1.7, dt 1/120): every 50
    steps the fluid volume (live particles x particle volume) and the kinetic energy track the oracle
    within 1%, and so does the post-projection residual (divergence of the projected field minus the
    drift source, over liquid cells, 1/s: the DIVERGENCE stage run again on the field the step left).
    Positions at the end within 1e-4 of the tank scale."""
    o, g = twin.make_twin(twin.small_scene(whitewater=False, iters=60, omega=1.7), threads=8)
    for step in range(1, 501):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
        if step % 50:
            continue
        st = g.stats()
        n = int(o.field("counters")[0])
        assert st["count"] == n, step
        gp, op = g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n]
        ke_g = 0.5 * (gp[:, 4:7].astype(np.float64) ** 2).sum()
        ke_o = 0.5 * (op[:, 4:7].astype(np.float64) ** 2).sum()
        assert abs(ke_g - ke_o) <= 0.01 * ke_o, (step, ke_g, ke_o)
        g.run_stage(suiri.S_DIVERGENCE, 0, 1 / 120)
        o.kernel("divergence")
        liquid = o.field("cellType") == 1
        assert np.array_equal(g.readback(suiri.F_CELLTYPE) == 1, liquid), step
        rg = np.abs(g.readback(suiri.F_DIVG)[liquid]).max()
        ro = np.abs(o.field("divg")[liquid]).max()
        assert abs(rg - ro) <= 0.01 * ro + 1e-6, (step, rg, ro)
        print(f"step {step}: count {n} KE {ke_g:.6g} (oracle {ke_o:.6g}) max residual {rg:.3g}/s "
              f"(oracle {ro:.3g}/s) max position error {np.abs(gp[:, :3] - op[:, :3]).max():.3g}", flush=True)
    assert_close("positions@500", gp[:, :3], op[:, :3], 1e-4)


def ripple_inputs(o, g, seed=3):
    """Random (asymmetric) spectral state and a random impulse grid, identical on both sides."""
    inf = o.info()
    R, M = int(inf["R"]), int(inf["M"])
    rng = np.random.default_rng(seed)
    src = np.zeros((R + 1) * (R + 1), np.int32)
    k = min(40, src.size // 2)  # impulses (half the texels of the smallest fields)
    idx = rng.choice(src.size, k, replace=False)
    src[idx] = rng.integers(1000, 2_000_000, k)
    st = rng.normal(0, 1e-6, (M * M, 2)).astype(np.float32)
    o.field("rsrc")[:] = src
    o.field("rstate")[:] = st
    g.upload(suiri.F_RSRC, src)
    g.upload(suiri.F_RSTATE, st)
    # the mean depth comes from the particle count (50-sim.js:332); give both the same count
    vp, h = inf["vp"], inf["h"]
    count = 40000
    md = count * vp * 1000 / 1000 / (inf["nx"] * inf["nz"] * h * h)
    md = min(max(md, 0.003), inf["rim"])
    g.set_scalar("count", count)
    return md


def per_component(name, got, want, tol):
    for c in range(got.shape[1]):
        if np.abs(want[:, c]).max() > 0:
            assert_close(f"{name}[{c}]", got[:, c], want[:, c], tol)


@pytest.mark.parametrize("R", [4, 8, 64, 256, 512])
def test_ripple_transforms_and_revolve(R):
    """One unmasked frame from an arbitrary (asymmetric) state, at M = 8, 16, 128, 512 and 1024 (each M
    has its own column tiling, and below 1024 a row-transform workgroup takes several rows: all 8 or
    16 rows of the smallest fields, 8 rows of 128, 2 rows of 512): the forward transform of the source
    (bit for bit) and REVOLVE match the oracle; the
    inverse transform matches to 1e-5 of each component except where the real part holds only the
    rounding noise of a 1e5x larger imaginary part."""
    o, g = twin.make_twin(twin.small_scene(R=R))
    md = ripple_inputs(o, g)
    o.set("mask_bound", 0)
    g.set_ripple_mask(0)
    o.write_params(1 / 60)
    o.L.or_ripple_frame(o.ptr, 1 / 60, md)
    g.run_stage(suiri.S_RIPPLE, 0, 1 / 60)
    assert np.array_equal(g.readback(suiri.F_RHAT), o.field("rhat"))
    per_component("revolved state", g.readback(suiri.F_RSTATE), o.field("rstate"), TOL)
    gb, ob = g.readback(suiri.F_RFIELD), o.field("rb")
    assert_close("inverse field (velocity)", gb[:, 1], ob[:, 1], TOL)
    # real part: bounded by the conditioning, f32 eps x the imaginary part's magnitude x sqrt(M^2)
    M = int(o.info()["M"])
    bound = 2 * np.finfo(np.float32).eps * np.abs(ob[:, 1]).max() * M
    assert np.abs(gb[:, 0] - ob[:, 0]).max() <= bound
    assert np.all(g.readback(suiri.F_RSRC) == 0)


@pytest.mark.parametrize("R", [64, 512])
def test_ripple_frames_from_rest(R):
    """The reference's use: the state starts at rest and is driven by mirrored impulses, frame after
    frame, with the absorption mask, at M = 128 and 1024. Height field (texture), mask and state agree
    to 1e-5."""
    o, g = twin.make_twin(twin.small_scene(R=R))
    inf = o.info()
    R = int(inf["R"])
    vp, h = inf["vp"], inf["h"]
    count = 40000
    md = min(max(count * vp * 1000 / 1000 / (inf["nx"] * inf["nz"] * h * h), 0.003), inf["rim"])
    g.set_scalar("count", count)
    rng = np.random.default_rng(5)
    o.write_params(1 / 60)
    for f in range(6):
        src = np.zeros((R + 1) * (R + 1), np.int32)
        idx = rng.choice(src.size, 12, replace=False)
        src[idx] = rng.integers(1000, 2_000_000, 12)
        o.field("rsrc")[:] = src
        g.upload(suiri.F_RSRC, src)
        o.L.or_ripple_frame(o.ptr, 1 / 60, md)
        g.run_stage(suiri.S_RIPPLE, 0, 1 / 60)
    per_component("rtex", g.readback(suiri.F_RTEX), o.field("tex"), TOL)
    assert_close("rmask", g.readback(suiri.F_RMASK), o.field("mask"), TOL)
    per_component("rstate", g.readback(suiri.F_RSTATE), o.field("rstate"), TOL)


def test_whitewater_multiset():
    """Spray spawned by G2P and advanced by WW: the same multiset of whitewater records, bit for bit."""
    sc = twin.small_scene(whitewater=True)
    o, g = twin.make_twin(sc)
    for _ in range(25):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
    st = g.stats()
    wo = int(o.info()["ww_count"])
    assert st["ww"] == wo
    assert wo > 0
    # the same records bit for bit; only their slots differ (appends follow atomics)
    assert np.array_equal(ww_multiset(g.readback(suiri.F_WW)[:wo]), ww_multiset(o.field("ww")[:wo]))


def test_top_height_map_feeds_the_mask():
    """With top_height on, each frame's end draws the top-surface height map from the particles
    (63-wgsl-water.js:5-55) and the next frame's ripple mask reads it: the map is bit-identical to
    the oracle's after every frame, and so are the mask and the ripple state it shapes."""
    o, g = twin.make_twin(twin.small_scene(top_height=1))
    for f in range(6):
        twin.oracle_step(o, 1 / 120)
        twin.gpu_step(g, 1 / 120)
        got, want = g.readback(suiri.F_TOPH), o.field("topH")
        assert (want >= 0).sum() > 100, "the water should cover part of the map"
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"top height, frame {f}"
    assert np.array_equal(g.readback(suiri.F_RMASK).view(np.uint32), o.field("mask").view(np.uint32))
    assert_close("ripple state", g.readback(suiri.F_RSTATE), o.field("rstate"), 1e-5)
