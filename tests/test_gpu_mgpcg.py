"""The converged pressure solve (SOLVER_MGPCG; no counterpart in the reference): its pressure against
a direct sparse solve of the SOR system assembled from the GPU's own buffers, the post-projection
residual of whole steps, and the reference solver after switching back."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import importlib

import twin
from twin import suiri

residual = importlib.import_module("hahenn-_amd.residual")

pytestmark = pytest.mark.gpu


def sor_system(g):
    """The system of the PRESSURE pass (src/js/30-wgsl-sim.js:267-297) from readbacks: unknowns are
    liquid cells with an open face; den p_c - sum f p_n = -h^2 div over open faces to liquid cells,
    with den the exact sum of the open fractions and div DIVERGENCE's formula in double precision."""
    nx, ny, nz, h = g.nx, g.ny, g.nz, g.h
    fr = g.readback(suiri.F_FRAC).reshape(nz + 1, ny + 1, nx + 1, 4).transpose(2, 1, 0, 3).astype(np.float64)
    liquid = (g.readback(suiri.F_CELLTYPE) == 1).reshape(nz, ny, nx).transpose(2, 1, 0)
    div = residual.divergence_f64(g)[0].transpose(2, 1, 0)
    pos = lambda a: np.where(a > 0, a, 0.0)
    fx, fy, fz = pos(fr[:, :ny, :nz, 0]), pos(fr[:nx, :, :nz, 1]), pos(fr[:nx, :ny, :, 2])
    den = fx[:-1] + fx[1:] + fy[:, :-1] + fy[:, 1:] + fz[:, :, :-1] + fz[:, :, 1:]
    act = liquid & (den > 0)
    idx = -np.ones(act.shape, np.int64)
    idx[act] = np.arange(act.sum())
    rows, cols, vals = [idx[act]], [idx[act]], [den[act]]
    for a, f in ((0, fx), (1, fy), (2, fz)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, -1), slice(1, None)
        w = np.take(f, range(1, act.shape[a]), axis=a)  # the faces between cells lo and hi along axis a
        both = act[tuple(lo)] & act[tuple(hi)] & (w > 0)
        i0, i1 = idx[tuple(lo)][both], idx[tuple(hi)][both]
        rows += [i0, i1]
        cols += [i1, i0]
        vals += [-w[both], -w[both]]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(int(act.sum()),) * 2)
    b = -h * h * div[act]
    return A, b, act


def run_to_state(sc, frames):
    o, g = twin.make_twin(sc)
    for _ in range(frames):
        twin.gpu_step(g, 1 / 120)
    return g


def test_converged_pressure_matches_a_direct_solve():
    g = run_to_state(twin.small_scene(whitewater=False, iters=20), 30)
    g.set_pressure_solver(suiri.SOLVER_MGPCG, 1e-7, 60)
    for st in (suiri.S_P2G, suiri.S_NORMFORCE, suiri.S_DIVERGENCE, suiri.S_SOR):
        g.run_stage(st, 0, 1 / 120)
    stats = g.pressure_solver_stats()
    A, b, act = sor_system(g)
    want = spl.spsolve(A.tocsc(), b)
    got = g.readback(suiri.F_PRES)
    # the pressure readback is in the reference's cell layout
    got = got.reshape(g.nz, g.ny, g.nx).transpose(2, 1, 0)[act].astype(np.float64)
    print(stats, "unknowns", act.sum(), "max |p|", np.abs(want).max(), "max err", np.abs(got - want).max())
    assert not stats["breakdown"] and 1 <= stats["iters"] <= 60
    assert stats["resid_per_s"] <= 1e-7 < stats["resid0_per_s"]
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


def post_projection_residual(g):
    """Over liquid cells with an open face: |div| by the DIVERGENCE stage (single precision), and by
    the same formula in double precision less the bound of storing face velocities in f32."""
    g.run_stage(suiri.S_DIVERGENCE, 0, 1 / 120)
    div, liquid, has_open, floor = residual.divergence_f64(g)
    sel = (liquid & has_open).ravel()
    f32 = np.abs(g.readback(suiri.F_DIVG)[sel].astype(np.float64))
    above = np.maximum(np.abs(div.ravel()[sel]) - floor.ravel()[sel], 0.0)
    return f32, above


def test_converged_steps_leave_a_small_residual():
    """A dam break with obstacles, 40 steps with each solver from the same state: the reference's 8
    SOR sweeps leave a residual above 1e-2/s; the converged solve one near the f32 spacing of the
    velocities, and in double precision within its tolerance of the bound that storing the face
    velocities in f32 sets; both runs keep their particles."""
    sc = twin.small_scene(tank=(0.5, 0.25, 0.25), h=1 / 96, whitewater=False, iters=8, fill_cells=(18, 15, 24),
                          cap=200000)
    res = {}
    for solver in (suiri.SOLVER_SOR, suiri.SOLVER_MGPCG):
        o, g = twin.make_twin(sc)
        g.set_pressure_solver(solver, 1e-6, 40)
        n0 = g.stats()["count"]
        worst = above_worst = 0.0
        for step in range(40):
            twin.gpu_step(g, 1 / 120)
            if step % 10 == 9:
                r, above = post_projection_residual(g)
                worst = max(worst, r.max())
                if solver == suiri.SOLVER_MGPCG:
                    st = g.pressure_solver_stats()
                    assert not st["breakdown"] and st["resid_per_s"] <= 1e-6
                    above_worst = max(above_worst, above.max())
        res[solver] = (worst, g.stats()["count"] / n0, above_worst)
        if solver == suiri.SOLVER_MGPCG:
            st = g.pressure_solver_stats()
            print("mgpcg", st)
            assert not st["breakdown"] and st["resid_per_s"] <= 1e-6
        g.close()
    print("worst residual per solver (1/s):", res)
    assert res[suiri.SOLVER_SOR][1] >= 0.999 and res[suiri.SOLVER_MGPCG][1] >= 0.999
    assert res[suiri.SOLVER_SOR][0] > 1e-2
    assert res[suiri.SOLVER_MGPCG][0] <= 1e-3
    assert res[suiri.SOLVER_MGPCG][2] <= 2e-6


def test_switching_back_gives_the_reference_sweeps():
    """After converged steps, SOLVER_SOR runs the reference's sweeps again: from identical state the
    next step is bit-identical to the oracle's."""
    o, g = twin.make_twin(twin.small_scene(whitewater=False, iters=20))
    g.set_pressure_solver(suiri.SOLVER_MGPCG, 1e-6, 30)
    for _ in range(5):
        twin.gpu_step(g, 1 / 120)
    g.set_pressure_solver(suiri.SOLVER_SOR)
    # hand the GPU's state to the oracle, then one step on each
    n = g.stats()["count"]
    o.field("parts")[:] = 0
    o.field("parts")[:n] = g.readback(suiri.F_PARTICLES)[:n]
    o.set("count", n)
    o.field("counters")[0] = n
    o.field("pres")[:] = g.readback(suiri.F_PRES)
    for name in ("time", "frame", "vmax"):
        o.set(name, g.scalar(name))
    twin.oracle_step(o, 1 / 120)
    twin.gpu_step(g, 1 / 120)
    assert g.stats()["count"] == int(o.field("counters")[0])
    assert np.array_equal(g.readback(suiri.F_PARTICLES)[:n], o.field("parts")[:n])
    assert np.array_equal(g.readback(suiri.F_PRES), o.field("pres"))


def test_solver_choice_errors():
    o, g = twin.make_twin(twin.small_scene(whitewater=False))
    with pytest.raises(suiri.SuiriError):
        g.set_pressure_solver(7)
    with pytest.raises(suiri.SuiriError):
        g.set_pressure_solver(suiri.SOLVER_MGPCG, -1.0, 10)
    with pytest.raises(suiri.SuiriError):
        g.set_pressure_solver(suiri.SOLVER_MGPCG, 1e-5, 0)
    with pytest.raises(suiri.SuiriError):
        g.pressure_solver_stats()   # not chosen yet


def test_single_level_grid_solves_directly():
    """A grid of at most 1024 cells is its own coarsest level: the preconditioner is the one-workgroup
    symmetric sweeps alone. Its pressure still agrees with a direct solve."""
    sc = twin.small_scene(tank=(0.125, 0.125, 0.125), h=1 / 64, iters=20, whitewater=False, fill_cells=(8, 4, 8),
                          cap=20000)
    g = run_to_state(sc, 20)
    assert g.nx * g.ny * g.nz <= 1024
    g.set_pressure_solver(suiri.SOLVER_MGPCG, 1e-7, 200)
    for st in (suiri.S_P2G, suiri.S_NORMFORCE, suiri.S_DIVERGENCE, suiri.S_SOR):
        g.run_stage(st, 0, 1 / 120)
    stats = g.pressure_solver_stats()
    A, b, act = sor_system(g)
    want = spl.spsolve(A.tocsc(), b)
    got = g.readback(suiri.F_PRES).reshape(g.nz, g.ny, g.nx).transpose(2, 1, 0)[act].astype(np.float64)
    print(stats, "unknowns", act.sum(), "max err", np.abs(got - want).max())
    assert act.sum() > 50 and not stats["breakdown"] and stats["resid_per_s"] <= 1e-7
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


def test_batch_tanks_refuse_the_converged_solve():
    cfg = suiri.config("standard", tank_w=0.25, tank_h=0.125, tank_d=0.125, h=1 / 32, cap=20000, ww_cap=2000,
                       ripple_R=0)
    b = suiri.Batch([cfg])
    with pytest.raises(suiri.SuiriError):
        b.tanks[0].set_pressure_solver(suiri.SOLVER_MGPCG, 1e-5, 20)
    b.tanks[0].set_pressure_solver(suiri.SOLVER_SOR)
    b.close()
