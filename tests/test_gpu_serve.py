"""Serving loop: device-memory accounting, admission against the budget, cancellation, and the
result of a job being the same alone and in a mix of other jobs (continuous batching)."""
import importlib

import numpy as np
import pytest

suiri = importlib.import_module("hahenn-_amd")
scene = importlib.import_module("hahenn-_amd.scene")

pytestmark = pytest.mark.gpu


def tank_cfg(w, h, d, cell, iters=20, R=32, reorder=3, cap=30000):
    return suiri.config("standard", tank_w=w, tank_h=h, tank_d=d, h=cell, iters=iters, ripple_R=R, cap=cap,
                       ww_cap=20000, flip=0.95, omega=1.7, whitewater=1, reorder_every=reorder)


def fill_block(cfg, i1, j1, k1, seed):
    nx = int(round(cfg.tank_w / cfg.h))
    nz = int(round(cfg.tank_d / cfg.h))
    hh = min(cfg.tank_w / nx, cfg.tank_h / int(round(cfg.tank_h / cfg.h)), cfg.tank_d / nz)
    origin = [-nx * hh / 2, 0.0, -nz * hh / 2]
    return [scene.cell_block_emit(hh, origin, 0, i1, 0, j1, 0, k1, seed=seed)]


def obstacle_rows(x):
    items = [dict(scene.make_obstacle("sphere", pos=[x, 0.05, 0.0], dims={"r": 0.03}), id=1)]
    rows, _ = scene.Scene().pack(items, 0)
    return rows


def drops(dt, t):
    """A falling ball of water every third frame, seeded by the frame index."""
    f = int(round(t / dt))
    if f % 3:
        return []
    return [dict(a=[0.05, 0.2, 0.0, 0.02], b=[0.0, -0.5, 0.0, 1.0], shape=1, count=250, seed=1000 + f)]


J = dict(cfg=lambda: tank_cfg(0.375, 0.25, 0.25, 1 / 64), rows=obstacle_rows(-0.05), steps=12, dt=1 / 120)


def submit_j(srv):
    cfg = J["cfg"]()
    return srv.submit(cfg, rows=J["rows"], fill=fill_block(cfg, 10, 8, 16, 77), steps=J["steps"], dt=J["dt"],
                      source=drops, keep=True)


def run_until_idle(srv, arrivals=(), limit=400):
    """Ticks until nothing is queued or running; `arrivals` = [(tick, fn)] submits at given ticks."""
    pending = sorted(arrivals, key=lambda a: a[0])
    for tick in range(limit):
        while pending and pending[0][0] <= tick:
            pending.pop(0)[1]()
        srv.tick()
        st = srv.stats()
        if not pending and st["queued"] == 0 and st["running"] == 0:
            return tick
    raise AssertionError("server did not drain")


def state_of(srv, jid):
    sim = srv.job_sim(jid)
    st = sim.stats()
    n = st["count"]
    out = {"count": n, "frames": st["frames"], "time": st["time"], "vmax": st["vmax"]}
    out["particles"] = sim.readback(suiri.F_PARTICLES)[:n].view(np.uint32).copy()
    for f in (suiri.F_VEL, suiri.F_PRES, suiri.F_RSTATE, suiri.F_RTEX):
        out[f] = sim.readback(f).view(np.uint32).copy()
    ww = sim.readback(suiri.F_WW)[: st["ww"]]
    out["ww"] = ww[np.lexsort(ww.T[::-1])].view(np.uint32).copy()
    return out


def test_job_bytes_is_what_creation_allocates():
    for cfg in (tank_cfg(0.375, 0.25, 0.25, 1 / 64), tank_cfg(0.5, 0.3, 0.2, 1 / 50, R=0, reorder=0),
                tank_cfg(0.3, 0.3, 0.3, 1 / 80, R=128, cap=100000)):
        sim = suiri.Sim(cfg)
        assert sim.scalar("bytes") == suiri.job_bytes(cfg)
        sim.close()


@pytest.mark.parametrize("frames_per_tick", [1, 4])
def test_job_alone_equals_job_in_mix(frames_per_tick):
    """Bit-identical results alone and in a mix, also when ticks enqueue several frames back to back
    (the mix's jobs end at 9, 7, 12 and 5 frames, so bursts stop at every job's last frame)."""
    print(f"server: job alone (mix at {frames_per_tick} frames per tick)", flush=True)
    alone = suiri.Server(budget=4 << 30, max_jobs=4, autotune_after=0)
    ja = submit_j(alone)
    run_until_idle(alone)
    assert alone.info(ja)["status"] == "done"
    want = state_of(alone, ja)
    assert want["frames"] == J["steps"]
    alone.close()

    print("server: job in a mix", flush=True)
    cfg_a = tank_cfg(0.4, 0.2, 0.2, 1 / 80, iters=30, R=64)          # 32x16x16
    cfg_b = tank_cfg(0.5, 0.3, 0.3, 1 / 64, iters=40, R=0, reorder=0)  # 32x19x19
    cfg_c = tank_cfg(0.25, 0.25, 0.25, 1 / 64, iters=24)
    budget = suiri.job_bytes(J["cfg"]()) + suiri.job_bytes(cfg_a) + suiri.job_bytes(cfg_c) + (1 << 20)
    mix = suiri.Server(budget=budget, max_jobs=3, autotune_after=3, frames_per_tick=frames_per_tick)
    ids = {}
    arrivals = [
        (0, lambda: ids.__setitem__("a", mix.submit(cfg_a, fill=fill_block(cfg_a, 20, 8, 16, 5), steps=9,
                                                     dt=1 / 240, source=drops))),
        (0, lambda: ids.__setitem__("b", mix.submit(cfg_b, rows=obstacle_rows(0.1), fill=fill_block(cfg_b, 16, 10, 19, 6),
                                                     steps=7, dt=1 / 100))),
        (1, lambda: ids.__setitem__("j", submit_j(mix))),
        (2, lambda: ids.__setitem__("c", mix.submit(cfg_c, fill=fill_block(cfg_c, 8, 8, 16, 7), steps=50,
                                                     dt=1 / 120))),
        (5, lambda: mix.cancel(ids["c"])),
        (6, lambda: ids.__setitem__("d", mix.submit(cfg_a, fill=fill_block(cfg_a, 12, 6, 16, 8), steps=5,
                                                     dt=1 / 120))),
    ]
    run_until_idle(mix, arrivals)
    info = {k: mix.info(v) for k, v in ids.items()}
    assert info["j"]["status"] == "done" and info["j"]["frames"] == J["steps"]
    assert info["c"]["status"] == "cancelled"
    assert all(info[k]["status"] == "done" for k in ("a", "b", "d"))
    assert info["j"]["t_submit"] <= info["j"]["t_admit"] <= info["j"]["t_done"]
    got = state_of(mix, ids["j"])
    for k in ("count", "frames", "time", "vmax"):
        assert got[k] == want[k], k
    for k in want:
        if isinstance(want[k], np.ndarray):
            assert np.array_equal(got[k], want[k]), f"field {k} differs alone vs in the mix"
    st = mix.stats()
    assert st["done"] == 4 and st["used"] == suiri.job_bytes(J["cfg"]())   # only the kept job stays resident
    mix.release(ids["j"])
    assert mix.stats()["used"] == 0
    mix.close()


def test_budget_refusal_queueing_and_cancel():
    small = tank_cfg(0.25, 0.25, 0.25, 1 / 64, iters=10)
    big = tank_cfg(0.6, 0.4, 0.4, 1 / 100, cap=2_000_000)
    budget = suiri.job_bytes(small) * 2 + (1 << 16)
    srv = suiri.Server(budget=budget, max_jobs=8, autotune_after=0)
    jb = srv.submit(big, steps=5)
    assert srv.info(jb)["status"] == "refused"
    ids = [srv.submit(small, fill=fill_block(small, 6, 6, 16, i), steps=4) for i in range(5)]
    srv.tick()
    st = srv.stats()
    assert st["running"] == 2 and st["queued"] == 3 and st["used"] <= budget   # the budget holds two
    srv.cancel(ids[0])
    srv.cancel(ids[4])   # still queued
    srv.tick()
    assert srv.info(ids[0])["status"] == "cancelled" and srv.info(ids[4])["status"] == "cancelled"
    run_until_idle(srv)
    done = [i for i in ids if srv.info(i)["status"] == "done"]
    assert sorted(done) == sorted(ids[1:4])
    statuses = {d["id"]: d["status"] for d in srv.poll()}
    assert statuses[jb] == "refused" and statuses[ids[0]] == "cancelled"
    assert srv.stats()["used"] == 0
    srv.close()
