"""Bake output of a tank's frame (the reference's Blender cache layout, bake.py:226-249): the npz
arrays are the tank's live particles and whitewater in Blender axes, the surface EXR holds the
top-height map and the ripple slopes at its texel centres; server jobs write them at admission,
every bake_every frames and at their last frame."""
import importlib
import os

import numpy as np
import pytest

import exrspec
import twin

suiri = importlib.import_module("hahenn-_amd")

pytestmark = pytest.mark.gpu


def to_blender(p):
    return np.stack([p[:, 0], -p[:, 2], p[:, 1]], 1).astype(np.float32)


def test_bake_frame_matches_the_tank(tmp_path):
    print("bake: small scene, top height on", flush=True)
    _, g = twin.make_twin(twin.small_scene(top_height=1))
    for _ in range(8):
        g.step(1 / 120)
    st = g.stats()
    g.bake(str(tmp_path), 8)
    z = np.load(tmp_path / "frame_00008.npz")
    parts = g.readback(suiri.F_PARTICLES)[: st["count"]]
    live = parts[:, 3] >= 0.5
    assert np.array_equal(z["pos"], to_blender(parts[live, :3]))
    assert np.array_equal(z["vel"], to_blender(parts[live, 4:7]))
    ww = g.readback(suiri.F_WW)[: st["ww"]]
    for k, name in enumerate(("spray", "foam", "bubble")):
        assert np.array_equal(z[name], to_blender(ww[np.rint(ww[:, 7]) == k, :3]))
    assert int(z["particles"]) == st["count"]
    assert float(z["time"]) == np.float32(st["time"])
    assert abs(float(z["liters"]) - st["count"] * g.vp * 1000) <= 1e-6 * max(1.0, st["count"] * g.vp * 1000)
    _, chans, planes = exrspec.read(str(tmp_path / "surface_00008.exr"))
    top = g.readback(suiri.F_TOPH)
    W = int(np.sqrt(top.size))
    assert np.array_equal(planes["R"], top.reshape(W, W))
    # slopes: bilinear in the ripple vertices at the texel centres (solver.py:909-919)
    R = g.grid.R
    tex = g.readback(suiri.F_RTEX).reshape(R + 1, R + 1, 4).astype(np.float64)
    q = np.clip((np.arange(W) + 0.5) / W * R, 0, R)
    b = np.minimum(np.floor(q).astype(int), R - 1)
    f = q - b
    for c, plane, sign in ((1, "G", 1.0), (2, "B", -1.0)):
        F = tex[..., c]
        bz, bx, fz, fx = b[:, None], b[None, :], f[:, None], f[None, :]
        want = (F[bz, bx] * (1 - fx) + F[bz, bx + 1] * fx) * (1 - fz) + (F[bz + 1, bx] * (1 - fx) + F[bz + 1, bx + 1] * fx) * fz
        assert np.abs(planes[plane] - sign * want).max() <= 1e-6 * max(1.0, np.abs(want).max())


def test_server_jobs_write_bake_frames(tmp_path):
    print("bake: server job every 2 frames", flush=True)
    sc = twin.small_scene(R=32)
    cfg = suiri.config("standard", tank_w=sc["tank"][0], tank_h=sc["tank"][1], tank_d=sc["tank"][2], h=sc["h"],
                       iters=20, ripple_R=32, cap=sc["cap"], ww_cap=sc["ww_cap"], top_height=1)
    srv = suiri.Server(budget=2 << 30, max_jobs=2, autotune_after=0)
    sim_h = 0.25 / round(0.25 / sc["h"])
    origin = [-round(0.5 / sc["h"]) * sim_h / 2, 0.0, -round(0.25 / sc["h"]) * sim_h / 2]
    fill = [twin.scene.cell_block_emit(sim_h, origin, 0, 10, 0, 8, 0, 12, seed=3)]
    jid = srv.submit(cfg, fill=fill, steps=5, dt=1 / 120, bake_dir=str(tmp_path), bake_every=2)
    for _ in range(20):
        srv.tick()
        if srv.info(jid)["status"] == "done":
            break
    assert srv.info(jid)["status"] == "done"
    names = sorted(os.listdir(tmp_path))
    assert names == ["frame_%05d.npz" % f for f in (0, 2, 4, 5)] + ["surface_%05d.exr" % f for f in (0, 2, 4, 5)]
    z = np.load(tmp_path / "frame_00005.npz")
    assert int(z["particles"]) == srv.info(jid)["count"] and z["pos"].shape == (int(z["particles"]), 3)
    srv.close()
