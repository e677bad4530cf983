"""Synthetic scenes of the benchmark workloads, in the reference's scene vocabulary (tank, items,
emitters), as configurations + object lists + fill records for the library.

* dam_break: 64x32x32 dam break (1.0 x 0.5 x 0.5 m, h = 1/64), fluid block x < 0.4 m of the tank
  length and below 0.3 m, one box obstacle 0.1 x 0.15 x 0.2 m at the tank centre, FLIP 0.95,
  SOR omega 1.7 with 60 iterations, dt = 1/120 s, ripple and whitewater off.
* faucet: 256x128x128 faucet into an obstacle field (h = 1/256), half full; sphere r 5 cm,
  cylinder r 3 cm h 20 cm, torus R 6 cm r 1.5 cm, drain r 2 cm; 0.5 L/s faucet at 1.5 m/s angled
  30 degrees down; omega 1.8, 60 iterations, dt = 1/240 s; whitewater on, ripple M = 1024.
* tank_mix: independent tanks drawn per seed, grids 32^3 .. 128x64x64 (h = 1/96), fill fractions
  0.2 .. 0.7 as a dam-break block, 0 .. 3 random primitives, 40 .. 80 SOR iterations,
  dt = 1/120 s; whitewater on, ripple M = 256.
* ripple_droplets: the droplet layer on a 128^3 tank (0.5 m cube, h = 1/256) filled 0.1 m deep with
  a sphere standing out of the water; ripple M = 1024 over the 0.5 x 0.5 m surface with the fixed depth
  0.1 m, sigma 0.072 N/m, rho 1000 kg/m^3; 4096 droplets per step, lognormal diameter (median 1.5 mm,
  sigma_ln 0.4), speed 1 .. 4 m/s; whitewater when K = We Oh^-0.4 > 60; SOR omega 1.8 x 60, dt 1/240 s.

Particles come from the reference's jittered half-cell lattice emitter (shape 2,
src/js/31-wgsl-sim2.js:295-306) with recorded seeds; velocities start at zero.
"""
import math

import numpy as np

from . import scene


def _cfg(tank, h, iters, R, cap, ww_cap, flip, omega, whitewater, reorder_every=16):
    return dict(tank_w=tank[0], tank_h=tank[1], tank_d=tank[2], h=h, iters=iters, ripple_R=R, cap=cap,
                ww_cap=ww_cap, flip=flip, omega=omega, whitewater=1 if whitewater else 0,
                reorder_every=reorder_every, reorder_phase=0)


def dam_break(steps=200):
    h = 1.0 / 64
    tank = (1.0, 0.5, 0.5)
    # cells whose centre lies in x < -0.1 (0.4 m from the wall at -0.5) and y < 0.3: 26 x 19 x 32
    block = (0, 26, 0, 19, 0, 32)
    items = [scene.make_obstacle("box", pos=[0.0, 0.25, 0.0], dims={"w": 0.1, "h": 0.15, "d": 0.2})]
    cap = 26 * 19 * 32 * 8 + 4096
    return dict(name="dam-break-64x32x32", grid=(64, 32, 32), steps=steps, frame_dt=1.0 / 120,
                cfg=_cfg(tank, h, 60, 0, cap, 1024, 0.95, 1.7, False, reorder_every=0),
                items=items, blocks=[block], seed=20240, sources=False)


def faucet(grid=(256, 128, 128), fill_layers=None, iters=60, R=512, ww=True, steps=1000):
    nx, ny, nz = grid
    h = 1.0 / 256 * (256 / nx)
    tank = (nx * h, ny * h, nz * h)
    fill = fill_layers if fill_layers is not None else ny // 2
    items = [
        scene.make_obstacle("sphere", pos=[-0.25 * tank[0], 0.2 * tank[1], -0.2 * tank[2]], dims={"r": 0.05}),
        scene.make_obstacle("cylinder", pos=[0.0, 0.1, 0.2 * tank[2]], dims={"r": 0.03, "h": 0.2}),
        scene.make_obstacle("torus", pos=[0.25 * tank[0], 0.12, -0.1 * tank[2]], dims={"R": 0.06, "r": 0.015}),
        scene.make_drain(pos=[0.4 * tank[0], 0.0, 0.3 * tank[2]], r=0.02),
        scene.make_source(pos=[-0.42 * tank[0], 0.9 * tank[1], 0.0], mode="stream", flow=30.0, speed=1.5,
                          pitch=-30, yaw=90),
    ]
    cap = nx * fill * nz * 8 + 400_000
    return dict(name="faucet-obstacles", grid=grid, steps=steps, frame_dt=1.0 / 240,
                cfg=_cfg(tank, h, iters, R, cap, 260_000, 0.95, 1.8, ww), items=items,
                blocks=[(0, nx, 0, fill, 0, nz)], seed=12345, sources=True)


MIX_NX = (32, 48, 64, 80, 96, 112, 128)
MIX_NYZ = (32, 40, 48, 56, 64)


def mix_tank(seed, steps=500):
    """One tank of the batched mix, drawn from `seed`."""
    rng = np.random.default_rng(seed)
    nx = int(rng.choice(MIX_NX))
    ny = int(rng.choice(MIX_NYZ))
    nz = int(rng.choice(MIX_NYZ))
    h = 1.0 / 96
    tank = (nx * h, ny * h, nz * h)
    f = float(rng.uniform(0.2, 0.7))
    ax = float(rng.uniform(max(f / 0.9, 0.4), 1.0))   # block length fraction; height fraction f / ax <= 0.9
    fx = max(1, int(round(ax * nx)))
    fy = max(1, min(ny - 1, int(round(f / ax * ny))))
    items = []
    for _ in range(int(rng.integers(0, 4))):
        kind = str(rng.choice(["sphere", "box", "cylinder"]))
        s = float(rng.uniform(0.06, 0.14)) * min(tank[1], tank[2])
        x = float(rng.uniform(-0.35, 0.35)) * tank[0]
        z = float(rng.uniform(-0.3, 0.3)) * tank[2]
        if kind == "sphere":
            items.append(scene.make_obstacle("sphere", pos=[x, s * 1.2, z], dims={"r": s}))
        elif kind == "box":
            items.append(scene.make_obstacle("box", pos=[x, s, z], dims={"w": 1.5 * s, "h": 2 * s, "d": s},
                                             rot=[0, float(rng.uniform(0, 90)), 0]))
        else:
            items.append(scene.make_obstacle("cylinder", pos=[x, s * 1.5, z], dims={"r": 0.6 * s, "h": 3 * s}))
    iters = int(rng.integers(40, 81))
    cap = fx * fy * nz * 8 + 4096
    return dict(name=f"mix-{nx}x{ny}x{nz}", grid=(nx, ny, nz), steps=steps, frame_dt=1.0 / 120,
                cfg=_cfg(tank, h, iters, 128, cap, 60_000, 0.95, 1.7, True),
                items=items, blocks=[(0, fx, 0, fy, 0, nz)], seed=int(seed) * 7919 + 1, sources=False, fill=f)


def tank_mix(n=64, seed=7, steps=500):
    """`n` tanks; their spatial re-orderings (every 16 frames) are staggered over the period."""
    mix = [mix_tank(seed * 100_003 + i, steps) for i in range(n)]
    for i, wl in enumerate(mix):
        wl["cfg"]["reorder_phase"] = i % 16
    return mix


def grid_of(cfg):
    """Grid of a tank configuration as the library sizes it (50-sim.js:33-45): (nx, ny, nz, h, origin)."""
    def rnd(x):  # C round(): halves away from zero
        return int(math.floor(x + 0.5))
    h = cfg["h"]
    nx, ny, nz = (max(8, rnd(cfg[k] / h)) for k in ("tank_w", "tank_h", "tank_d"))
    hh = min(cfg["tank_w"] / nx, cfg["tank_h"] / ny, cfg["tank_d"] / nz)
    return nx, ny, nz, hh, [-nx * hh / 2, 0.0, -nz * hh / 2]


def ripple_droplets(steps=2000, R=512, droplets=4096):
    """The droplet layer over a 128^3 tank filled 0.1 m deep (see the module docstring)."""
    n = 128
    h = 0.5 / n
    tank = (0.5, 0.5, 0.5)
    layers = int(round(0.1 / h))  # 26 layers = 0.1016 m of water
    items = [scene.make_obstacle("sphere", pos=[0.1, 0.06, -0.08], dims={"r": 0.07})]
    cfg = _cfg(tank, h, 60, R, n * layers * n * 8 + 200_000, 1_000_000, 0.95, 1.8, True)
    cfg.update(ripple_depth=0.1, r_sigma_rho=0.072 / 1000.0, r_rho_kick=1000.0, splash_rho=1000.0,
               splash_sigma=0.072, splash_mu=0.001, splash_k_jet=60.0, droplets=droplets, drop_seed=2024,
               drop_d50=0.0015, drop_sigma_ln=0.4, drop_vmin=1.0, drop_vmax=4.0)
    return dict(name="ripple-droplets-128^3", grid=(n, n, n), steps=steps, frame_dt=1.0 / 240, cfg=cfg,
                items=items, blocks=[(0, n, 0, layers, 0, n)], seed=777, sources=False)


def job_inputs(wl):
    """(scene items, fill records) of a workload for a serving job: the job packs the items' obstacles
    and drains at admission and emits their sources every frame (sl_job.items)."""
    nx, ny, nz, hh, origin = grid_of(wl["cfg"])
    fill = [scene.cell_block_emit(hh, origin, *blk, seed=wl["seed"] + i) for i, blk in enumerate(wl["blocks"])]
    return wl["items"], fill


def setup(sim, wl):
    """Objects and the initial fill of a workload on a Sim (or a tank of a Batch)."""
    sim.set_items(wl["items"], 0)
    recs = [scene.cell_block_emit(sim.h, sim.origin, *blk, seed=wl["seed"] + i) for i, blk in enumerate(wl["blocks"])]
    sim.emit_now(recs)


def step_bytes(nx, ny, nz, P, iters):
    """Algorithmic HBM bytes of one step: 92P + 252N + 40C + iters (24N + 18C)."""
    N = (nx + 1) * (ny + 1) * (nz + 1)
    C = nx * ny * nz
    return 92 * P + 252 * N + 40 * C + iters * (24 * N + 18 * C)
