"""Golden vectors for one G2P pass, written by the reference's NumPy solver (Solver._g2p,
blender_addon/suiri_lab/solver.py:527-631): a restatement of G2P (src/js/31-wgsl-sim2.js:149-251)
independent of the oracle's, in float64. On these inputs its semantics are the WGSL's: the masked
PIC/FLIP transfer (FLIP 0.96), advection, the tank clamp and the kill rules, the free-flight mark
(the cell count tests), drop merging with its ripple deposit and impact count, and vmax. Whitewater
is off (its spawns and crown splashes draw from another random generator) and there are no
obstacles (the port's obstacle lattice differs from the analytic distances). Survivors keep their
order.

Run where the reference is checked out (the path is the first argument); it writes g2p_pass.npz
beside this script. The reference does not travel with the fixtures.
"""
import importlib.util
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

spec = importlib.util.spec_from_file_location("solver", os.path.join(REF, "blender_addon/suiri_lab/solver.py"))
solver = importlib.util.module_from_spec(spec)
spec.loader.exec_module(solver)

TANK = (0.15, 0.12, 0.12)
DT = 1.0 / 120.0
N = 1500

rng = np.random.default_rng(31415)
s = solver.Solver(*TANK, quality="STANDARD", whitewater=False, seed=1)
nx, ny, nz = s.nx, s.ny, s.nz
height = rng.integers(ny // 4, ny // 2, size=(nx, nz))
count = np.zeros((nx, ny, nz), np.int64)
for i in range(nx):
    for k in range(nz):
        hgt = height[i, k]
        count[i, :hgt, k] = rng.integers(4, 12, size=hgt)
        count[i, hgt, k] = rng.integers(0, 4)
        if rng.random() < 0.3:                               # sparse drops above the surface
            count[i, min(hgt + 2 + rng.integers(0, 4), ny - 1), k] = rng.integers(1, 3)
velg = (rng.standard_normal((3,) + s.ns) * 0.4).astype(np.float32)
velold = (velg + rng.standard_normal((3,) + s.ns).astype(np.float32) * 0.05).astype(np.float32)
mask = rng.random((3,) + s.ns) < 0.85
lo, hi = s.origin, s.origin + np.array([nx, ny, nz]) * s.h
pos = lo + rng.random((N, 3)) * (hi - lo)
pos[: N // 5, 1] = s.rim + rng.random(N // 5) * 0.03              # above the grid: ballistic
pos[N // 5: N // 5 + 40, 0] = hi[0] + 0.001 + rng.random(40) * 0.004   # beside the tank
pos[N // 5 + 40: N // 5 + 60, 1] = 0.0015                            # beside and near the floor: killed
pos[N // 5 + 40: N // 5 + 60, 0] = lo[0] - 0.003
pos[N // 5 + 60: N // 5 + 160, 0] = lo[0] + 0.0005                   # against a wall
vel = rng.standard_normal((N, 3)) * 0.6
vel[: N // 2, 1] = -0.35 - rng.random(N // 2) * 2.5               # falling
air = (rng.random(N) < 0.5).astype(np.float64)
pos, vel = (a.astype(np.float32).astype(np.float64) for a in (pos, vel))
# the counts hold every in-grid particle in its own cell, as P2G's do (an empty own cell would make the
# WGSL's u32 `countAt(c1) - 1` wrap where NumPy's int64 gives -1)
g = (pos - s.origin) / s.h
inside = np.all((g >= 0) & (g < [nx, ny, nz]), axis=1)
c = np.floor(g[inside]).astype(np.int64)
np.add.at(count, (c[:, 0], c[:, 1], c[:, 2]), 1)

s.count = count
s.velg = velg.astype(np.float64)
s.velold = velold.astype(np.float64)
s.mask = mask
s.pos, s.vel, s.air = pos.copy(), vel.copy(), air.copy()
s._g2p(DT)
nzi = np.nonzero(s.rsrc)
np.savez_compressed(
    os.path.join(HERE, "g2p_pass.npz"),
    tank=np.array(TANK), dt=np.float64(DT), grid=np.array([nx, ny, nz]), h=np.float64(s.h),
    count=count.astype(np.uint8), velg=velg, velold=velold, mask=mask, pos=pos, vel=vel, air=air,
    out_pos=s.pos, out_vel=s.vel, out_air=s.air, impacts=np.int64(s.impacts), vmax=np.float64(s.vmax),
    rsrc_j=nzi[0].astype(np.int32), rsrc_i=nzi[1].astype(np.int32), rsrc_v=s.rsrc[nzi])
print("wrote", os.path.join(HERE, "g2p_pass.npz"), len(s.pos), "survivors of", N, "impacts", s.impacts,
      "airborne", int(s.air.sum()), "deposits", len(nzi[0]))
