"""Golden vectors for one whitewater pass, written by the reference's NumPy solver
(Solver._whitewater, blender_addon/suiri_lab/solver.py:675-713): an executable restatement of the WW
kernel (src/js/31-wgsl-sim2.js:361-409) that is independent of the oracle's. On these inputs its
semantics are the WGSL's: no obstacles (its lattice distance is then 1e9, so the collision and the
drain test do nothing, as with an empty object list), classification from the cell counts, the
landing deposit into the ripple source, spray / foam / bubble motion, the tank clamp and the
survival rule. It runs in complex128/float64 where the WGSL runs f32.

Run where the reference is checked out (the path is the first argument); it writes ww_pass.npz
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

TANK = (0.2, 0.12, 0.15)
DT = 1.0 / 60.0
N = 400

rng = np.random.default_rng(4242)
s = solver.Solver(*TANK, quality="STANDARD", whitewater=True, seed=1)
nx, ny, nz = s.nx, s.ny, s.nz
# a water body with a ragged surface: full cells below a height field, a partial cell on top, holes
height = rng.integers(ny // 3, 2 * ny // 3, size=(nx, nz))
count = np.zeros((nx, ny, nz), np.int64)
for i in range(nx):
    for k in range(nz):
        hgt = height[i, k]
        count[i, :hgt, k] = 8
        count[i, hgt, k] = rng.integers(1, 8)
count[rng.random(count.shape) < 0.05] = 0
velg = (rng.standard_normal((3,) + s.ns) * 0.4).astype(np.float32)
# whitewater: inside the water, around the surface, in the air above the rim, outside the footprint
lo, hi = s.origin, s.origin + np.array([nx, ny, nz]) * s.h
pos = lo + rng.random((N, 3)) * (hi - lo)
pos[: N // 4, 1] = s.rim + rng.random(N // 4) * 0.05                       # above the grid
pos[N // 4: N // 4 + 20, 0] = hi[0] + 0.002 + rng.random(20) * 0.01       # beside the tank
pos[N // 4 + 20: N // 4 + 30, 1] = 0.0005                                   # near the floor
vel = rng.standard_normal((N, 3)) * 0.5
vel[: N // 3, 1] = -0.3 - rng.random(N // 3) * 1.5                          # falling: landing deposits
life = 0.01 + rng.random(N) * 1.5
life[::10] = 0.004 + rng.random(len(life[::10])) * 0.01                   # some expire this pass
kind = rng.integers(0, 3, N).astype(np.float64)
pos, vel, life = (a.astype(np.float32).astype(np.float64) for a in (pos, vel, life))

s.count = count
s.velg = velg.astype(np.float64)
s.wpos, s.wvel, s.wlife, s.wkind = pos.copy(), vel.copy(), life.copy(), kind.copy()
s._whitewater(DT)
nzi = np.nonzero(s.rsrc)
np.savez_compressed(
    os.path.join(HERE, "ww_pass.npz"),
    tank=np.array(TANK), dt=np.float64(DT), grid=np.array([nx, ny, nz]), h=np.float64(s.h),
    ripple_R=np.int32(s.ripple.R), count=count.astype(np.uint8), velg=velg,
    pos=pos, vel=vel, life=life, kind=kind,
    out_pos=s.wpos, out_vel=s.wvel, out_life=s.wlife, out_kind=s.wkind,
    rsrc_j=nzi[0].astype(np.int32), rsrc_i=nzi[1].astype(np.int32), rsrc_v=s.rsrc[nzi])
print("wrote", os.path.join(HERE, "ww_pass.npz"), len(s.wpos), "survivors of", N, "deposits", len(nzi[0]))
