"""Golden vectors for one whole substep, written by the reference's NumPy solver (Solver._substep,
blender_addon/suiri_lab/solver.py:438-525, then Solver._g2p): a restatement of the WGSL chain P2G,
NORMALIZE, FORCES, DIVERGENCE, PRESSURE (red-black SOR), PROJECT and G2P (src/js/30-wgsl-sim.js,
31-wgsl-sim2.js) independent of the oracle's, with its own constants (G, PPC, DRIFT, OMEGA,
solver.py:11-17), iteration count (QUALITY['STANDARD'], solver.py:25) and velocity clamp. It differs
from the WGSL only in rounding: float64 weights where P2G sums fixed point, an SOR update written as
p += w((sum - rhs)/den - p) in f32, float64 elsewhere. Whitewater is off and there are no obstacles
(GEOM is then the empty tank in both). The SOR starts from a warm pressure in the liquid cells.

Run where the reference is checked out (the path is the first argument); it writes substep.npz
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

TANK = (0.1, 0.08, 0.08)
DT = 1.0 / 180.0

rng = np.random.default_rng(271828)
s = solver.Solver(*TANK, quality="STANDARD", whitewater=False, seed=1)
nx, ny, nz, h = s.nx, s.ny, s.nz, s.h
# a sloshing body: a ragged surface, some crowded cells (the drift term), a few drops above
height = rng.integers(3, 6, size=(nx, nz))
P = []
for i in range(nx):
    for k in range(nz):
        for j in range(height[i, k] + 1):
            n = int(rng.integers(5, 9)) if j < height[i, k] else int(rng.integers(1, 5))
            if rng.random() < 0.12:
                n = int(rng.integers(11, 16))                      # crowded: count/PPC - 1 > 0
            P.append(np.array([i, j, k]) + rng.random((n, 3)))
g = np.concatenate(P)
drops = np.column_stack([rng.random(40) * nx, 7 + rng.random(40) * 1.8, rng.random(40) * nz])
g = np.concatenate([g, drops])
pos = s.origin + g * h
N = len(pos)
x = (g / np.array([nx, ny, nz]))
vel = np.column_stack([0.3 * np.sin(3 * x[:, 1] + 2 * x[:, 2]), 0.2 * np.cos(4 * x[:, 0]) - 0.1,
                       -0.25 * np.sin(5 * x[:, 0] - x[:, 1])]) + rng.standard_normal((N, 3)) * 0.03
vel[-40:, 1] = -0.5 - rng.random(40)                               # the drops fall
vel[:5] = [[40.0, -35.0, 31.0]] * 5                                # beyond the +-30 clamp
air = np.zeros(N)
air[-40:] = 1.0
pos, vel = (a.astype(np.float32).astype(np.float64) for a in (pos, vel))
# warm start: roughly hydrostatic in the liquid cells, zero elsewhere
jj = np.arange(ny)[None, :, None]
pres = (solver.G * DT * h * np.maximum(height[:, None, :] - jj, 0) * 0.9
        + rng.standard_normal((nx, ny, nz)) * 1e-5).astype(np.float32)
s.pos, s.vel, s.air = pos.copy(), vel.copy(), air.copy()
s.pres = pres.astype(np.float64)
s._substep(DT, [])
nzi = np.nonzero(s.rsrc)
np.savez_compressed(
    os.path.join(HERE, "substep.npz"),
    tank=np.array(TANK), dt=np.float64(DT), grid=np.array([nx, ny, nz]), h=np.float64(h), iters=np.int32(s.iters),
    pos=pos.astype(np.float32), vel=vel.astype(np.float32), air=air.astype(np.uint8), pres=pres,
    out_pos=s.pos.astype(np.float32), out_vel=s.vel.astype(np.float32), out_air=s.air.astype(np.uint8),
    out_pres=s.pres.astype(np.float32), impacts=np.int64(s.impacts), vmax=np.float64(s.vmax),
    rsrc_j=nzi[0].astype(np.int32), rsrc_i=nzi[1].astype(np.int32), rsrc_v=s.rsrc[nzi])
print("wrote", os.path.join(HERE, "substep.npz"), N, "particles,", len(s.pos), "survive; impacts", s.impacts,
      "fluid cells", int((s.count > 0).sum()), "crowded", int((s.count > 8).sum()), "vmax", s.vmax)
