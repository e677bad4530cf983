"""Golden vectors for the ripple mask, written by the reference's NumPy solver (Solver._ripple_mask,
blender_addon/suiri_lab/solver.py:781-794), a restatement of RMASKBUILD (src/js/64-ripple.js:106-126)
independent of the oracle's: per ripple node, the top-height texel under it, dry -> 0.75, wet ->
smoothstep(-0.003, 0.006) of the obstacle distance 2 mm below the surface. The two solvers draw
different top-height maps, so the map is an input here, and their obstacle models differ, so the
obstacle is one whose distance both give exactly: a large box whose top face (y = 0.04) is the only
surface near the sampled points, given to the port as a lattice holding that plane's distance
(trilinear interpolation reproduces a plane).

Run where the reference is checked out (the path is the first argument); it writes ripple_mask.npz
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

TANK = (0.4, 0.2, 0.3)
R, W = 128, 199
FACE = 0.04   # top face of the box: centre (0, 0, 0), half extents (1.0, 0.04, 1.0)

s = solver.Solver(*TANK, quality="STANDARD", whitewater=False, seed=1)
s.ripple = solver.RippleLayer(R, s.nx * s.h, s.nz * s.h)
shape = s.lattice_shape()
ys = s.origin[1] + 0.5 * s.h * np.arange(shape[1])
sdf = np.broadcast_to((ys - FACE)[None, :, None], shape).copy()
s.set_solids(solver.Solids(sdf, np.zeros(shape, np.int64), [(np.zeros(3), np.zeros(3), np.zeros(3))], []))
rng = np.random.default_rng(777)
top = (FACE + 0.002 + rng.uniform(-0.012, 0.018, (W, W))).astype(np.float32).astype(np.float64)
top[rng.random((W, W)) < 0.25] = -1.0
s.top = top
mask = s._ripple_mask()
np.savez_compressed(os.path.join(HERE, "ripple_mask.npz"), tank=np.array(TANK), R=np.int32(R), W=np.int32(W),
                    face=np.float64(FACE), top=top.astype(np.float32), mask=mask)
band = np.mean((mask > 0.001) & (mask < 0.999) & (mask != 0.75))
print("wrote", os.path.join(HERE, "ripple_mask.npz"), mask.shape, "in the band:", round(float(band), 3))
