"""Golden vectors for one ripple frame, written by the reference's NumPy RippleLayer
(blender_addon/suiri_lab/solver.py:855-907): an executable restatement of RippleLayer.encodeFrame
(src/js/64-ripple.js:222-253) that is independent of the oracle's. Its semantics match the WGSL's for
these inputs: the same constants (solver.py:11-13 against the uniform of 64-ripple.js), numpy's
complex128 transforms in place of the f32 Stockham passes, dead modes zeroed after the kick instead
of before (the kick on them is zero), and the mask given as an input (the two top-height producers
differ, so the mask is not built here).

Run where the reference is checked out (the path is the first argument); it writes
ripple_frame.npz beside this script. The reference does not travel with the fixtures.
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

R = 16                  # M = 32
LX, LZ = 0.5, 0.25      # the footprint of a 0.5 x 0.25 x 0.25 m tank at h = 1/64 (nx = 32, nz = 16)
CASES = [               # (dt, mean depth, mask kind)
    (1.0 / 120.0, 0.05, "ones"),
    (1.0 / 120.0, 0.05, "partial"),
    (1.0 / 60.0, 0.001, "partial"),   # depth below the 0.002 floor
    (0.0, 0.2, "partial"),            # dt = 0: kick and mask only
]

rng = np.random.default_rng(20261203)
out = {}
M = 2 * R
for c, (dt, depth, mk) in enumerate(CASES):
    S0 = ((rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) * 1e-4).astype(np.complex64)
    rsrc = rng.integers(-2000, 2000, size=(R + 1, R + 1)).astype(np.int32)   # impulses x 1e9, as RSRC holds them
    rsrc[rng.random((R + 1, R + 1)) < 0.6] = 0
    if mk == "ones":
        mask = np.ones((R + 1, R + 1))
    else:
        mask = np.clip(rng.random((R + 1, R + 1)) * 1.6 - 0.3, 0.0, 1.0).astype(np.float32).astype(np.float64)
    layer = solver.RippleLayer(R, LX, LZ)
    layer.S = S0.astype(np.complex128)
    layer.frame(dt, depth, rsrc.astype(np.float64) / 1e9, mask)
    out[f"c{c}_S0"] = S0
    out[f"c{c}_rsrc"] = rsrc
    out[f"c{c}_mask"] = mask
    out[f"c{c}_dt"] = np.float64(dt)
    out[f"c{c}_depth"] = np.float64(depth)
    out[f"c{c}_S1"] = layer.S
    out[f"c{c}_field"] = layer.field.copy()
out["R"] = np.int32(R)
out["L"] = np.array([LX, LZ])
np.savez_compressed(os.path.join(HERE, "ripple_frame.npz"), **out)
print("wrote", os.path.join(HERE, "ripple_frame.npz"), len(CASES), "cases")
