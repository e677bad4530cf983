"""Golden counts for the crown splash, written by the reference's NumPy solver (Solver._crown,
blender_addon/suiri_lab/solver.py:632-657), a restatement of splashCrown (src/js/31-wgsl-sim2.js:
119-148) independent of the oracle's. Its random draws come from another generator, so only what
they do not touch is recorded: for each (U, m) the ring droplets (life != 0.7) and jets (life 0.7)
appended, and the drop diameter D the ring radius 0.6 D follows from.

Run where the reference is checked out (the path is the first argument); it writes crown.npz beside
this script. The reference does not travel with the fixtures.
"""
import importlib.util
import math
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

spec = importlib.util.spec_from_file_location("solver", os.path.join(REF, "blender_addon/suiri_lab/solver.py"))
solver = importlib.util.module_from_spec(spec)
spec.loader.exec_module(solver)

rows = []
for h in (1 / 256, 1 / 128, 0.0065, 0.0085, 1 / 64, 0.012, 0.02):
    m = float(np.float32(998.0 * (0.5 * h) ** 3))       # the particle mass G2P passes
    for U in np.linspace(0.2, 12.0, 60):
        U = float(np.float32(U))
        s = solver.Solver(0.2, 0.12, 0.15, quality="LOW", whitewater=True, seed=3)
        s._crown(np.array([0.01, 0.05, -0.02]), U, m)
        life = s.wlife
        nj = int(np.sum(life == 0.7))
        n = len(life) - nj
        D = (6 * m / (math.pi * solver.RHO)) ** (1 / 3)
        rows.append((U, m, n, nj, D))
a = np.array(rows)
np.savez_compressed(os.path.join(HERE, "crown.npz"), U=a[:, 0], m=a[:, 1], n=a[:, 2].astype(np.int32),
                    nj=a[:, 3].astype(np.int32), D=a[:, 4])
print("wrote", os.path.join(HERE, "crown.npz"), len(rows), "cases; rings", int(a[:, 2].sum()), "jets", int(a[:, 3].sum()))
