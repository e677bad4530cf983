"""The divergence a pressure solve leaves, measured from a simulation's readbacks in double precision.

The formula is DIVERGENCE's (src/js/30-wgsl-sim.js:240-264): per cell the net flux f v + (1 - f) s
through its six faces over h, minus drift * max(count / ppc - 1, 0). Evaluated here in float64 from
the stored single-precision velocities, face fractions, solid velocities and particle counts, it is
the divergence of the field itself, free of the rounding of the single-precision formula's
subtractions (of the order of ulp(flux) / h per face).
"""
import numpy as np


def divergence_f64(g):
    """Per cell (z, y, x order) the divergence in 1/s, the liquid mask, the mask of cells with an open
    face, and the bound on how far storing each face velocity in single precision (half an ulp,
    times the face fraction) can move the cell's divergence."""
    from . import F_CELLCOUNT, F_FRAC, F_SVEL, F_VEL
    nx, ny, nz, h = g.nx, g.ny, g.nz, g.h
    shape = (nz + 1, ny + 1, nx + 1, 4)
    fr32 = g.readback(F_FRAC).reshape(shape)
    v32 = g.readback(F_VEL).reshape(shape)
    f = fr32.astype(np.float64)
    F = f * v32.astype(np.float64) + (1.0 - f) * g.readback(F_SVEL).reshape(shape).astype(np.float64)

    def net(a):
        return ((a[:nz, :ny, 1:, 0] - a[:nz, :ny, :nx, 0]) + (a[:nz, 1:, :nx, 1] - a[:nz, :ny, :nx, 1]) +
                (a[1:, :ny, :nx, 2] - a[:nz, :ny, :nx, 2]))

    def total(a):
        return ((a[:nz, :ny, 1:, 0] + a[:nz, :ny, :nx, 0]) + (a[:nz, 1:, :nx, 1] + a[:nz, :ny, :nx, 1]) +
                (a[1:, :ny, :nx, 2] + a[:nz, :ny, :nx, 2]))

    count = g.readback(F_CELLCOUNT).reshape(nz, ny, nx)
    crowd = count.astype(np.float64) / float(g.cfg.ppc) - 1.0
    div = net(F) / h - float(g.cfg.drift) * np.maximum(crowd, 0.0)
    floor = total(np.abs(f) * 0.5 * np.spacing(np.abs(v32)).astype(np.float64)) / h
    fx, fy, fz = fr32[:nz, :ny, :, 0] > 0, fr32[:nz, :, :nx, 1] > 0, fr32[:, :ny, :nx, 2] > 0
    has_open = fx[:, :, :-1] | fx[:, :, 1:] | fy[:, :-1] | fy[:, 1:] | fz[:-1] | fz[1:]
    return div, count > 0, has_open, floor
