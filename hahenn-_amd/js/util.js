'use strict';
// Vector / rotation helpers with the conventions of the reference (src/js/00-util.js:5-61):
// row-major 3x3 rotation matrices, world = R * local, Euler angles in degrees applied X, Y, Z.

const DEG = Math.PI / 180;
const clamp = (x, a, b) => Math.min(b, Math.max(a, x));

const v3 = {
  add: (a, b) => [a[0] + b[0], a[1] + b[1], a[2] + b[2]],
  sub: (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]],
  mul: (a, s) => [a[0] * s, a[1] * s, a[2] * s],
  dot: (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2],
  cross: (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]],
  len: (a) => Math.hypot(a[0], a[1], a[2]),
  norm: (a) => {
    const l = Math.hypot(a[0], a[1], a[2]) || 1;
    return [a[0] / l, a[1] / l, a[2] / l];
  },
};

const m3 = {
  // R = Rz(rz) * Ry(ry) * Rx(rx), degrees
  euler(rx, ry, rz) {
    const cx = Math.cos(rx * DEG), sx = Math.sin(rx * DEG);
    const cy = Math.cos(ry * DEG), sy = Math.sin(ry * DEG);
    const cz = Math.cos(rz * DEG), sz = Math.sin(rz * DEG);
    return [
      cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
      sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
      -sy, cy * sx, cy * cx,
    ];
  },
  transpose: (R) => [R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]],
  mul(A, B) {
    const o = new Array(9);
    for (let r = 0; r < 3; r++) {
      for (let c = 0; c < 3; c++) o[r * 3 + c] = A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c] + A[r * 3 + 2] * B[6 + c];
    }
    return o;
  },
  applyT: (R, p) => [R[0] * p[0] + R[3] * p[1] + R[6] * p[2], R[1] * p[0] + R[4] * p[1] + R[7] * p[2],
    R[2] * p[0] + R[5] * p[1] + R[8] * p[2]],
  // rotation vector (axis * angle) of R
  log(R) {
    const c = clamp((R[0] + R[4] + R[8] - 1) / 2, -1, 1);
    const ang = Math.acos(c);
    if (ang < 1e-6) return [0, 0, 0];
    const s = 2 * Math.sin(ang);
    if (Math.abs(s) < 1e-6) return [0, 0, 0];
    return v3.mul([(R[7] - R[5]) / s, (R[2] - R[6]) / s, (R[3] - R[1]) / s], ang);
  },
};

// deterministic 32-bit seed stream (replaces Math.random() for emission seeds when a seed is given)
function seedStream(seed) {
  let s = seed >>> 0;
  return () => {
    s = (Math.imul(s, 1664525) + 1013904223) >>> 0;
    return s;
  };
}

module.exports = { DEG, clamp, v3, m3, seedStream };
