'use strict';
// FluidSim for Node: the interface of the reference's simulation driver (src/js/50-sim.js:28-376)
// over the MI355X library (N-API addon -> C ABI -> HIP kernels). Callers written against the
// reference's FluidSim (main loop, automation API, undo/reset) use it unchanged:
//   new FluidSim(tank, qualityKey, opts), setObjects(items, dtMove), emitNow(list, fill),
//   step(frameDt, emitFn), pollStats(), clearWater(), saveWater()/hasWaterCopy()/restoreWater(),
//   destroy(), waterLiters, and the host fields (nx, ny, nz, h, origin, rim, cap, vp, count, ...).
// The GPU buffers stay on the device; readback()/upload() move them in the reference layouts.
const path = require('path');
const { m3, v3, seedStream } = require('./util');
const { shapeParams, boundRadius, obstacleRot } = require('./scene');

const addon = require(path.join(__dirname, '..', 'suiri_napi.node'));

const MAX_OBJ = 64;

// obstacles and drains -> packed 28-float object records; velocities from the previous transforms
// `prevXf` (id -> {pos, R}) (50-sim.js:165-211). Returns {data, n, nextXf, moving}.
function packObjects(items, prevXf, dtMove) {
  const data = new Float32Array(MAX_OBJ * 28);
  let n = 0;
  const nextXf = new Map();
  let moving = false;
  for (const it of items) {
    if (n >= MAX_OBJ) break;
    if (it.kind !== 'obstacle' && it.kind !== 'drain') continue;
    const R = it.kind === 'obstacle' ? obstacleRot(it) : m3.euler(it.rot[0], it.rot[1], it.rot[2]);
    const RT = m3.transpose(R);
    let type, d, flag, bound;
    if (it.kind === 'obstacle') {
      const sp = shapeParams(it);
      type = sp[0]; d = sp.slice(1); flag = 1; bound = boundRadius(it) + 0.002;
    } else {
      type = 9; d = [it.r, 0.012, 0, 0]; flag = 2; bound = it.r + 0.02;
    }
    let lin = [0, 0, 0], ang = [0, 0, 0];
    const prev = prevXf.get(it.id);
    if (prev && dtMove > 0) {
      lin = v3.mul(v3.sub(it.pos, prev.pos), 1 / dtMove);
      ang = v3.mul(m3.log(m3.mul(R, m3.transpose(prev.R))), 1 / dtMove);
      if (v3.len(lin) > 1e-5 || v3.len(ang) > 1e-4) moving = true;
      if (v3.len(lin) > 4) lin = v3.mul(v3.norm(lin), 4);
      if (v3.len(ang) > 40) ang = v3.mul(v3.norm(ang), 40);
    }
    nextXf.set(it.id, { pos: [...it.pos], R });
    const o = n * 28;
    data.set([RT[0], RT[1], RT[2], 0, RT[3], RT[4], RT[5], 0, RT[6], RT[7], RT[8], 0], o);
    data.set([it.pos[0], it.pos[1], it.pos[2], type], o + 12);
    data.set([d[0], d[1], d[2], d[3]], o + 16);
    data.set([lin[0], lin[1], lin[2], bound], o + 20);
    data.set([ang[0], ang[1], ang[2], flag], o + 24);
    n++;
  }
  return { data: data.subarray(0, n * 28), n, nextXf, moving };
}

// grid of a tank configuration, 50-sim.js:33-45: {nx, ny, nz, h, origin, rim, vp, radius}
function gridOf(cfg) {
  const nx = Math.max(8, Math.round(cfg.tank_w / cfg.h));
  const ny = Math.max(8, Math.round(cfg.tank_h / cfg.h));
  const nz = Math.max(8, Math.round(cfg.tank_d / cfg.h));
  const h = Math.min(cfg.tank_w / nx, cfg.tank_h / ny, cfg.tank_d / nz);
  return { nx, ny, nz, h, origin: [-nx * h / 2, 0, -nz * h / 2], rim: ny * h, vp: Math.pow(h / 2, 3), radius: 0.5 * h * 1.05 };
}
const FIELDS = {
  particles: 0, ww: 1, counters: 2, acc: 3, cellCount: 4, vel: 5, velOld: 6, frac: 7, svel: 8, cellType: 9,
  divg: 10, pres: 11, objs: 12, rsrc: 13, rippleState: 14, rippleHat: 15, rippleField: 16, rippleMask: 17,
  rippleTex: 18, topH: 19, rippleTex16: 20, args: 21,
};
const QUALITY = ['low', 'standard', 'high'];

class FluidSim {
  // tank {w, h, d} [m]; qualityKey 'low' | 'standard' | 'high'; opts overrides solver settings
  // (h, cap, ww_cap, iters, ripple_R, flip, omega, whitewater, reorder_every, ...) and `seed`
  constructor(tank, qualityKey, opts = {}) {
    if (!QUALITY.includes(qualityKey)) throw new Error(`unknown quality ${qualityKey}`);
    const cfg = addon.configDefault(qualityKey);
    cfg.tank_w = tank.w; cfg.tank_h = tank.h; cfg.tank_d = tank.d;
    for (const [k, v] of Object.entries(opts)) if (k in cfg) cfg[k] = typeof v === 'boolean' ? (v ? 1 : 0) : v;
    this.cfg = cfg;
    this.handle = addon.create(cfg);
    const g = addon.gridInfo(this.handle);
    this.nx = g.nx; this.ny = g.ny; this.nz = g.nz; this.h = g.h;
    this.origin = g.origin; this.rim = g.rim; this.nodes = g.nodes; this.cells = g.cells;
    this.cap = g.cap; this.wwCap = g.wwCap; this.radius = g.radius;
    this.vp = Math.pow(this.h / 2, 3);  // 50-sim.js:44, in JS like the reference (the emitters use it)
    this.iters = g.iters;
    this.ripple = { R: g.R, M: g.M };
    this.emitAcc = new Map();
    this.prevXf = new Map();
    this.seed = seedStream(opts.seed === undefined ? ((Math.random() * 4294967295) >>> 0) : opts.seed);
    this.sceneItems = null;
    this.dead = false;
    this._pull();
  }

  _pull() {
    const st = addon.hostState(this.handle);
    this.count = st.count; this.time = st.time; this.frame = st.frame; this.vmax = st.vmax;
    this.full = !!st.full; this.impacts = st.impacts; this.wwCount = st.wwCount;
  }

  _seeded(list) {
    return list.map((e) => ({ ...e, seed: e.seed === undefined ? this.seed() : e.seed }));
  }

  // obstacles and drains -> packed 28-float object records; velocities from the previous transforms
  // (50-sim.js:165-211). Returns whether any object moved.
  setObjects(items, dtMove) {
    const { data, n, nextXf, moving } = packObjects(items, this.prevXf, dtMove);
    this.prevXf = nextXf;
    this.objCount = n;
    addon.setObjects(this.handle, data);
    return moving;
  }

  // frame_%05d.npz (+ surface_%05d.exr) of the current state, the reference's bake cache layout
  bake(dir, frame) { addon.bakeFrame(this.handle, dir, frame); }

  // emissions outside the time stepping (water blocks at a reset, or poured in) (50-sim.js:248-265)
  emitNow(list, fill) {
    const total = addon.emitNow(this.handle, this._seeded(list), !!fill);
    this._pull();
    return total;
  }

  // advances the liquid by frameDt seconds; emitFn(sim, dt, t) returns each substep's emission list
  // (50-sim.js:300-336)
  step(frameDt, emitFn) {
    if (this.dead || !(frameDt > 0)) return;
    const { n, dt } = addon.stepBegin(this.handle, frameDt);
    for (let s = 0; s < n; s++) {
      this._pull();
      const list = emitFn ? emitFn(this, dt, this.time) : [];
      addon.substep(this.handle, dt, this._seeded(list));
    }
    addon.stepEnd(this.handle, frameDt);
    this._pull();
  }

  // pins the substep count (fixed-dt workloads); 0 restores the adaptive count
  setSubsteps(n) { addon.setSubsteps(this.handle, n); }

  // pressure solver: 'sor' (the reference's sweeps, the default) or 'converged' (multigrid-
  // preconditioned CG until the largest residual is at most tolPerS 1/s, or maxIters iterations)
  setPressureSolver(kind, tolPerS = 1e-4, maxIters = 24) {
    const id = { sor: 0, converged: 1 }[kind];
    if (id === undefined) throw new Error(`unknown pressure solver '${kind}'`);
    addon.setPressureSolver(this.handle, id, tolPerS, maxIters);
  }
  pressureSolverStats() { return addon.pressureSolverStats(this.handle); }

  pollStats() {
    const updated = addon.pollStats(this.handle);
    this._pull();
    return updated;
  }

  // waits for the submitted work, then the counters (suiriLab.stats, 90-main.js:196)
  stats() {
    const st = addon.stats(this.handle);
    this._pull();
    return st;
  }

  idle() { addon.sync(this.handle); }

  clearWater() {
    addon.clearWater(this.handle);
    this.emitAcc.clear();
    this._pull();
  }

  saveWater() {
    const id = addon.saveWater(this.handle);
    this.saved = { id, acc: new Map([...this.emitAcc].map(([k, v]) => [k, { ...v }])) };
    return id;
  }
  hasWaterCopy(id) { return addon.hasWaterCopy(this.handle, id); }
  restoreWater(id) {
    if (!addon.restoreWater(this.handle, id)) return false;
    if (this.saved && this.saved.id === id) this.emitAcc = new Map([...this.saved.acc].map(([k, v]) => [k, { ...v }]));
    this._pull();
    return true;
  }

  noRippleMask() { addon.setRippleMask(this.handle, false); }
  reorder() { addon.reorder(this.handle); }

  // a device buffer in the reference layout: Float32Array / Int32Array / Uint32Array / Uint16Array
  readback(name) {
    const f = FIELDS[name];
    if (f === undefined) throw new Error(`unknown field ${name}`);
    const ab = addon.readback(this.handle, f);
    if (name === 'acc' || name === 'rsrc') return new Int32Array(ab);
    if (name === 'counters' || name === 'cellCount' || name === 'cellType' || name === 'args') return new Uint32Array(ab);
    if (name === 'rippleTex16') return new Uint16Array(ab);
    return new Float32Array(ab);
  }
  upload(name, data) {
    const f = FIELDS[name];
    if (f === undefined) throw new Error(`unknown field ${name}`);
    addon.upload(this.handle, f, data);
    this._pull();
  }

  destroy() {
    if (this.dead) return;
    this.dead = true;
    addon.destroy(this.handle);
  }

  get waterLiters() { return this.count * this.vp * 1000; }
}

module.exports = { FluidSim, FIELDS, addon, packObjects, gridOf, QUALITY };
