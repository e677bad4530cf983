'use strict';
// FluidSim for Node: the interface of the reference's simulation driver (src/js/50-sim.js:28-376)
// over the MI355X library (N-API addon -> C ABI -> HIP kernels). Callers written against the
// reference's FluidSim (main loop, automation API, undo/reset) use it unchanged:
//   new FluidSim(tank, qualityKey, opts), setObjects(items, dtMove), emitNow(list, fill),
//   step(frameDt, emitFn), pollStats(), clearWater(), saveWater()/hasWaterCopy()/restoreWater(),
//   destroy(), waterLiters, and the host fields (nx, ny, nz, h, origin, rim, cap, vp, count, ...).
// The GPU buffers stay on the device; readback()/upload() move them in the reference layouts.
const path = require('path');

const addon = require(path.join(__dirname, '..', 'suiri_napi.node'));

const MAX_OBJ = 64;
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
    this.vp = g.vp;
    this.iters = g.iters;
    this.ripple = { R: g.R, M: g.M };
    // the library keeps the sources' progress, the objects' previous poses and the seed stream of the
    // emission records (cfg.seed); records handed in without a seed take one from a stream here
    this.scene = addon.simScene(this.handle);
    this._seedState = (opts.seed === undefined ? ((Math.random() * 4294967295) >>> 0) : opts.seed) >>> 0;
    this.objCount = 0;
    this.sceneItems = null;
    this.dead = false;
    this._pull();
  }

  _pull() {
    const st = addon.hostState(this.handle);
    this.count = st.count; this.time = st.time; this.frame = st.frame; this.vmax = st.vmax;
    this.full = !!st.full; this.impacts = st.impacts; this.wwCount = st.wwCount;
  }

  _nextSeed() {
    this._seedState = (Math.imul(this._seedState, 747796405) + 2891336453) >>> 0;
    return this._seedState;
  }
  _seeded(list) {
    return list.map((e) => (e.seed === undefined ? { ...e, seed: this._nextSeed() } : e));
  }

  // obstacles and drains -> object records with velocities from the previous poses, packed and
  // uploaded by the library (50-sim.js:165-211). Returns whether any object moved.
  setObjects(items, dtMove) {
    const moving = addon.setItems(this.handle, items, dtMove || 0);
    this.objCount = items.filter((it) => it.kind === 'obstacle' || it.kind === 'drain').length;
    if (this.objCount > MAX_OBJ) this.objCount = MAX_OBJ;
    return moving;
  }

  // emission records of the sources among `items` for one substep of dt (sourceEmits, 55-emit.js:13-69)
  sourceEmits(items, dt) { return addon.sourceEmits(this.scene, items, dt); }
  // lattice records of the blocks among `items` (blockEmits, 55-emit.js:83-106)
  blockEmits(items, fill) { return addon.blockEmits(this.scene, items, !!fill); }
  // a frame whose substeps emit the sources among `items`, all in the library
  stepItems(frameDt, items) {
    if (this.dead || !(frameDt > 0)) return;
    addon.stepItems(this.handle, frameDt, items);
    this._pull();
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

  // a Promise of the same stats, resolved once the submitted work is done (the wait runs on a worker
  // thread, so the JS thread stays free)
  statsAsync() {
    return addon.statsAsync(this.handle).then((st) => { this._pull(); return st; });
  }

  idle() { addon.sync(this.handle); }

  // the library resets the sources' progress with the water, and saves / restores it with the copy
  clearWater() {
    addon.clearWater(this.handle);
    this._pull();
  }

  saveWater() { return addon.saveWater(this.handle); }
  hasWaterCopy(id) { return addon.hasWaterCopy(this.handle, id); }
  restoreWater(id) {
    if (!addon.restoreWater(this.handle, id)) return false;
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

module.exports = { FluidSim, FIELDS, addon, QUALITY, MAX_OBJ };
