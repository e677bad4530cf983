'use strict';
// Node entry point: the FluidSim driver, the scene model, and a headless counterpart of the
// reference's automation API (window.suiriLab, src/js/90-main.js:145-205) without the DOM parts.
// Object packing and the emitters run in the library (FluidSim.setObjects / sourceEmits / blockEmits).
const scene = require('./scene');
const { FluidSim, FIELDS, addon } = require('./fluid_sim');
const { BakeServer } = require('./server');

// one scene + one simulation, driven like the reference's headless runs
class SuiriLab {
  constructor(sceneObj, quality = 'standard', opts = {}) {
    this.scene = sceneObj;
    this.quality = quality;
    this.opts = opts;
    this.rebuildSim(true);
  }

  // rebuildSim / restartWater, 90-main.js:5-19
  rebuildSim(refill) {
    if (this.sim) this.sim.destroy();
    this.sim = new FluidSim(this.scene.tank, this.quality, { ...this.opts, ...(this.scene.solver || {}) });
    this.sim.sceneItems = this.scene.items;
    this.sim.setObjects(this.scene.items, 0);
    if (refill) this.sim.emitNow(this.sim.blockEmits(this.scene.items));
  }
  restartWater() {
    this.sim.clearWater();
    this.sim.setObjects(this.scene.items, 0);
    this.sim.emitNow(this.sim.blockEmits(this.scene.items));
  }
  resetSim() {
    const copy = this.sim.saveWater();
    this.restartWater();
    return copy;
  }
  loadPreset(id) {
    const p = scene.PRESETS.find((x) => x.id === id);
    if (!p) throw new Error(`unknown preset ${id}`);
    this.scene = p.make();
    this.rebuildSim(true);
  }

  // run(seconds, dt): setObjects + step + wait + pollStats per frame (90-main.js:151-165)
  run(seconds, dt = 1 / 120) {
    const emitFn = (sim, d) => sim.sourceEmits(this.scene.items, d);
    for (let t = 0; t < seconds - 1e-9; t += dt) {
      this.sim.setObjects(this.scene.items, dt);
      this.sim.step(dt, emitFn);
      this.sim.idle();
      this.sim.pollStats();
    }
    return this.stats();
  }

  stats() {
    const st = this.sim.stats();
    return { frames: st.frames, impacts: st.impacts, count: st.count, ww: st.ww, vmax: st.vmax, time: st.time,
      liters: st.liters, grid: [this.sim.nx, this.sim.ny, this.sim.nz], gpuErrors: st.gpuErrors };
  }

  // max |h|, max and mean slope of the ripple texture (90-main.js:175-194); f32 values
  rippleStats() {
    const tex = this.sim.readback('rippleTex');
    let maxH = 0, maxS = 0, sumS = 0;
    const n = tex.length / 4;
    for (let i = 0; i < n; i++) {
      maxH = Math.max(maxH, Math.abs(tex[4 * i]));
      const sl = Math.hypot(tex[4 * i + 1], tex[4 * i + 2]);
      maxS = Math.max(maxS, sl);
      sumS += sl;
    }
    return { maxH, maxS, meanS: n ? sumS / n : 0 };
  }
  noRippleMask() { this.sim.noRippleMask(); }
  destroy() { if (this.sim) this.sim.destroy(); }
}

// litres of a block inside a tank {w, h, d} (blockLiters, 55-emit.js:108-113)
const blockLiters = (tank, block) => addon.blockLiters(block, tank);

module.exports = { FluidSim, FIELDS, addon, SuiriLab, BakeServer, blockLiters, ...scene };
