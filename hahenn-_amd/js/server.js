'use strict';
// Serving loop for Node: scene-bake jobs (a scene JSON of the reference's schema + frame count)
// admitted against a device-memory budget and stepped together on the GPU, one frame per tick
// (sl_server_* through the N-API addon). The library packs a job's objects and fills its blocks at
// admission and emits its sources every frame (sl_job.items), with the emitters a FluidSim uses, so
// a job's results are the same as a plain FluidSim run of the scene with one substep per frame.
const { addon, QUALITY, FIELDS } = require('./fluid_sim');

class BakeServer {
  // opts: {budget (bytes), maxJobs, device, autotuneAfter, graphAfter, framesPerTick}
  constructor(opts = {}) {
    this.handle = addon.serverCreate(opts);
    this.jobs = new Map();
  }

  // scene {tank, items, solver?}; opts {quality, frames, dt, seed, keep, bakeDir, bakeEvery}
  submitScene(scene, opts = {}) {
    const quality = opts.quality || 'standard';
    if (!QUALITY.includes(quality)) throw new Error(`unknown quality ${quality}`);
    const cfg = addon.configDefault(quality);
    cfg.tank_w = scene.tank.w; cfg.tank_h = scene.tank.h; cfg.tank_d = scene.tank.d;
    for (const [k, v] of Object.entries(scene.solver || {})) if (k in cfg) cfg[k] = typeof v === 'boolean' ? (v ? 1 : 0) : v;
    if (opts.seed !== undefined) cfg.seed = opts.seed;  // the seed stream of the job's emission records
    const id = addon.serverSubmit(this.handle, {
      config: cfg, items: scene.items, steps: opts.frames || 100, dt: opts.dt || 1 / 120, keep: opts.keep ? 1 : 0,
      bakeDir: opts.bakeDir, bakeEvery: opts.bakeEvery || 0,
    });
    this.jobs.set(id, { scene, opts });
    return id;
  }

  tick() { return addon.serverTick(this.handle); }
  poll() { return addon.serverPoll(this.handle); }
  info(id) { return addon.serverInfo(this.handle, id); }
  cancel(id) { addon.serverCancel(this.handle, id); }
  release(id) { addon.serverRelease(this.handle, id); this.jobs.delete(id); }
  stats() { return addon.serverStats(this.handle); }

  // a kept job's buffer in the reference layout (see FluidSim.readback)
  readback(id, name) {
    const h = addon.serverJobSim(this.handle, id);
    if (!h) throw new Error(`job ${id} is not resident`);
    const f = FIELDS[name];
    if (f === undefined) throw new Error(`unknown field ${name}`);
    const ab = addon.readback(h, f);
    if (name === 'counters' || name === 'cellCount') return new Uint32Array(ab);
    return new Float32Array(ab);
  }

  // runs ticks until nothing is queued or running
  drain(maxTicks = 1e9) {
    for (let i = 0; i < maxTicks; i++) {
      const running = this.tick();
      const st = this.stats();
      if (!running && !st.queued) return i;
    }
    throw new Error('drain: ticks exhausted');
  }

  destroy() { addon.serverDestroy(this.handle); }
}

module.exports = { BakeServer };
