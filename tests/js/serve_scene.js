'use strict';
// A preset served by the BakeServer (N-API addon -> sl_server) next to the same preset run alone
// through FluidSim with one substep per frame: the served job's particles must equal the plain run's.
// The ripple layer is off: a plain run's ripple depth reads the count it last polled.
// usage: node serve_scene.js <preset> <frames> <bakedir>
const path = require('path');
const S = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'index.js'));

const [preset, framesArg, bakeDir] = process.argv.slice(2);
const frames = parseInt(framesArg, 10);
const dt = 1 / 120;
const sc = S.PRESETS.find((p) => p.id === preset).make();
sc.solver = { ripple_R: 0, cap: 150000, ww_cap: 20000 };

const sim = new S.FluidSim(sc.tank, 'low', { seed: 77, ...sc.solver });
sim.sceneItems = sc.items;
sim.setSubsteps(1);
sim.setObjects(sc.items, 0);
sim.emitNow(sim.blockEmits(sc.items));
for (let f = 0; f < frames; f++) sim.step(dt, (s, d) => s.sourceEmits(sc.items, d));
const plainCount = sim.stats().count;
const plain = sim.readback('particles').subarray(0, plainCount * 8);
sim.destroy();

const srv = new S.BakeServer({ budget: 4e9, maxJobs: 4, autotuneAfter: 3 });
srv.submitScene(sc, { quality: 'low', frames: frames + 7, seed: 5, dt });          // shares the device
const id = srv.submitScene(sc, { quality: 'low', frames, seed: 77, dt, keep: true, bakeDir, bakeEvery: frames });
const ticks = srv.drain();
const info = srv.info(id);
const served = srv.readback(id, 'particles').subarray(0, info.count * 8);
let equal = info.count === plainCount;
for (let i = 0; equal && i < plain.length; i++) {
  if (!Object.is(plain[i], served[i])) equal = false;
}
const done = srv.poll().map((j) => j.status);
const st = srv.stats();
srv.release(id);
srv.destroy();
console.log(JSON.stringify({ plainCount, servedCount: info.count, equal, ticks, done, frames: info.frames,
  latencyP50: st.latencyP50 }));
