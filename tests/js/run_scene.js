'use strict';
// Runs a preset through the Node host (FluidSim over the N-API addon) and writes what the oracle needs
// to replay it (object rows, emission records with seeds, frame count) plus the final state.
// usage: node run_scene.js <preset> <frames> <outdir>
const fs = require('fs');
const path = require('path');
const S = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'index.js'));

const [preset, framesArg, outdir] = process.argv.slice(2);
const frames = parseInt(framesArg, 10);
const sc = S.PRESETS.find((p) => p.id === preset).make();
const opts = { seed: 1234, ripple_R: 64, cap: 200000, ww_cap: 20000 };
const sim = new S.FluidSim(sc.tank, 'low', opts);
sim.sceneItems = sc.items;
// record what crosses the boundary
const trace = { preset, frames, opts, tank: sc.tank, emits: [], rows: null, substeps: [] };
sim.setObjects(sc.items, 0);
trace.rows = Array.from(sim.readback('objs').subarray(0, sim.objCount * 28));
const blocks = sim.blockEmits(sc.items);
trace.emits = blocks;
sim.emitNow(blocks);
for (let f = 0; f < frames; f++) {
  sim.setObjects(sc.items, 1 / 120);
  const subs = [];
  sim.step(1 / 120, (s, dt) => {
    const list = s.sourceEmits(sc.items, dt);
    subs.push(list);
    return list;
  });
  trace.substeps.push(subs);
  sim.idle();
  sim.pollStats();
}
const st = sim.stats();
trace.stats = st;
fs.writeFileSync(path.join(outdir, 'trace.json'), JSON.stringify(trace));
fs.writeFileSync(path.join(outdir, 'particles.bin'), Buffer.from(sim.readback('particles').buffer));
fs.writeFileSync(path.join(outdir, 'pres.bin'), Buffer.from(sim.readback('pres').buffer));
console.log(JSON.stringify({ count: st.count, frames: st.frames, vmax: st.vmax }));
// statsAsync (a Promise resolved once the submitted work is done) agrees with the blocking stats,
// also with work enqueued after the call
sim.step(1 / 120, null);
const pending = sim.statsAsync();
const stNow = sim.stats();
pending.then((sa) => {
  if (sa.count !== stNow.count || sa.frames !== stNow.frames || sa.ww !== stNow.ww) {
    console.error('statsAsync mismatch', JSON.stringify(sa), JSON.stringify(stNow));
    process.exit(5);
  }
  console.log(JSON.stringify({ statsAsync: 'ok' }));
  sim.destroy();
}, (e) => { console.error(e); process.exit(6); });
