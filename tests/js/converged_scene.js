'use strict';
// The dam preset through the Node host with the converged pressure solve: prints the solver's stats
// and the particle count after each of a few frames, and the error thrown for an unknown solver.
// usage: node converged_scene.js <frames>
const path = require('path');
const S = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'index.js'));

const frames = parseInt(process.argv[2], 10);
const sc = S.PRESETS.find((p) => p.id === 'dam').make();
const sim = new S.FluidSim(sc.tank, 'low', { seed: 1234, ripple_R: 64, cap: 200000, ww_cap: 20000 });
sim.setObjects(sc.items, 0);
sim.emitNow(sim.blockEmits(sc.items));
let unknown = '';
try { sim.setPressureSolver('multigrid'); } catch (e) { unknown = e.message; }
sim.setPressureSolver('converged', 1e-5, 40);
const out = [];
for (let f = 0; f < frames; f++) {
  sim.step(1 / 120, null);
  out.push({ ...sim.pressureSolverStats(), count: sim.stats().count });
}
console.log(JSON.stringify({ unknown, frames: out }));
sim.destroy();
