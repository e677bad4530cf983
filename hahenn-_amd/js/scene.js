'use strict';
// Scene model and scene-file schema of the reference (src/js/10-scene.js:5-87, :254-336; the file
// check of src/js/80-ui.js:430-460): item factories with their defaults, the six presets, and the
// table-driven validation that takes every field a file may set from the file when it has the right
// type and from a fresh item of the same kind otherwise. Object packing and the emitters run in the
// library (sl_scene_* of include/suiri.h), reached through FluidSim.

const clamp = (x, a, b) => Math.min(b, Math.max(a, x));
const has = (table, key) => typeof key === 'string' && Object.prototype.hasOwnProperty.call(table, key);

// obstacle shapes: `type` is the shape id of the signed-distance code (10-scene.js:5-15)
const SHAPES = {
  box: { type: 0, dims: { w: 0.10, h: 0.06, d: 0.08 } },
  plate: { type: 0, dims: { w: 0.22, d: 0.10, t: 0.012 } },
  sphere: { type: 1, dims: { r: 0.035 } },
  cylinder: { type: 2, dims: { r: 0.025, h: 0.12 } },
  cone: { type: 3, dims: { r: 0.04, h: 0.08 } },
  torus: { type: 4, dims: { R: 0.05, r: 0.014 } },
  cup: { type: 5, dims: { r: 0.04, h: 0.10, t: 0.006 } },
  container: { type: 6, dims: { w: 0.14, h: 0.06, d: 0.10, t: 0.008 } },
  gutter: { type: 7, dims: { len: 0.28, r: 0.03, t: 0.006 } },
};
const MATERIALS = { ceramic: 1, plastic: 1, metal: 1, stone: 1 };
const SOURCE_MODES = { stream: 1, drops: 1, shower: 1, burst: 1, rain: 1 };

let uid = 1;
function setNextId(n) { uid = n; }

function makeObstacle(shape, pos) {
  const s = SHAPES[shape];
  return {
    id: uid++, kind: 'obstacle', shape, name: shape,
    pos: pos ? [...pos] : [0, 0.05, 0], rot: [0, 0, 0], dims: { ...s.dims },
    color: shape === 'cup' ? '#e9e4da' : shape === 'gutter' ? '#8a8f94' : '#c9c2b4',
    material: shape === 'gutter' ? 'metal' : shape === 'sphere' || shape === 'cone' ? 'stone' : 'ceramic',
  };
}
function makeSource(pos) {
  return {
    id: uid++, kind: 'source', name: 'source', pos: pos ? [...pos] : [0, 0.4, 0],
    pitch: -90, yaw: 0, mode: 'stream', enabled: true,
    flow: 3.0, speed: 0.6, dropD: 5, dropRate: 3, spread: 18, burst: 0.5, burstEvery: 0,
    rainRate: 60, rainSize: 0.3, total: 0,
  };
}
function makeBlock(pos) {
  return { id: uid++, kind: 'block', name: 'block', pos: pos ? [...pos] : [0, 0.05, 0], dims: { w: 0.16, h: 0.10, d: 0.16 } };
}
function makeDrain(pos) {
  return { id: uid++, kind: 'drain', name: 'drain', pos: pos ? [...pos] : [0, 0, 0], rot: [0, 0, 0], r: 0.025 };
}

// the file format {v:1, app:'suiri-lab', scene, cam} (80-ui.js:121)
function sceneFile(scene, cam) { return { v: 1, app: 'suiri-lab', scene, cam: cam || null }; }

// How a field read from a file is checked: (file value, fresh item's value) -> value.
const RULES = {
  num: (v, d) => (Number.isFinite(v) ? v : d),
  bool: (v, d) => (typeof v === 'boolean' ? v : d),
  text: (v, d) => (typeof v === 'string' && v ? v : d),
  vec3: (v, d) => (Array.isArray(v) && v.length === 3 && v.every(Number.isFinite) ? [...v] : d),
  // every dimension of the fresh item's shape, each a finite number or the fresh value
  dims: (v, d) => Object.fromEntries(Object.keys(d).map((k) => [k, Number.isFinite(v && v[k]) ? v[k] : d[k]])),
  material: (v, d) => (has(MATERIALS, v) ? v : d),
  mode: (v, d) => (has(SOURCE_MODES, v) ? v : d),
};
// Per item kind: the fresh item a file's item is checked against, and the rule of every field of it
// besides id, kind and shape (the obstacle's shape picks the fresh item: an unknown shape is a box).
const SCHEMA = {
  obstacle: {
    fresh: (it) => makeObstacle(has(SHAPES, it.shape) ? it.shape : 'box'),
    fields: { name: 'text', pos: 'vec3', rot: 'vec3', dims: 'dims', color: 'text', material: 'material' },
  },
  source: {
    fresh: () => makeSource(),
    fields: {
      name: 'text', pos: 'vec3', pitch: 'num', yaw: 'num', mode: 'mode', enabled: 'bool', flow: 'num', speed: 'num',
      dropD: 'num', dropRate: 'num', spread: 'num', burst: 'num', burstEvery: 'num', rainRate: 'num', rainSize: 'num',
      total: 'num',
    },
  },
  block: { fresh: () => makeBlock(), fields: { name: 'text', pos: 'vec3', dims: 'dims' } },
  drain: { fresh: () => makeDrain(), fields: { name: 'text', pos: 'vec3', rot: 'vec3', r: 'num' } },
};
// tank sizes a file may set (80-ui.js:459)
const TANK_LIMITS = { w: [0.2, 1.2], h: [0.1, 0.6], d: [0.15, 1.0] };

// A scene read from a file {v, app, scene: {tank, items}, solver?}: null unless the tank has three
// positive sizes and the items are a list; items of unknown kinds are dropped; ids stay when they are
// positive integers not seen before, the others take fresh ids above the largest kept one.
function validScene(o) {
  if (!o || !o.scene || !o.scene.tank || !Array.isArray(o.scene.items)) return null;
  const t = o.scene.tank;
  if (!Object.keys(TANK_LIMITS).every((k) => Number.isFinite(t[k]) && t[k] > 0)) return null;
  const items = [];
  const ids = new Set();
  for (const it of o.scene.items) {
    if (!it || !has(SCHEMA, it.kind)) continue;
    const spec = SCHEMA[it.kind];
    const out = spec.fresh(it);
    for (const [field, rule] of Object.entries(spec.fields)) out[field] = RULES[rule](it[field], out[field]);
    out.id = Number.isInteger(it.id) && it.id > 0 && !ids.has(it.id) ? it.id : null;
    if (out.id !== null) ids.add(out.id);
    items.push(out);
  }
  let next = Math.max(0, ...ids) + 1;
  for (const it of items) if (it.id === null) it.id = next++;
  const tank = Object.fromEntries(Object.entries(TANK_LIMITS).map(([k, [lo, hi]]) => [k, clamp(t[k], lo, hi)]));
  const scene = { tank, items };
  // the solver block is an addition of this build (solver sizes and physics for batch jobs); files
  // without it load unchanged because unknown top-level keys are ignored
  if (o.solver && typeof o.solver === 'object') scene.solver = { ...o.solver };
  return scene;
}

// presets (10-scene.js:254-336)
function presetFaucetCup() {
  const cup = makeObstacle('cup', [0.06, 0.06, 0.02]); cup.dims = { r: 0.045, h: 0.12, t: 0.006 };
  const src = makeSource([0.06, 0.42, 0.02]); src.flow = 4; src.speed = 0.7;
  const block = makeBlock([0, 0.03, 0]); block.dims = { w: 0.6, h: 0.05, d: 0.4 };
  const rock = makeObstacle('sphere', [-0.14, 0.03, -0.06]); rock.dims = { r: 0.045 }; rock.material = 'stone';
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [cup, src, block, rock] };
}
function presetCascade() {
  const g1 = makeObstacle('gutter', [0.08, 0.24, -0.06]); g1.dims = { len: 0.34, r: 0.032, t: 0.006 }; g1.rot = [0, 0, -14];
  const g2 = makeObstacle('gutter', [-0.08, 0.15, 0.07]); g2.dims = { len: 0.34, r: 0.032, t: 0.006 }; g2.rot = [0, 0, 14];
  const box = makeObstacle('container', [0.16, 0.035, 0.1]); box.dims = { w: 0.16, h: 0.07, d: 0.12, t: 0.008 };
  const src = makeSource([0.21, 0.36, -0.06]); src.flow = 5; src.speed = 0.8;
  const drain = makeDrain([-0.24, 0, -0.14]); drain.r = 0.03;
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [g1, g2, box, src, drain] };
}
function presetDamBreak() {
  const block = makeBlock([-0.2, 0.11, 0]); block.dims = { w: 0.18, h: 0.22, d: 0.38 };
  const pillar = makeObstacle('cylinder', [0.08, 0.06, 0.04]); pillar.dims = { r: 0.035, h: 0.12 };
  const box = makeObstacle('box', [0.12, 0.025, -0.1]); box.dims = { w: 0.08, h: 0.05, d: 0.08 }; box.rot = [0, 30, 0];
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [block, pillar, box] };
}
function presetDrops() {
  const block = makeBlock([0, 0.03, 0]); block.dims = { w: 0.6, h: 0.06, d: 0.4 };
  const src = makeSource([0, 0.34, 0]); src.mode = 'drops'; src.dropD = 6; src.dropRate = 2; src.speed = 0.0;
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [block, src] };
}
function presetFountain() {
  const block = makeBlock([0, 0.025, 0]); block.dims = { w: 0.6, h: 0.05, d: 0.4 };
  const basin = makeObstacle('cylinder', [0, 0.03, 0]); basin.dims = { r: 0.05, h: 0.06 }; basin.material = 'stone';
  const src = makeSource([0, 0.07, 0]); src.pitch = 90; src.flow = 2.5; src.speed = 1.6;
  // the reference builds a torus here but does not add it to the scene (10-scene.js:313-315)
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [block, basin, src] };
}
function presetRain() {
  const block = makeBlock([0, 0.03, 0]); block.dims = { w: 0.6, h: 0.06, d: 0.4 };
  const rain = makeSource([0, 0.34, 0]); rain.mode = 'rain'; rain.rainRate = 120; rain.rainSize = 0.5;
  const cone = makeObstacle('cone', [0.12, 0.09, 0.05]); cone.dims = { r: 0.05, h: 0.1 }; cone.material = 'stone';
  return { tank: { w: 0.6, h: 0.32, d: 0.4 }, items: [block, rain, cone] };
}
const PRESETS = [
  { id: 'faucet', make: presetFaucetCup }, { id: 'cascade', make: presetCascade }, { id: 'dam', make: presetDamBreak },
  { id: 'drops', make: presetDrops }, { id: 'fountain', make: presetFountain }, { id: 'rain', make: presetRain },
];

module.exports = {
  SHAPES, MATERIALS, SOURCE_MODES, SCHEMA, makeObstacle, makeSource, makeBlock, makeDrain, setNextId, sceneFile,
  validScene, PRESETS,
};
