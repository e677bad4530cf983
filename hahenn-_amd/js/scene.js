'use strict';
// Scene model and scene-JSON schema of the reference (src/js/10-scene.js, src/js/80-ui.js:119-121,
// :430-460): item factories with their defaults, obstacle shape parameters for the GPU, the six
// presets, and the file validation that fills anything missing from a fresh item of the same kind.
const { clamp, m3 } = require('./util');

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

// outlet geometry from flow and speed (10-scene.js:91-99)
function sourceNozzle(src) {
  const Q = src.flow / 60000;
  const v = Math.max(src.speed, 0.05);
  return Math.sqrt(4 * Q / (Math.PI * v));
}
function sourceDir(src) {
  const p = src.pitch * Math.PI / 180, y = src.yaw * Math.PI / 180;
  return [Math.cos(p) * Math.sin(y), Math.sin(p), Math.cos(p) * Math.cos(y)];
}

// GPU parameters of an obstacle: [type, d0, d1, d2, d3] (10-scene.js:143-157)
function shapeParams(o) {
  const d = o.dims;
  switch (o.shape) {
    case 'box': return [0, d.w / 2, d.h / 2, d.d / 2, Math.min(0.003, d.w / 4, d.h / 4, d.d / 4)];
    case 'plate': return [0, d.w / 2, d.t / 2, d.d / 2, Math.min(0.002, d.t / 4)];
    case 'sphere': return [1, d.r, 0, 0, 0];
    case 'cylinder': return [2, d.r, d.h / 2, 0, 0];
    case 'cone': return [3, d.r, d.h / 2, 0, 0];
    case 'torus': return [4, d.R, d.r, 0, 0];
    case 'cup': return [5, d.r, d.h / 2, Math.min(d.t, d.r * 0.8), 0];
    case 'container': return [6, d.w / 2, d.h / 2, d.d / 2, Math.min(d.t, d.w / 3, d.d / 3, d.h * 0.8)];
    case 'gutter': return [7, d.len / 2, d.r, d.t, 0];
    default: return [0, 0.05, 0.05, 0.05, 0];
  }
}
// bounding-sphere radius of an obstacle (10-scene.js:177-188)
function boundRadius(o) {
  const [t, a, b, c] = shapeParams(o);
  switch (t) {
    case 0: case 6: return Math.hypot(a, b, c);
    case 1: return a;
    case 2: case 3: case 5: return Math.hypot(a, b);
    case 4: return a + b;
    case 7: return Math.hypot(a, b + c);
    default: return 0.1;
  }
}
function obstacleRot(o) { return m3.euler(o.rot[0], o.rot[1], o.rot[2]); }

// the file format {v:1, app:'suiri-lab', scene, cam} (80-ui.js:121)
function sceneFile(scene, cam) { return { v: 1, app: 'suiri-lab', scene, cam: cam || null }; }

// a scene read from a file: every field is checked against a fresh item of the same kind, and
// anything missing or unusable takes that item's value; tank sizes are clamped (80-ui.js:430-460)
function validScene(o) {
  if (!o || !o.scene || !o.scene.tank || !Array.isArray(o.scene.items)) return null;
  const t = o.scene.tank;
  if (![t.w, t.h, t.d].every((x) => Number.isFinite(x) && x > 0)) return null;
  const fresh = {
    obstacle: (it) => makeObstacle(SHAPES[it.shape] ? it.shape : 'box'),
    source: () => makeSource(), block: () => makeBlock(), drain: () => makeDrain(),
  };
  const num = (v, d) => (Number.isFinite(v) ? v : d);
  const vec = (v, d) => (Array.isArray(v) && v.length === 3 && v.every(Number.isFinite) ? [...v] : d);
  const items = [];
  const ids = new Set();
  for (const it of o.scene.items) {
    if (!it || !fresh[it.kind]) continue;
    const base = fresh[it.kind](it);
    const out = { ...base };
    for (const [k, v] of Object.entries(base)) {
      if (k === 'id') continue;
      if (typeof v === 'number') out[k] = num(it[k], v);
      else if (typeof v === 'boolean') out[k] = typeof it[k] === 'boolean' ? it[k] : v;
      else if (typeof v === 'string') out[k] = typeof it[k] === 'string' && it[k] ? it[k] : v;
      else if (Array.isArray(v)) out[k] = vec(it[k], v);
      else if (v && typeof v === 'object') {
        out[k] = Object.fromEntries(Object.entries(v).map(([dk, dv]) => [dk, num(it[k] && it[k][dk], dv)]));
      }
    }
    if (it.kind === 'obstacle') { out.shape = base.shape; if (!MATERIALS[out.material]) out.material = base.material; }
    if (it.kind === 'source' && !SOURCE_MODES[out.mode]) out.mode = base.mode;
    out.id = Number.isInteger(it.id) && it.id > 0 && !ids.has(it.id) ? it.id : null;
    if (out.id !== null) ids.add(out.id);
    items.push(out);
  }
  let next = Math.max(0, ...ids) + 1;
  for (const it of items) if (it.id === null) it.id = next++;
  const scene = { tank: { w: clamp(t.w, 0.2, 1.2), h: clamp(t.h, 0.1, 0.6), d: clamp(t.d, 0.15, 1.0) }, items };
  // the solver block is an addition of this build (solver sizes and physics for batch jobs); old
  // files without it load unchanged because unknown top-level keys are ignored
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
  SHAPES, MATERIALS, SOURCE_MODES, makeObstacle, makeSource, makeBlock, makeDrain, setNextId,
  sourceNozzle, sourceDir, shapeParams, boundRadius, obstacleRot, sceneFile, validScene, PRESETS,
};
