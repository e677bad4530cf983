'use strict';
// Host-side logic of the Node binding that needs no GPU: the scene-file check (table-driven
// validScene), the presets, and the library's host-only scene functions through the addon (object
// packing, emitters), pinned by invariants. Exits non-zero on the first failed check.
const assert = require('assert');
const path = require('path');
const lib = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'scene.js'));
const addon = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'suiri_napi.node'));

// validScene: rejects non-scenes, fills missing fields from fresh items, dedups ids, clamps tanks
assert.strictEqual(lib.validScene(null), null);
assert.strictEqual(lib.validScene({ scene: { tank: { w: 0, h: 1, d: 1 }, items: [] } }), null);
assert.strictEqual(lib.validScene({ scene: { tank: { w: 1, h: 1 }, items: [] } }), null);
const s = lib.validScene({ v: 1, scene: { tank: { w: 5, h: 0.01, d: 0.5 }, items: [
  { kind: 'obstacle', shape: 'hexagon', id: 4, pos: [0.1, 0.2] },
  { kind: 'source', id: 4, mode: 'geyser', flow: 7, enabled: 'yes', name: '' },
  { kind: 'block', dims: { w: 0.3, h: 'tall' }, extra: 1 },
  { kind: 'martian' },
  { kind: 'toString' },
  { kind: 'drain', id: 9, r: 0.04, rot: [1, 2, NaN] },
  { kind: 'obstacle', shape: 'torus', id: 2.5, material: 'constructor', dims: { R: 0.07 } },
] }, solver: { iters: 9 } });
assert.deepStrictEqual(s.tank, { w: 1.2, h: 0.1, d: 0.5 });
assert.strictEqual(s.items.length, 5);
assert.strictEqual(s.items[0].shape, 'box');
assert.deepStrictEqual(s.items[0].pos, [0, 0.05, 0]);
assert.deepStrictEqual(s.items[0].dims, lib.SHAPES.box.dims);
assert.strictEqual(s.items[1].mode, 'stream');
assert.strictEqual(s.items[1].flow, 7);
assert.strictEqual(s.items[1].enabled, true);
assert.strictEqual(s.items[1].name, lib.makeSource().name);
assert.deepStrictEqual(s.items[2].dims, { w: 0.3, h: 0.10, d: 0.16 });
assert.strictEqual(s.items[2].extra, undefined);
assert.strictEqual(s.items[3].r, 0.04);
assert.deepStrictEqual(s.items[3].rot, [0, 0, 0]);
assert.strictEqual(s.items[4].shape, 'torus');
assert.deepStrictEqual(s.items[4].dims, { R: 0.07, r: lib.SHAPES.torus.dims.r });
assert.strictEqual(s.items[4].material, lib.makeObstacle('torus').material);
assert.deepStrictEqual(s.solver, { iters: 9 });
const ids = s.items.map((it) => it.id);
assert.strictEqual(new Set(ids).size, 5);
assert.strictEqual(ids[0], 4);
assert.strictEqual(ids[3], 9);
assert.deepStrictEqual([ids[1], ids[2], ids[4]], [10, 11, 12]);  // fresh ids above the largest kept one
// the schema covers every field of a fresh item of each kind (besides id, kind, shape)
const fresh = { obstacle: lib.makeObstacle('cup'), source: lib.makeSource(), block: lib.makeBlock(), drain: lib.makeDrain() };
for (const [kind, item] of Object.entries(fresh)) {
  const want = Object.keys(item).filter((k) => !['id', 'kind', 'shape'].includes(k)).sort();
  assert.deepStrictEqual(Object.keys(lib.SCHEMA[kind].fields).sort(), want, kind);
}
// a valid file round-trips unchanged
const round = lib.validScene(lib.sceneFile(s));
assert.deepStrictEqual(round.items, s.items);

// presets load and pass the check
for (const p of lib.PRESETS) {
  const sc = p.make();
  assert.ok(sc.items.length >= 2, p.id);
  assert.strictEqual(lib.validScene(lib.sceneFile(sc)).items.length, sc.items.length, p.id);
}

// library scene functions (host only)
const cfg = addon.configDefault('standard');
cfg.tank_w = 1.0; cfg.tank_h = 0.5; cfg.tank_d = 0.5; cfg.h = 1 / 128;
const scene = addon.sceneCreate(cfg);
const vp = Math.pow(1 / 128 / 2, 3);
// stream source: the particle count tracks the flow (Q dt / Vp per substep, the fraction carried)
const src = lib.makeSource([0, 0.3, 0]); src.flow = 30; src.speed = 1.5; src.pitch = -30; src.yaw = 90;
let n = 0;
for (let i = 0; i < 240; i++) for (const e of addon.sourceEmits(scene, [src], 1 / 240)) n += e.count;
assert.ok(Math.abs(n - (30 / 60000) / vp) <= 1, `${n}`);
// block lattice: floor(extent / (h/2)) per axis
const blk = lib.makeBlock([0, 0.05, 0]); blk.dims = { w: 0.1, h: 0.1, d: 0.1 };
const recs = addon.blockEmits(scene, [blk], false);
assert.strictEqual(recs.length, 1);
assert.deepStrictEqual(recs[0].b.slice(0, 3), [25, 25, 25]);
assert.strictEqual(recs[0].count, 25 * 25 * 25);
// packing: velocities from the previous pose over dtMove
const ob = lib.makeObstacle('sphere', [0, 0.1, 0]);
addon.packObjects(scene, [ob], 0);
ob.pos = [0.01, 0.1, 0];
const pk = addon.packObjects(scene, [ob], 0.01);
assert.ok(pk.moving && pk.n === 1);
assert.ok(Math.abs(pk.rows[20] - 1.0) < 1e-6 && pk.rows[21] === 0 && pk.rows[27] === 1);
assert.ok(Math.abs(addon.blockLiters(blk, { w: 1, h: 0.5, d: 0.5 }) - 1.0) < 1e-12);
console.log('host_cpu ok');
