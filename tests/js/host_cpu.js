'use strict';
// Host-side logic of the Node binding that needs no GPU: scene validation, shape parameters,
// emitters. Exits non-zero on the first failed check.
const assert = require('assert');
const path = require('path');
const lib = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'scene.js'));
const emit = require(path.join(__dirname, '..', '..', 'hahenn-_amd', 'js', 'emit.js'));

// validScene: rejects non-scenes, fills missing fields from fresh items, dedups ids, clamps tanks
assert.strictEqual(lib.validScene(null), null);
assert.strictEqual(lib.validScene({ scene: { tank: { w: 0, h: 1, d: 1 }, items: [] } }), null);
const s = lib.validScene({ v: 1, scene: { tank: { w: 5, h: 0.01, d: 0.5 }, items: [
  { kind: 'obstacle', shape: 'hexagon', id: 4, pos: [0.1, 0.2] },
  { kind: 'source', id: 4, mode: 'geyser', flow: 7 },
  { kind: 'block', dims: { w: 0.3 } },
  { kind: 'martian' },
  { kind: 'drain', id: 9, r: 0.04 },
] } });
assert.deepStrictEqual(s.tank, { w: 1.2, h: 0.1, d: 0.5 });
assert.strictEqual(s.items.length, 4);
assert.strictEqual(s.items[0].shape, 'box');
assert.deepStrictEqual(s.items[0].pos, [0, 0.05, 0]);
assert.strictEqual(s.items[1].mode, 'stream');
assert.strictEqual(s.items[1].flow, 7);
assert.deepStrictEqual(s.items[2].dims, { w: 0.3, h: 0.10, d: 0.16 });
const ids = s.items.map((it) => it.id);
assert.strictEqual(new Set(ids).size, 4);
assert.strictEqual(ids[0], 4);
assert.strictEqual(ids[3], 9);
assert.ok(ids[1] > 9 && ids[2] > 9);

// shape parameters and bounding radii
const box = lib.makeObstacle('box'); box.dims = { w: 0.1, h: 0.06, d: 0.08 };
assert.deepStrictEqual(lib.shapeParams(box), [0, 0.05, 0.03, 0.04, 0.003]);
assert.ok(Math.abs(lib.boundRadius(box) - Math.hypot(0.05, 0.03, 0.04)) < 1e-15);
const tor = lib.makeObstacle('torus'); tor.dims = { R: 0.06, r: 0.015 };
assert.strictEqual(lib.boundRadius(tor), 0.075);

// stream source: the accumulated particle volume tracks the flow (Q dt / Vp per substep)
const sim = { vp: Math.pow(1 / 128 / 2, 3), h: 1 / 128, origin: [-0.5, 0, -0.25], emitAcc: new Map() };
const src = lib.makeSource([0, 0.3, 0]); src.flow = 30; src.speed = 1.5; src.pitch = -30; src.yaw = 90;
let n = 0;
for (let i = 0; i < 240; i++) for (const e of emit.sourceEmits(sim, 1 / 240, i / 240, [src])) n += e.count;
const expected = (30 / 60000) / sim.vp;
assert.ok(Math.abs(n - expected) <= 1, `${n} vs ${expected}`);

// block lattice: n = floor(extent / (h/2)) per axis, centred, clipped to the tank
const blk = lib.makeBlock([0, 0.05, 0]); blk.dims = { w: 0.1, h: 0.1, d: 0.1 };
const recs = emit.blockEmits(sim, [blk]);
assert.strictEqual(recs.length, 1);
assert.deepStrictEqual(recs[0].b.slice(0, 3), [25, 25, 25]);
assert.strictEqual(recs[0].count, 25 * 25 * 25);

// presets load
for (const p of lib.PRESETS) { const sc = p.make(); assert.ok(sc.items.length >= 2, p.id); }
console.log('host_cpu ok');
