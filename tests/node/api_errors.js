'use strict';
// Host-side behaviour of the JavaScript Hush class that needs no GPU: hiddenOf, the limit check and the
// constructor / resize error messages of the reference (hush.js:735, 743, 765).
const assert = require('assert');
const path = require('path');
const { Hush, hiddenOf, decodeWeights, defaultWeights } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

for (let h = 1; h <= 256; h++) {
  assert.strictEqual(hiddenOf(h * h + 34 * h + 12), h);
  assert.strictEqual(hiddenOf(h * h + 34 * h + 13), 0);
}
assert.strictEqual(hiddenOf(812), 16);
assert.deepStrictEqual(Object.assign({}, Hush.requiredLimits), { maxStorageTexturesPerShaderStage: 8 });
assert.ok(Object.isFrozen(Hush.requiredLimits));

const w = defaultWeights();
assert.strictEqual(w.length, 812);
const b64 = Buffer.from(w.buffer).toString('base64');
assert.deepStrictEqual(Array.from(decodeWeights(b64)), Array.from(w));
assert.deepStrictEqual(Array.from(decodeWeights(w.buffer)), Array.from(w));

const dev = { ordinal: 0, limits: { maxStorageTexturesPerShaderStage: 8 } };
assert.throws(() => new Hush({ ordinal: 0, limits: { maxStorageTexturesPerShaderStage: 4 } }, { width: 8, height: 8 }),
  /^Error: Hush needs the device limit maxStorageTexturesPerShaderStage >= 8 \(this device: 4\); create the device with requiredLimits: Hush.requiredLimits$/);
assert.throws(() => new Hush(dev, { width: 8, height: 8, weights: new Uint16Array(5) }),
  /^Error: hush weights: 5 f16 values fit no MLP size$/);
assert.throws(() => new Hush(dev, { width: 0, height: 8 }), /^Error: Hush: bad size 0x8$/);
assert.throws(() => new Hush(dev, { height: 8 }), /^Error: Hush: bad size undefinedx8$/);
// the native layer reports its own errors with the same texts
const native = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush_napi.node'));
assert.throws(() => native.create(0, 8, 8, 1, new Uint16Array(7), 0), /hush weights: 7 f16 values fit no MLP size/);
assert.throws(() => native.create(0, -1, 8, 1, defaultWeights(), 0), /Hush: bad size -1x8/);
// SVGF: options and size are checked before any device call
const { SVGF } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));
assert.throws(() => new SVGF(dev, 0, 8), /SVGF: bad size 0x8/);
assert.throws(() => new SVGF(dev, 8, 8, { iterations: 6 }), /6 a-trous iterations/);
console.log('api_errors: ok');
