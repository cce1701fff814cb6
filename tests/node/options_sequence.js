'use strict';
// The JavaScript Hush class's options on HIP device 0, over the inputs written for denoise_sequence.js:
//   A  exposure 1, set to 2.5 before frame 2, reset() before frame 3
//   B  weights given as a base64 string (<dir>/weights.b64, the reference's decodeWeights form)
//   C  precision 'fp16-mfma'      D  f16: true (the reference's option)      E  compensated: true
//   F  precision 'fp8-mfma' (the calibrated input scales stored for the built-in weights)
//   G  precision 'fp8-mfma', fp8Scales [1/16, 4, 16]
// usage: node options_sequence.js <dir> <width> <height> <frames>; writes <dir>/out<case><i>.bin
const fs = require('fs');
const path = require('path');
const { Hush, initDevice } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

const [dir, W, H, F] = [process.argv[2], +process.argv[3], +process.argv[4], +process.argv[5]];
const device = initDevice(0);
const keys = ['radiance', 'normalDepth', 'motion', 'material', 'emission'];
const tex = {};
for (const k of keys) tex[k] = device.createTexture(W, H);
function load(f) {
  for (const k of keys) {
    const b = fs.readFileSync(path.join(dir, `f${f}_${k}.bin`));
    tex[k].upload(new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)));
  }
}
const b64 = fs.readFileSync(path.join(dir, 'weights.b64'), 'utf8').trim();
const cases = {
  A: { width: W, height: H, exposure: 1 },
  B: { width: W, height: H, weights: b64 },
  C: { width: W, height: H, precision: 'fp16-mfma' },
  D: { width: W, height: H, f16: true },
  E: { width: W, height: H, compensated: true },
  F: { width: W, height: H, precision: 'fp8-mfma' },
  G: { width: W, height: H, precision: 'fp8-mfma', fp8Scales: [0.0625, 4, 16] },
};
for (const [name, opts] of Object.entries(cases)) {
  const hush = new Hush(device, opts);
  for (let f = 0; f < F; f++) {
    if (name === 'A' && f === 2) hush.exposure = 2.5;
    if (name === 'A' && f === 3) hush.reset();
    load(f);
    const out = hush.denoise(tex);
    fs.writeFileSync(path.join(dir, `out${name}${f}.bin`), Buffer.from(out.download().buffer));
  }
  if (name === 'A' && hush.exposure !== 2.5) throw new Error('exposure getter');
  if (name === 'B' && hush.hidden !== 8) throw new Error('hidden width of the base64 weights: ' + hush.hidden);
  if (name === 'D' && !hush.f16) throw new Error('f16 flag');
  if (name === 'G' && hush.fp8Scales.join() !== '0.0625,4,16') throw new Error('fp8Scales getter: ' + hush.fp8Scales);
  if (name === 'G') {
    let threw = false;
    try { hush.fp8Scales = [3, 1, 1]; } catch (e) { threw = true; }
    if (!threw || hush.fp8Scales.join() !== '0.0625,4,16') throw new Error('fp8Scales accepted a non-power of two');
  }
  hush.destroy();
  process.stdout.write(`case ${name} ok\n`);
}
for (const k of keys) tex[k].destroy();
console.log('options_sequence: ok');
