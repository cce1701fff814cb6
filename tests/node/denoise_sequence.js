'use strict';
// Runs a G-buffer sequence through the JavaScript Hush class on HIP device 0.
// usage: node denoise_sequence.js <dir> <width> <height> <frames>
// reads <dir>/f<i>_<input>.bin (binary16, width*height*4 values) and writes <dir>/out<i>.bin plus
// <dir>/timing.json (per-pass times of the last frame).
const fs = require('fs');
const path = require('path');
const { Hush, initDevice } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

const [dir, W, H, F] = [process.argv[2], +process.argv[3], +process.argv[4], +process.argv[5]];
const device = initDevice(0);
const hush = new Hush(device, { width: W, height: H, exposure: 1 });
const keys = ['radiance', 'normalDepth', 'motion', 'material', 'emission'];
const tex = {};
for (const k of keys) tex[k] = device.createTexture(W, H);
let passes = null;
const timer = { record(p) { passes = p; } };
for (let f = 0; f < F; f++) {
  for (const k of keys) {
    const b = fs.readFileSync(path.join(dir, `f${f}_${k}.bin`));
    tex[k].upload(new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)));
  }
  const out = hush.denoise({ ...tex, extra: 'ignored' }, f === F - 1 ? timer : undefined);
  if (out.width !== W || out.height !== H) throw new Error('bad output size');
  const same = hush.output;
  if (same.surface.ptr !== out.surface.ptr) throw new Error('output getter differs from the returned surface');
  fs.writeFileSync(path.join(dir, `out${f}.bin`), Buffer.from(out.download().buffer));
  process.stdout.write(`frame ${f} ok\n`);
}
// a size mismatch raises the reference's message
const small = device.createTexture(W - 1, H);
let msg = '';
try { hush.denoise({ ...tex, motion: small }); } catch (e) { msg = e.message; }
if (msg !== `Hush.denoise: 'motion' is ${W - 1}x${H}, the denoiser ${W}x${H} (call resize)`) throw new Error('size message: ' + msg);
try { hush.denoise({ ...tex, emission: undefined }); } catch (e) { msg = e.message; }
if (msg !== "Hush.denoise: missing input 'emission'") throw new Error('missing message: ' + msg);
fs.writeFileSync(path.join(dir, 'timing.json'), JSON.stringify(passes));
hush.resize(W, H);               // same size: no-op, history kept
if (hush.frame !== F) throw new Error('resize to the same size must be a no-op');
hush.resize(W + 2, H);
if (hush.frame !== 0 || hush.width !== W + 2) throw new Error('resize must reset');
hush.destroy();
for (const k of keys) tex[k].destroy();
small.destroy();
console.log('denoise_sequence: ok');
