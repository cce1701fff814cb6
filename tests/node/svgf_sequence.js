'use strict';
// Runs a G-buffer sequence through the JavaScript SVGF class on HIP device 0.
// usage: node svgf_sequence.js <dir> <width> <height> <frames>
// reads <dir>/f<i>_<input>.bin (binary16, width*height*4 values), ping-pongs normalDepth / motion as the
// previous frame's, writes <dir>/out<i>.bin and <dir>/timing.json (per-pass times of the last frame).
const fs = require('fs');
const path = require('path');
const { SVGF, initDevice } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

const [dir, W, H, F] = [process.argv[2], +process.argv[3], +process.argv[4], +process.argv[5]];
const device = initDevice(0);
const svgf = new SVGF(device, W, H, { iterations: 5 });
if (svgf.iterations !== 5 || svgf.params.sigmaN !== 128 || svgf.params.taa !== true) throw new Error('options');
const keys = ['radiance', 'normalDepth', 'motion', 'material', 'emission'];
const cur = {}, prev = {};
for (const k of keys) cur[k] = device.createTexture(W, H);
prev.normalDepth = device.createTexture(W, H);
prev.motion = device.createTexture(W, H);
let passes = null;
const timer = { record(p) { passes = p; } };
const read = (f, k) => {
  const b = fs.readFileSync(path.join(dir, `f${f}_${k}.bin`));
  return new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
};
for (let f = 0; f < F; f++) {
  for (const k of keys) cur[k].upload(read(f, k));
  prev.normalDepth.upload(read(f > 0 ? f - 1 : f, 'normalDepth'));
  prev.motion.upload(read(f > 0 ? f - 1 : f, 'motion'));
  const out = svgf.encode({ ...cur, prevNormalDepth: prev.normalDepth, prevMotion: prev.motion }, f === F - 1 ? timer : undefined);
  if (svgf.frame !== f + 1) throw new Error('frame counter');
  fs.writeFileSync(path.join(dir, `out${f}.bin`), Buffer.from(out.download().buffer));
}
fs.writeFileSync(path.join(dir, 'timing.json'), JSON.stringify(passes));
let msg = '';
try { svgf.encode({ ...cur, prevNormalDepth: prev.normalDepth }); } catch (e) { msg = e.message; }
if (msg !== "SVGF.encode: missing input 'prevMotion'") throw new Error('missing message: ' + msg);
svgf.destroy();
for (const k of keys) cur[k].destroy();
prev.normalDepth.destroy();
prev.motion.destroy();
console.log('svgf_sequence: ok');
