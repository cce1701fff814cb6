'use strict';
// The reference's timer protocol (src/index.d.ts:31-34, hush.js:846-849, svgf.js:353, demo.js:126) through
// the JavaScript classes on HIP device 0.
// usage: node timer_sequence.js <dir> <width> <height> <frames>
// reads <dir>/f<i>_<input>.bin (binary16, width*height*4 values); writes <dir>/timer.json with the read()
// results of the per-pass and the span frames.
const fs = require('fs');
const path = require('path');
const { Hush, SVGF, GpuTimer, initDevice } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

async function main() {
  const [dir, W, H, F] = [process.argv[2], +process.argv[3], +process.argv[4], +process.argv[5]];
  const device = initDevice(0);
  const timed = new Hush(device, { width: W, height: H });
  const plain = new Hush(device, { width: W, height: H });
  const svgf = new SVGF(device, W, H, { iterations: 3 });
  const keys = ['radiance', 'normalDepth', 'motion', 'material', 'emission'];
  const tex = {};
  for (const k of keys) tex[k] = device.createTexture(W, H);
  const timer = new GpuTimer(device);
  if (!timer.enabled) throw new Error('GpuTimer must be enabled on a HIP device');
  const perPass = { desc: (l) => timer.desc(`hush.${l}`) };     // the demo's wrapper (demo.js:126)
  const span = { span: true, spanBegin: (l) => timer.spanBegin(l), spanEnd: () => timer.spanEnd() };
  const results = [];
  for (let f = 0; f < F; f++) {
    for (const k of keys) {
      const b = fs.readFileSync(path.join(dir, `f${f}_${k}.bin`));
      tex[k].upload(new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)));
    }
    timer.begin();
    const a = timed.denoise(tex, f % 2 === 0 ? perPass : span);
    if (f % 2 === 0) svgf.encode({ ...tex, prevNormalDepth: tex.normalDepth, prevMotion: tex.motion }, timer);
    if (!timer.resolve()) throw new Error('resolve');
    results.push(await timer.read());
    // timing runs the eight passes one by one: the same bits as the untimed frame
    const b = plain.denoise(tex);
    const x = Buffer.from(a.download().buffer), y = Buffer.from(b.download().buffer);
    if (!x.equals(y)) throw new Error(`frame ${f}: the timed frame differs from the untimed one`);
  }
  // a disabled timer (desc returns undefined) times nothing and raises nothing
  timed.denoise(tex, { desc: () => undefined });
  // a query set that is not a HIP timestamp set is refused with a message
  let msg = '';
  try {
    timed.denoise(tex, { desc: () => ({ timestampWrites: { querySet: {}, beginningOfPassWriteIndex: 0 } }) });
  } catch (e) { msg = e.message; }
  if (!msg.includes('HIP timestamps')) throw new Error('foreign query set: ' + msg);
  fs.writeFileSync(path.join(dir, 'timer.json'), JSON.stringify(results));
  timed.destroy();
  plain.destroy();
  svgf.destroy();
  for (const k of keys) tex[k].destroy();
  console.log('timer_sequence: ok');
}

main().catch((e) => { console.error(e); process.exit(1); });
