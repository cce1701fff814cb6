'use strict';
// Runs the JavaScript Post class on HIP device 0.
// usage: node post_sequence.js <dir> <width> <height>
// reads <dir>/left.bin and <dir>/right.bin (binary16, width*height*4 values), writes <dir>/target.bin
// (width*height*4 bytes, bgra8unorm) and checks the error message of a missing image.
const fs = require('fs');
const path = require('path');
const { Post, initDevice } = require(path.join(__dirname, '..', '..', 'hush_amd', 'node', 'hush.js'));

const [dir, W, H] = [process.argv[2], +process.argv[3], +process.argv[4]];
const device = initDevice(0);
const read = (k) => {
  const b = fs.readFileSync(path.join(dir, `${k}.bin`));
  return new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
};
const left = device.createTexture(W, H).upload(read('left'));
const right = device.createTexture(W, H).upload(read('right'));
const target = device.createCanvasTexture(W, H, 'bgra8unorm');
const post = new Post(device, 'bgra8unorm');
if (post.bloom !== 0.03 || post.vignette !== 0.35) throw new Error('defaults');
post.encode({ left, right, split: 0.4, exposure: 1.25, target, width: W, height: H });
fs.writeFileSync(path.join(dir, 'target.bin'), Buffer.from(target.download().buffer));
let msg = '';
try { post.encode({ left, target }); } catch (e) { msg = e.message; }
if (msg !== "Post.encode: missing image 'right'") throw new Error('missing message: ' + msg);
post.destroy();
left.destroy();
right.destroy();
target.destroy();
console.log('post_sequence: ok');
