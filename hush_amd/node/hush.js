'use strict';
// Hush denoiser for MI355X: the reference's JavaScript API (src/denoise/hush.js:725-875, src/index.d.ts:36-60)
// over the N-API addon (hush_napi.node) and the HIP C ABI (include/hush.h).
//
// Device surfaces stand in for GPUTextures: rgba16float row-major buffers in HBM, created with
// `device.createTexture(width, height)` and filled with `upload(Uint16Array)`.  Everything else keeps its
// meaning: `new Hush(device, {width, height, exposure, weights, f16})`, `denoise(inputs, timer)`, `output`,
// `resize`, `reset`, `exposure`, `destroy`, `hiddenOf`, `Hush.requiredLimits`, and the error messages.

const fs = require('fs');
const path = require('path');
const native = require('./hush_napi.node');

const N_FEAT = 20;
const N_OUT = 12;
const INPUTS = ['radiance', 'normalDepth', 'motion', 'material', 'emission'];
const PRECISION = { fp32: 0, f16: 1, 'fp16-mfma': 2, 'bf16-mfma': 3, 'fp8-mfma': 4 };
const DEFAULT_WEIGHTS_PATH = path.join(__dirname, '..', 'data', 'hush-weights.bin');
const FP8_SCALES_PATH = path.join(__dirname, '..', 'data', 'fp8_scales.json');

// the calibrated fp8 input scales stored for these weights (their SHA-256 must match), or null
function storedFp8Scales(u16) {
  let d;
  try { d = JSON.parse(fs.readFileSync(FP8_SCALES_PATH, 'utf8')); } catch (e) { return null; }
  const digest = require('crypto').createHash('sha256')
    .update(Buffer.from(u16.buffer, u16.byteOffset, u16.byteLength)).digest('hex');
  return d.weights_sha256 === digest && Array.isArray(d.scales) && d.scales.length === 3 ? d.scales : null;
}

// hiddenOf (hush.js:22-25): the MLP width h with h*h + (N_FEAT + N_OUT + 2) h + N_OUT == n, 0 if none
function hiddenOf(n) {
  for (let h = 1; h <= 256; h++) if (h * h + (N_FEAT + N_OUT + 2) * h + N_OUT === n) return h;
  return 0;
}

// decodeWeights (hush.js:713-717): base64 string, typed array or ArrayBuffer -> Uint16Array
function decodeWeights(w) {
  if (typeof w === 'string') {
    const b = Buffer.from(w, 'base64');
    return new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
  }
  if (ArrayBuffer.isView(w)) return new Uint16Array(w.buffer.slice(w.byteOffset, w.byteOffset + w.byteLength));
  return new Uint16Array(w);
}

function defaultWeights() {
  const b = fs.readFileSync(DEFAULT_WEIGHTS_PATH);
  return new Uint16Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

// An rgba16float device surface (the GPUTexture of the reference's API).
class DeviceTexture {
  constructor(device, width, height, surface) {
    this.device = device;
    this.width = width;
    this.height = height;
    this.format = 'rgba16float';
    this._s = surface || native.alloc(device.ordinal, width, height);
    this._owned = !surface;
  }
  get surface() { return this._s; }
  // binary16 bit patterns, width * height * 4 values, row-major
  upload(data) { native.upload(this._s, data); return this; }
  download() { return native.download(this._s); }
  destroy() {
    if (this._owned && this._s) native.free(this._s);
    this._s = null;
  }
}

// The device: a HIP ordinal with the limits object the reference checks (gpu.js:3-22 initDevice).
function initDevice(ordinal = 0) {
  const n = native.deviceCount();
  if (!(ordinal >= 0 && ordinal < n)) throw new Error(`HIP device ${ordinal} not available (${n} devices)`);
  const device = {
    ordinal,
    limits: { maxStorageTexturesPerShaderStage: 8 },
    features: new Set(['shader-f16', 'timestamp-query']),
    createTexture(width, height) { return new DeviceTexture(device, width, height); },
    // an 8-bit colour target for Post (the canvas texture of the reference's demo)
    createCanvasTexture(width, height, format) { return new CanvasTexture(device, width, height, format); },
  };
  return device;
}

const REQUIRED_LIMITS = Object.freeze({ maxStorageTexturesPerShaderStage: 8 });

class Hush {
  // options: { width, height, exposure = 1, weights = built-in, f16 = false, precision = 'fp32', compensated = false,
  //            fp8Scales = calibrated for the built-in weights }
  constructor(device, options = {}) {
    const { width, height, exposure = 1, weights, f16 = false } = options;
    for (const [k, v] of Object.entries(REQUIRED_LIMITS)) {
      if (!(device.limits && device.limits[k] >= v)) {
        throw new Error(`Hush needs the device limit ${k} >= ${v} (this device: ${device.limits && device.limits[k]}); create the device with requiredLimits: Hush.requiredLimits`);
      }
    }
    this.device = device;
    const u16 = weights === undefined ? defaultWeights() : decodeWeights(weights);
    this.hidden = hiddenOf(u16.length);
    if (!this.hidden) throw new Error(`hush weights: ${u16.length} f16 values fit no MLP size`);
    if (!(width > 0 && height > 0)) throw new Error(`Hush: bad size ${width}x${height}`);
    const precision = options.precision !== undefined ? options.precision : (f16 ? 'f16' : 'fp32');
    this.f16 = precision === 'f16';
    this._h = native.create(device.ordinal, width, height, exposure, u16, PRECISION[precision]);
    if (options.compensated) native.setCompensated(this._h, true);   // compensated moment sums (a precision form)
    // fp8 form: static per-layer input scales (powers of two); default the calibration stored for these weights
    const fp8 = options.fp8Scales !== undefined ? options.fp8Scales : (this.hidden === 16 ? storedFp8Scales(u16) : null);
    if (fp8) native.setFp8Scales(this._h, fp8);
    this._exposure = exposure;
    this._timed = false;
  }

  // the fp8 form's static input scales of the three layers (each layer's input x enters as E4M3 of x / s)
  get fp8Scales() { return native.getFp8Scales(this._h); }
  set fp8Scales(s) { native.setFp8Scales(this._h, s); }

  get width() { return native.info(this._h).width; }
  get height() { return native.info(this._h).height; }
  get frame() { return native.info(this._h).frame; }

  // Display exposure the network sees (index.d.ts:42-43); read at the next frame.
  get exposure() { return this._exposure; }
  set exposure(e) { this._exposure = e; native.setExposure(this._h, e); }

  // Reallocates for a new input size (no-op if unchanged) and resets the history (hush.js:764-794).
  resize(width, height) {
    if (!(width > 0 && height > 0)) throw new Error(`Hush: bad size ${width}x${height}`);
    native.resize(this._h, width, height);
  }

  // Output of the last denoise() (hush.js:798): two surfaces alternate between frames.
  get output() { return new DeviceTexture(this.device, this.width, this.height, native.output(this._h)); }

  // Drops the history (camera cut, scene switch) (hush.js:801).
  reset() { native.reset(this._h); }

  destroy() {
    if (this._h) native.destroy(this._h);
    this._h = null;
  }

  // One frame (hush.js:813-822). inputs = { radiance, normalDepth, motion, material, emission }: DeviceTextures
  // of the denoiser's size (extra fields are ignored). timer (optional, index.d.ts:31-34): { desc(label) } times
  // every pass, { span: true, spanBegin, spanEnd } the whole denoiser as one span, as with the reference's
  // GpuTimer (this module's GpuTimer receives the times); an object with `record(passes)` receives
  // [{label, ms}] for the 8 passes of this frame.
  denoise(inputs, timer, stream) {
    for (const k of INPUTS) {
      const t = inputs[k];
      if (!t) throw new Error(`Hush.denoise: missing input '${k}'`);
      if (t.width !== this.width || t.height !== this.height) {
        throw new Error(`Hush.denoise: '${k}' is ${t.width}x${t.height}, the denoiser ${this.width}x${this.height} (call resize)`);
      }
    }
    return this.encode(inputs, timer, stream);
  }

  // denoise() without the input checks (hush.js:835-874).
  encode(inputs, timer, stream) {
    const surf = {};
    for (const k of INPUTS) surf[k] = inputs[k] && (inputs[k].surface || inputs[k]);
    // the timer's pass descriptors, asked for in the reference's order and with its labels (hush.js:846-849)
    const descs = timer ? passDescriptors(timer) : null;
    const writes = descs ? descs.filter((d) => d && d.timestampWrites) : [];
    const record = timer && typeof timer.record === 'function';
    const out = native.denoise(this._h, surf, stream || 0n, true, writes.length > 0 || record);
    if (writes.length || record) {
      const passes = native.passTimes(this._h);
      if (writes.length) writeTimestamps(descs, passes);
      if (record) timer.record(passes);
    }
    return new DeviceTexture(this.device, out.width, out.height, out);
  }
}

const PASS_LABELS = ['temporal', 'infer+mid', 'infer+mid', 'infer+mid', 'coarse', 'coarse', 'fine+final', 'taa'];

// The reference's timer protocol (hush.js:846-849, index.d.ts:31-34): a span timer gets spanBegin('total') on
// the first pass and spanEnd() on 'taa'; any other timer gets desc(label) for each of the 8 passes.
function passDescriptors(timer) {
  if (timer.span) {
    return PASS_LABELS.map((label, i) => {
      if (i === 0) return timer.spanBegin('total');
      return label === 'taa' ? timer.spanEnd() : undefined;
    });
  }
  if (typeof timer.desc === 'function') return PASS_LABELS.map((label) => timer.desc(label));
  return PASS_LABELS.map(() => undefined);
}

// Writes each pass's begin / end timestamps (ns from the frame's first event) where its descriptor's
// timestampWrites asks for them, as a compute pass's timestampWrites do (WebGPU); the query set must be a
// HIP timestamp set (GpuTimer below), whose `timestamps` array receives them.
function writeTimestamps(descs, passes) {
  let t = 0;
  passes.forEach((p, i) => {
    const begin = t;
    t += p.ms;
    const w = descs[i] && descs[i].timestampWrites;
    if (!w) return;
    const qs = w.querySet;
    if (!qs || !qs.timestamps) throw new Error('Hush.denoise: the timer\'s query set holds no HIP timestamps (use GpuTimer of this module)');
    if (w.beginningOfPassWriteIndex !== undefined) qs.timestamps[w.beginningOfPassWriteIndex] = begin * 1e6;
    if (w.endOfPassWriteIndex !== undefined) qs.timestamps[w.endOfPassWriteIndex] = t * 1e6;
  });
}

// Pass timer with the interface of the reference's GpuTimer (src/gpu.js:25-77): begin(), desc(label),
// spanBegin(label) / spanEnd(), resolve(encoder), read() -> Promise<{label: ms}> summed per label.  Its query
// set holds the timestamps in host memory (ns, Float64), filled by Hush.denoise from the frame's hipEvents.
class GpuTimer {
  constructor(device, max = 64) {
    this.device = device;
    this.enabled = !!(device && device.features && device.features.has('timestamp-query'));
    this.max = max;
    if (!this.enabled) return;
    this.qs = { type: 'timestamp', count: max * 2, timestamps: new Float64Array(max * 2) };
    this.labels = [];
    this.busy = false;
    this.resolved = false;
  }
  begin() { this.labels = []; this.resolved = false; }
  desc(label) {
    if (!this.enabled || this.busy || this.labels.length >= this.max) return undefined;
    const i = this.labels.length;
    this.labels.push(label);
    return { timestampWrites: { querySet: this.qs, beginningOfPassWriteIndex: i * 2, endOfPassWriteIndex: i * 2 + 1 } };
  }
  spanBegin(label) {
    if (!this.enabled || this.busy || this.labels.length >= this.max) return undefined;
    this.spanIndex = this.labels.length;
    this.labels.push(label);
    return { timestampWrites: { querySet: this.qs, beginningOfPassWriteIndex: this.spanIndex * 2 } };
  }
  spanEnd() {
    if (this.spanIndex === undefined) return undefined;
    const i = this.spanIndex;
    this.spanIndex = undefined;
    return { timestampWrites: { querySet: this.qs, endOfPassWriteIndex: i * 2 + 1 } };
  }
  // the timestamps are already in host memory once denoise() returns; kept for the reference's call sequence
  resolve() {
    if (!this.enabled || this.busy || !this.labels.length) return false;
    this.resolved = true;
    return true;
  }
  async read() {
    if (!this.enabled || this.busy || !this.labels.length) return null;
    const t = this.qs.timestamps;
    const out = {};
    this.labels.forEach((l, i) => { out[l] = (out[l] === undefined ? 0 : out[l]) + (t[i * 2 + 1] - t[i * 2]) / 1e6; });
    return out;
  }
}
Hush.requiredLimits = REQUIRED_LIMITS;

const SVGF_INPUTS = ['radiance', 'normalDepth', 'motion', 'material', 'emission', 'prevNormalDepth', 'prevMotion'];

// The SVGF baseline the reference exports beside Hush (src/denoise/svgf.js:273-379, src/index.d.ts:62-90):
// new SVGF(device, width, height, { sigmaL = 4, sigmaN = 128, sigmaZ = 1, alphaColor = 0.2,
// alphaMoments = 0.2, taaAlpha = 0.1, taa = true, iterations = 5 }); encode(g, timer, stream) with
// g = { radiance, normalDepth, motion, material, emission, prevNormalDepth, prevMotion } (the caller
// ping-pongs the previous frame's normalDepth / motion) returns this frame's output; `output` is the
// reference's getter, taa[frame & 1] after the frame advanced.
class SVGF {
  constructor(device, width, height, opts = {}) {
    this.device = device;
    this._s = native.svgfCreate(device.ordinal, width, height, opts);
    const i = native.svgfInfo(this._s);
    this.iterations = i.iterations;
    this.params = {
      sigmaL: i.sigmaL, sigmaN: i.sigmaN, sigmaZ: i.sigmaZ, alphaColor: i.alphaColor, alphaMoments: i.alphaMoments,
      taaAlpha: i.taaAlpha, taa: i.taa,
    };
  }

  get width() { return native.svgfInfo(this._s).width; }
  get height() { return native.svgfInfo(this._s).height; }
  get frame() { return native.svgfInfo(this._s).frame; }
  get output() { return new DeviceTexture(this.device, this.width, this.height, native.svgfOutput(this._s)); }

  // reallocates for a new size (no-op if unchanged), frame 0, history reset (svgf.js:298-319)
  resize(width, height) { native.svgfResize(this._s, width, height); }
  reset() { native.svgfReset(this._s); }
  destroy() {
    if (this._s) native.svgfDestroy(this._s);
    this._s = null;
  }

  // timer (optional, svgf.js:342-353): desc(label) is asked for each pass (temporal, variance, atrous per
  // iteration, resolve) and a GpuTimer of this module receives the times; `record(passes)` receives
  // [{label, ms}]
  encode(g, timer, stream) {
    const surf = {};
    for (const k of SVGF_INPUTS) surf[k] = g[k] && (g[k].surface || g[k]);
    const labels = ['temporal', 'variance', ...Array(this.iterations).fill('atrous'), 'resolve'];
    const descs = timer && typeof timer.desc === 'function' ? labels.map((l) => timer.desc(l)) : [];
    const writes = descs.filter((d) => d && d.timestampWrites);
    const record = timer && typeof timer.record === 'function';
    const out = native.svgfEncode(this._s, surf, stream || 0n, writes.length > 0 || record);
    if (writes.length || record) {
      const passes = native.svgfPassTimes(this._s);
      if (writes.length) writeTimestamps(descs, passes);
      if (record) timer.record(passes);
    }
    return new DeviceTexture(this.device, out.width, out.height, out);
  }
}

const POST_FORMATS = { rgba8unorm: 0, bgra8unorm: 1 };

// An 8-bit colour target (the canvas texture the reference's demo renders into): width x height 4-byte
// texels in `format` byte order.
class CanvasTexture {
  constructor(device, width, height, format = 'bgra8unorm') {
    if (!(format in POST_FORMATS)) throw new Error(`Post: unsupported target format '${format}'`);
    this.device = device;
    this.width = width;
    this.height = height;
    this.format = format;
    this._s = native.alloc8(device.ordinal, width, height);
  }
  get surface() { return this._s; }
  download() { return native.download8(this._s); }
  destroy() {
    if (this._s) native.free(this._s);
    this._s = null;
  }
}

// The display post of the reference's demo (src/demo/post.js:131-210): split-screen composite of two (or
// three) linear HDR images, bloom, vignette, AgX tone map, sRGB, divider lines.
// new Post(device, format, { bloom = 0.03, vignette = 0.35 }); encode({ left, right, mid = right,
// split = 0.5, split2 = split, same = false, exposure = 1, target, width, height, line = 1.5 }, stream)
// with DeviceTextures for the images and a CanvasTexture (or {ptr, pitch, width, height}) as the target.
class Post {
  constructor(device, format = 'bgra8unorm', opts = {}) {
    if (!(format in POST_FORMATS)) throw new Error(`Post: unsupported target format '${format}'`);
    const bloom = opts.bloom === undefined ? 0.03 : opts.bloom;
    const vignette = opts.vignette === undefined ? 0.35 : opts.vignette;
    this.device = device;
    this.format = format;
    this.bloom = bloom;
    this.vignette = vignette;
    this._p = native.postCreate(device.ordinal, { bloom, vignette });
  }

  // HDR and bloom surfaces for images of this size (no-op if unchanged); encode() calls it (post.js:149-158)
  resize(width, height) { native.postResize(this._p, width, height); }

  destroy() {
    if (this._p) native.postDestroy(this._p);
    this._p = null;
  }

  encode(args, stream) {
    const s = (t) => t && (t.surface || t);
    const split = args.split === undefined ? 0.5 : args.split;
    const target = args.target;
    native.postEncode(this._p, {
      left: s(args.left), right: s(args.right), mid: s(args.mid || args.right), split,
      split2: args.split2 === undefined ? split : args.split2, same: !!args.same,
      exposure: args.exposure === undefined ? 1 : args.exposure, line: args.line === undefined ? 1.5 : args.line,
      target: s(target), width: args.width === undefined ? target && target.width : args.width,
      height: args.height === undefined ? target && target.height : args.height, format: POST_FORMATS[this.format],
    }, stream || 0n);
    return target;
  }
}

module.exports = {
  Hush, SVGF, Post, GpuTimer, hiddenOf, initDevice, DeviceTexture, CanvasTexture, decodeWeights, defaultWeights, N_FEAT, N_OUT,
};
