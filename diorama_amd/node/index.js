'use strict';
// The reference's operator classes for Node, over the N-API addon and the C ABI
// (include/diorama.h). Method and getter names follow src/prefixSum.ts,
// src/radixSort.ts, src/tileRasterizer.ts, src/gaussianRenderer.ts,
// src/trainingLoss.ts and src/trainer.ts, and so do the argument lists: encode
// calls take the reference's GPUCommandEncoder (or compute pass) first. Those come
// from device.createCommandEncoder() / beginComputePass() here. The library
// enqueues each call on the context's HIP stream at once, in call order, so an
// encoder is a token of the device it was made on, and finish() and
// queue.submit() keep the reference's call sequence without deferring anything.
// Targets the reference gives as rgba8unorm texture views are device.createTexture()
// views (one u32 per pixel) or plain Buffers. Methods the reference makes async
// (afterSubmit, read, step, ...) return Promises that resolve once the GPU work
// they wait for is done. Plain ES2019.

const path = require('path');
const native = require(path.join(__dirname, 'diorama.node'));

const FLOATS_PER_GAUSSIAN = 16; // struct Gaussian (src/gaussians.ts:32)
const FLOATS_PER_SPLAT = 12; // struct Splat2D
const SH_REST_FLOATS = 45;
const SH_C0 = 0.28209479177387814;

/** A device allocation; `owned` ones are freed by destroy(). */
class Buffer {
  constructor(handle, owned, keep) {
    this.handle = handle;
    this.owned = owned;
    this.keep = keep;
  }
  get size() {
    return native.buf_size(this.handle);
  }
  /** queue.writeBuffer: copies a TypedArray or ArrayBuffer in at `offset` bytes. */
  write(data, offset) {
    native.buf_upload(this.handle, offset || 0, data);
  }
  /** Reads `bytes` (default: all) from `offset` into a new ArrayBuffer, as mapAsync + getMappedRange would; waits for the GPU. */
  read(offset, bytes) {
    const from = offset || 0;
    return native.buf_download(this.handle, from, bytes === undefined ? this.size - from : bytes);
  }
  /** encoder.clearBuffer */
  clear() {
    native.buf_clear(this.handle);
  }
  /** encoder.copyBufferToBuffer(src, srcOffset, this, dstOffset, bytes) */
  copyFrom(src, srcOffset, dstOffset, bytes) {
    const from = srcOffset || 0;
    native.buf_copy(this.handle, dstOffset || 0, src.handle, from, bytes === undefined ? src.size - from : bytes);
  }
  destroy() {
    if (this.owned && this.handle) native.buf_destroy(this.handle);
    this.handle = null;
  }
}

function view(handle, keep) {
  return handle ? new Buffer(handle, false, keep) : undefined;
}

/** GPUComputePassEncoder: what PrefixSum/RadixSort.encode and the preprocess hooks receive. */
class ComputePassEncoder {
  constructor(encoder) {
    this.encoder = encoder;
    this.device = encoder.device;
  }
  end() {}
}

/** GPUCommandBuffer: what finish() returns and queue.submit() takes. */
class CommandBuffer {
  constructor(encoder) {
    this.encoder = encoder;
  }
}

/** GPUCommandEncoder. Work is already enqueued on the device's stream when recorded. */
class CommandEncoder {
  constructor(device) {
    this.device = device;
  }
  beginComputePass() {
    return new ComputePassEncoder(this);
  }
  copyBufferToBuffer(source, sourceOffset, destination, destinationOffset, size) {
    destination.copyFrom(source, sourceOffset, destinationOffset, size);
  }
  /** clearBuffer over the whole buffer (the only form the reference uses). */
  clearBuffer(buffer, offset, size) {
    if ((offset || 0) !== 0 || (size !== undefined && size !== buffer.size)) {
      throw new RangeError('clearBuffer clears whole buffers only');
    }
    buffer.clear();
  }
  finish() {
    return new CommandBuffer(this);
  }
}

/** An rgba8unorm GPUTexture: width x height pixels, one u32 each, row by row. */
class Texture {
  constructor(device, width, height) {
    this.width = width;
    this.height = height;
    this.format = 'rgba8unorm';
    this.buffer = device.createBuffer(Math.max(width * height, 1) * 4);
  }
  createView() {
    return new TextureView(this);
  }
  destroy() {
    this.buffer.destroy();
  }
}

class TextureView {
  constructor(texture) {
    this.texture = texture;
  }
}

/** The encoder argument of an encode call, checked: a misplaced argument list fails here. */
function encoderOf(encoder, what) {
  if (!(encoder instanceof CommandEncoder)) {
    throw new TypeError(what + ' takes a command encoder (device.createCommandEncoder()) as its first argument');
  }
  return encoder;
}

function passOf(pass, what) {
  if (!(pass instanceof ComputePassEncoder)) {
    throw new TypeError(what + ' takes a compute pass (encoder.beginComputePass()) as its first argument');
  }
  return pass;
}

/** The device buffer behind a render target or photo: a texture, its view, a Buffer, or none. */
function surface(target) {
  if (target === null || target === undefined) return null;
  if (target instanceof TextureView) return target.texture.buffer.handle;
  if (target instanceof Texture) return target.buffer.handle;
  if (target instanceof Buffer) return target.handle;
  throw new TypeError('a target must be a texture, a texture view or a Buffer');
}

/** A training photo as host rgba8 words: Uint32Array as given, or read back from a texture. */
function hostImage(image) {
  if (image instanceof Texture) return new Uint32Array(image.buffer.read(0, image.width * image.height * 4));
  return image;
}

/** One device and its in-order stream: the counterpart of initGpu (src/gpu.ts:8-47). */
class Device {
  constructor(index) {
    this.handle = native.ctx_create(index || 0);
  }
  /** createBuffer({size, usage, label}) as in WebGPU, or createBuffer(bytes). */
  createBuffer(descriptor) {
    const bytes = typeof descriptor === 'number' ? descriptor : descriptor.size;
    return new Buffer(native.buf_create(this.handle, bytes), true);
  }
  /** createTexture({size: [w, h] | {width, height}, format: 'rgba8unorm'}). */
  createTexture(descriptor) {
    const size = descriptor.size;
    const width = Array.isArray(size) ? size[0] : size.width;
    const height = Array.isArray(size) ? (size[1] === undefined ? 1 : size[1]) : (size.height === undefined ? 1 : size.height);
    if (descriptor.format !== undefined && descriptor.format !== 'rgba8unorm') {
      throw new TypeError('only rgba8unorm textures are supported');
    }
    return new Texture(this, width, height);
  }
  createCommandEncoder() {
    return new CommandEncoder(this);
  }
  get queue() {
    const device = this;
    return {
      submit(commandBuffers) {
        for (const c of commandBuffers || []) {
          if (!(c instanceof CommandBuffer) || c.encoder.device !== device) {
            throw new TypeError('submit takes command buffers finished on this device');
          }
        }
      },
      /** writeBuffer(buffer, bufferOffset, data, dataOffset, size): offsets and size in elements of a TypedArray. */
      writeBuffer(buffer, bufferOffset, data, dataOffset, size) {
        const element = ArrayBuffer.isView(data) ? data.BYTES_PER_ELEMENT || 1 : 1;
        const base = ArrayBuffer.isView(data) ? data.byteOffset : 0;
        const total = ArrayBuffer.isView(data) ? data.byteLength / element : data.byteLength;
        const from = dataOffset || 0;
        const count = size === undefined ? total - from : size;
        const source = ArrayBuffer.isView(data) ? data.buffer : data;
        buffer.write(new Uint8Array(source, base + from * element, count * element), bufferOffset);
      },
      /** writeTexture({texture}, data, {bytesPerRow}, [w, h]) for whole rgba8 images. */
      writeTexture(destination, data, layout, size) {
        const texture = destination.texture;
        const width = Array.isArray(size) ? size[0] : size.width;
        const height = Array.isArray(size) ? size[1] : size.height;
        if (width !== texture.width || height !== texture.height || (layout.bytesPerRow && layout.bytesPerRow !== width * 4)) {
          throw new RangeError('writeTexture writes whole, tightly packed rgba8 images');
        }
        texture.buffer.write(data, 0);
      },
      onSubmittedWorkDone() {
        return device.onSubmittedWorkDone();
      },
    };
  }
  upload(data) {
    const b = this.createBuffer(data.byteLength);
    b.write(data);
    return b;
  }
  synchronize() {
    native.ctx_synchronize(this.handle);
  }
  /** Timestamp queries: record slot `slot` (0-63), then read the ms between two slots. */
  timestamp(slot) {
    native.ctx_timestamp(this.handle, slot);
  }
  elapsed(fromSlot, toSlot) {
    return native.ctx_elapsed(this.handle, fromSlot, toSlot);
  }
  onSubmittedWorkDone() {
    this.synchronize();
    return Promise.resolve();
  }
  destroy() {
    if (this.handle) native.ctx_destroy(this.handle);
    this.handle = null;
  }
}

function initGpu(index) {
  return new Device(index);
}

/** src/prefixSum.ts:10-103 */
class PrefixSum {
  constructor(device, data, capacity, count) {
    this.handle = native.prefix_sum_create(device.handle, data.handle, capacity, count ? count.handle : null);
    this.keep = [data, count];
  }
  get total() {
    return view(native.prefix_sum_total(this.handle), this);
  }
  encode(pass) {
    passOf(pass, 'PrefixSum.encode');
    native.prefix_sum_encode(this.handle);
  }
  destroy() {
    if (this.handle) native.prefix_sum_destroy(this.handle);
    this.handle = null;
  }
}

/** src/radixSort.ts:27-142 */
class RadixSort {
  constructor(device, capacity, options) {
    const opts = options || {};
    this.handle = native.radix_sort_create(device.handle, capacity, opts.highBits || 0,
      opts.count ? opts.count.handle : null);
    this.keep = opts.count;
  }
  get capacity() {
    return native.radix_sort_capacity(this.handle);
  }
  get keys() {
    return view(native.radix_sort_keys(this.handle), this);
  }
  get values() {
    return view(native.radix_sort_values(this.handle), this);
  }
  encode(pass) {
    passOf(pass, 'RadixSort.encode');
    native.radix_sort_encode(this.handle);
  }
  destroy() {
    if (this.handle) native.radix_sort_destroy(this.handle);
    this.handle = null;
  }
}

/**
 * src/tileRasterizer.ts:29-317, for `count` splats already on the device.
 * `camera` is the Camera uniform buffer (144 bytes: view, proj, viewport) as in
 * the reference; the tile stage takes its viewport from encode()'s `viewport`,
 * which the reference's callers always pass equal to the camera's.
 */
class TileRasterizer {
  constructor(device, camera, splats, count, initialPairCapacity) {
    if (!(camera instanceof Buffer) || camera.size < 144 || !(splats instanceof Buffer) || typeof count !== 'number') {
      throw new TypeError('TileRasterizer takes (device, camera buffer of 144 bytes, splats buffer, count, ' +
        'initialPairCapacity?)');
    }
    this.handle = native.tiles_create(device.handle, splats.handle, count, initialPairCapacity || 0);
    this.owned = true;
    this.camera = camera;
    this.keep = splats;
  }
  /** The rasterizer a renderer owns: a view that destroy() leaves alone. */
  static wrap(handle, keep) {
    const t = Object.create(TileRasterizer.prototype);
    t.handle = handle;
    t.owned = false;
    t.keep = keep;
    return t;
  }
  get lastPairCount() {
    return native.tiles_last_pair_count(this.handle);
  }
  get pairCapacity() {
    return native.tiles_pair_capacity(this.handle);
  }
  get debugBuffers() {
    return { pairSplats: view(native.tiles_pair_splats(this.handle), this), ranges: view(native.tiles_ranges(this.handle), this) };
  }
  get pixelColor() {
    return view(native.tiles_pixel_color(this.handle), this);
  }
  get pixelLast() {
    return view(native.tiles_pixel_last(this.handle), this);
  }
  /** encode(encoder, target, viewport, background, before?): `before(pass)` runs ahead of the tile stage. */
  encode(encoder, target, viewport, background, before) {
    const enc = encoderOf(encoder, 'TileRasterizer.encode');
    if (before) before(enc.beginComputePass());
    native.tiles_encode(this.handle, surface(target), viewport[0], viewport[1], background);
  }
  encodeBackward(encoder, viewport, pixelGrads, splatGrads) {
    encoderOf(encoder, 'TileRasterizer.encodeBackward');
    native.tiles_encode_backward(this.handle, viewport[0], viewport[1], pixelGrads.handle, splatGrads.handle);
  }
  afterSubmit() {
    native.tiles_after_submit(this.handle);
    return Promise.resolve();
  }
  /** Launch settings (see dio_tile_rasterizer_set_tuning); results do not depend on them. */
  setTuning(k) {
    native.tiles_set_tuning(this.handle, [k.tileSize || 16, k.splatBatch || 64, k.digitBits || 8, k.persistent || 0,
      k.segment || 0, k.perTileDepth || 0]);
  }
  destroy() {
    if (this.owned && this.handle) native.tiles_destroy(this.handle);
    this.handle = null;
  }
}

/** src/gaussianRenderer.ts:55-311 */
class GaussianRenderer {
  constructor(device, options) {
    this.device = device;
    const opts = options || {};
    this.handle = native.renderer_create(device.handle, opts.initialPairCapacity || 0);
    this.owned = true;
  }
  /** The renderer a trainer owns: a view that destroy() leaves alone. */
  static wrap(device, handle, keep) {
    const r = Object.create(GaussianRenderer.prototype);
    r.device = device;
    r.handle = handle;
    r.owned = false;
    r.keep = keep;
    return r;
  }
  /** setGaussians(Float32Array of 16 floats each, {rest, degree}) */
  setGaussians(data, sh) {
    native.renderer_set_gaussians(this.handle, data, sh ? sh.rest : null, sh ? sh.degree : 0);
  }
  adoptGaussians(count, gaussians, sh, degree) {
    native.renderer_adopt_gaussians(this.handle, count, gaussians.handle, sh.handle, degree);
    gaussians.owned = false;
    sh.owned = false;
  }
  setShDegree(degree) {
    native.renderer_set_sh_degree(this.handle, degree);
  }
  /**
   * encode(encoder, target, camera, background, afterPreprocess?): the hook gets
   * a compute pass between the preprocess and the tiles, as in the reference,
   * and only when there are Gaussians (with none the target is just cleared).
   */
  encode(encoder, target, camera, background, afterPreprocess) {
    const enc = encoderOf(encoder, 'GaussianRenderer.encode');
    native.renderer_encode_preprocess(this.handle, camera);
    if (afterPreprocess && native.renderer_splat_buffer(this.handle)) afterPreprocess(enc.beginComputePass());
    native.renderer_encode_tiles(this.handle, surface(target), background);
  }
  encodeDraw(encoder, target, splats, count, viewport, background) {
    encoderOf(encoder, 'GaussianRenderer.encodeDraw');
    native.renderer_encode_draw(this.handle, surface(target), splats.handle, count, viewport[0], viewport[1],
      background);
  }
  encodeBackward(encoder, pixelGrads) {
    encoderOf(encoder, 'GaussianRenderer.encodeBackward');
    native.renderer_encode_backward(this.handle, pixelGrads.handle);
  }
  afterSubmit() {
    native.renderer_after_submit(this.handle);
    return Promise.resolve();
  }
  get count() {
    return native.renderer_count(this.handle);
  }
  get splatBuffer() {
    return view(native.renderer_splat_buffer(this.handle), this);
  }
  get gaussianBuffer() {
    return view(native.renderer_gaussian_buffer(this.handle), this);
  }
  get shBuffer() {
    return view(native.renderer_sh_buffer(this.handle), this);
  }
  get tiles() {
    const h = native.renderer_tiles(this.handle);
    return h ? TileRasterizer.wrap(h, this) : undefined;
  }
  get pixelBuffer() {
    return view(native.renderer_pixel_buffer(this.handle), this);
  }
  get pixelLastBuffer() {
    return view(native.renderer_pixel_last_buffer(this.handle), this);
  }
  get gradients() {
    const g = native.renderer_gradient_gaussians(this.handle);
    if (!g) return undefined;
    return {
      gaussians: view(g, this),
      sh: view(native.renderer_gradient_sh(this.handle), this),
      splats: view(native.renderer_gradient_splats(this.handle), this),
    };
  }
  destroy() {
    if (this.owned && this.handle) native.renderer_destroy(this.handle);
    this.handle = null;
  }
}

/** src/trainingLoss.ts:16-130 */
class TrainingLoss {
  constructor(device, width, height) {
    this.width = width;
    this.height = height;
    this.handle = native.loss_create(device.handle, width, height);
  }
  /** encode(encoder, rendered, target): `target` is the photo's texture view (or a Buffer of rgba8 words). */
  encode(encoder, rendered, target) {
    encoderOf(encoder, 'TrainingLoss.encode');
    native.loss_encode(this.handle, rendered.handle, surface(target));
  }
  read() {
    return Promise.resolve(native.loss_read(this.handle));
  }
  get pixelGrads() {
    return view(native.loss_pixel_grads(this.handle), this);
  }
  get pixelLoss() {
    return view(native.loss_pixel_loss(this.handle), this);
  }
  destroy() {
    if (this.handle) native.loss_destroy(this.handle);
    this.handle = null;
  }
}

/** The shortest decimal that rounds to the same float32 (the C ABI stores floats). */
function shortestFloat32(x) {
  for (let p = 1; p <= 9; p++) {
    const y = Number(x.toPrecision(p));
    if (Math.fround(y) === Math.fround(x)) return y;
  }
  return x;
}

/** DEFAULT_SETTINGS (src/trainer.ts:65-82), without the extent. */
const DEFAULT_SETTINGS = (() => {
  const s = native.default_settings(1);
  delete s.extent;
  for (const key of Object.keys(s)) {
    if (Array.isArray(s[key])) s[key] = s[key].map(shortestFloat32);
    else s[key] = shortestFloat32(s[key]);
  }
  return Object.freeze(s);
})();

/**
 * src/trainer.ts:103-446. `raw` = {params: Float32Array (16 per Gaussian), sh:
 * Float32Array (45 per Gaussian)}; `views` = [{name, camera, image}] where the
 * image is an rgba8unorm texture (as in the reference) or a Uint32Array of rgba8 words.
 * options.batch > 0 makes a multi-view batch trainer: that many views (of any
 * sizes) per step, their gradients summed into one Adam step.
 */
class Trainer {
  constructor(device, raw, views, settings, options) {
    this.device = device;
    this.settings = settings;
    const batch = options && options.batch ? options.batch : 0;
    const hostViews = views.map((v) => Object.assign({}, v, { image: hostImage(v.image) }));
    this.handle = native.trainer_create(device.handle, raw.params, raw.sh, hostViews, settings, batch);
  }
  get batchSize() {
    return native.trainer_batch_size(this.handle);
  }
  /** mode 'plain' | 'log' | 'learned' (see dio_trainer_set_densify_mode). */
  setDensifyMode(mode, mlp, options) {
    const o = options || {};
    const code = { plain: 0, log: 1, learned: 2 }[mode];
    native.trainer_set_densify_mode(this.handle, code === undefined ? mode : code, mlp ? mlp.handle : null,
      o.budget || 0, o.confidence === undefined ? 0.5 : o.confidence, o.samplesPerDensify || 80000);
  }
  get tuning() {
    return native.trainer_get_tuning(this.handle);
  }
  set tuning(k) {
    native.trainer_set_tuning(this.handle, [k.tileSize || 16, k.splatBatch || 64, k.digitBits || 8, k.persistent || 0,
      k.segment || 0, k.perTileDepth || 0]);
  }
  /** Times every candidate setting over `steps` training steps and keeps the fastest. */
  autotune(steps) {
    return native.trainer_autotune(this.handle, steps || 3);
  }
  samples() {
    const s = native.trainer_samples(this.handle);
    return { features: view(s.features, this), labels: view(s.labels, this), count: s.count };
  }
  features() {
    return view(native.trainer_features(this.handle), this);
  }
  get count() {
    return native.trainer_count(this.handle);
  }
  get iteration() {
    return native.trainer_iteration(this.handle);
  }
  set iteration(value) {
    native.trainer_set_iteration(this.handle, value);
  }
  get shDegree() {
    return native.trainer_sh_degree(this.handle);
  }
  set shDegree(value) {
    native.trainer_set_sh_degree(this.handle, value);
  }
  get densifications() {
    return native.trainer_densifications(this.handle);
  }
  get buffers() {
    return {
      params: view(native.trainer_params(this.handle), this),
      moments: view(native.trainer_moments(this.handle), this),
      stats: view(native.trainer_stats(this.handle), this),
    };
  }
  get renderer() {
    return GaussianRenderer.wrap(this.device, native.trainer_renderer(this.handle), this);
  }
  /** writeGaussianPly(await trainer.exportRaw()) without the host round trip. */
  writePly() {
    return writeGaussianPly({ params: this.buffers.params, sh: this.renderer.shBuffer }, null, this.count);
  }
  step() {
    native.trainer_step(this.handle);
    return Promise.resolve();
  }
  optimize() {
    native.trainer_optimize(this.handle);
  }
  maintain() {
    native.trainer_maintain(this.handle);
    return Promise.resolve();
  }
  densify(pruneLarge) {
    native.trainer_densify(this.handle, !!pruneLarge);
    return Promise.resolve();
  }
  resetOpacity() {
    native.trainer_reset_opacity(this.handle);
  }
  exportRaw() {
    return Promise.resolve(native.trainer_export_raw(this.handle));
  }
  lastLoss() {
    return Promise.resolve(native.trainer_last_loss(this.handle));
  }
  evaluate(views, withSsim) {
    const hostViews = views.map((v) => Object.assign({}, v, { image: hostImage(v.image) }));
    return Promise.resolve(native.trainer_evaluate(this.handle, hostViews, withSsim === undefined ? true : withSsim));
  }
  /** measurePsnr (src/insideView.ts:233-267): views [{camera, image}] whose images are device Buffers. */
  measurePsnr(views) {
    return Promise.resolve(native.trainer_measure_psnr(this.handle, views.map((v) => v.camera),
      views.map((v) => surface(v.image))));
  }
  destroy() {
    if (this.handle) native.trainer_destroy(this.handle);
    this.handle = null;
  }
}

/**
 * The learned densification proposal (no counterpart in the reference): a
 * 12 -> 64 -> 64 -> 4 MLP with BF16 weights and FP32 accumulation.
 */
class DensifyMLP {
  constructor(device, seed) {
    this.handle = native.mlp_create(device.handle, seed || 1);
  }
  get weights() {
    return native.mlp_get_weights(this.handle);
  }
  set weights(w) {
    native.mlp_set_weights(this.handle, w);
  }
  setInputNorm(mean, scale) {
    native.mlp_set_input_norm(this.handle, mean, scale);
  }
  setClassWeights(weights) {
    native.mlp_set_class_weights(this.handle, weights);
  }
  forward(features, count, logits) {
    native.mlp_forward(this.handle, features.handle, count, logits.handle);
  }
  gradients(features, labels, count) {
    return native.mlp_gradients(this.handle, features.handle, labels.handle, count);
  }
  train(features, labels, count, steps, batch, lr, seed) {
    return native.mlp_train(this.handle, features.handle, labels.handle, count, steps, batch || 4096, lr || 1e-3,
      seed || 1);
  }
  destroy() {
    if (this.handle) native.mlp_destroy(this.handle);
    this.handle = null;
  }
}

/**
 * parseColmap (src/colmap.ts:18-32): {cameras (TrainingCamera objects sorted by
 * name), positions (Float32Array xyz), colors (Uint8Array rgb)} from the three
 * binary files' ArrayBuffers.
 */
function parseColmap(camerasBin, imagesBin, pointsBin) {
  return native.colmap_parse(camerasBin, imagesBin, pointsBin);
}

/** trainingCameraData (src/loadDataset.ts:80-86): the renderer's camera for a TrainingCamera at `viewport`. */
function trainingCameraData(camera, viewport) {
  return native.colmap_training_camera(camera, viewport[0], viewport[1]);
}

/**
 * meanNeighborDistance2 (src/loadDataset.ts:126-149): per point (Float32Array of
 * xyz) the mean squared distance to its 3 nearest other points.
 */
function meanNeighborDistance2(device, positions) {
  const n = positions.length / 3;
  const points = device.upload(positions);
  const out = device.createBuffer(Math.max(n, 1) * 4);
  try {
    native.mean_neighbor_distance2(device.handle, points.handle, n, out.handle);
    return Promise.resolve(new Float32Array(out.read(0, n * 4)));
  } finally {
    points.destroy();
    out.destroy();
  }
}

/**
 * initialGaussians (src/loadDataset.ts:87-110): RawGaussians {params, sh} for
 * structure-from-motion points (Float32Array xyz) and colours (Uint8Array rgb).
 */
function initialGaussians(device, positions, colors) {
  const n = positions.length / 3;
  const points = device.upload(positions);
  const rgb = device.upload(colors);
  const params = device.createBuffer(Math.max(n, 1) * FLOATS_PER_GAUSSIAN * 4);
  const sh = device.createBuffer(Math.max(n, 1) * SH_REST_FLOATS * 4);
  try {
    native.initial_gaussians(device.handle, points.handle, rgb.handle, n, params.handle, sh.handle);
    return Promise.resolve({
      params: new Float32Array(params.read(0, n * FLOATS_PER_GAUSSIAN * 4)),
      sh: new Float32Array(sh.read(0, n * SH_REST_FLOATS * 4)),
    });
  } finally {
    for (const b of [points, rgb, params, sh]) b.destroy();
  }
}

/**
 * writeGaussianPly (src/plyWriter.ts:14-43). `raw` is {params, sh}: Float32Arrays
 * (uploaded on `device`) or device Buffers (then `count` Gaussians; the body is
 * packed where they live). Returns a Blob-like {size, type, arrayBuffer()}, as the
 * reference's Blob is read with `await blob.arrayBuffer()`.
 */
function writeGaussianPly(raw, device, count) {
  let params = raw.params;
  let sh = raw.sh;
  const staged = [];
  let n = count;
  if (!(params instanceof Buffer)) {
    if (!device) throw new Error('writeGaussianPly needs a Device for host arrays');
    n = n === undefined ? params.length / FLOATS_PER_GAUSSIAN : n;
    params = device.upload(raw.params);
    sh = device.upload(raw.sh);
    staged.push(params, sh);
  } else if (n === undefined) {
    n = params.size / (FLOATS_PER_GAUSSIAN * 4);
  }
  try {
    const file = native.ply_write(params.handle, sh.handle, n);
    return { size: file.byteLength, type: 'application/octet-stream', arrayBuffer: () => Promise.resolve(file) };
  } finally {
    for (const b of staged) b.destroy();
  }
}

/**
 * parseGaussianPly (src/plyLoader.ts:21-98): {count, gaussians, shRest, shDegree}
 * with Float32Arrays, as the reference returns; with {onDevice: true} the two
 * arrays stay device Buffers, ready for GaussianRenderer.adoptGaussians.
 */
function parseGaussianPly(buffer, device, options) {
  const o = native.ply_parse(device.handle, buffer);
  const gaussians = new Buffer(o.gaussians, true);
  const shRest = new Buffer(o.shRest, true);
  if (options && options.onDevice) return { count: o.count, gaussians, shRest, shDegree: o.shDegree };
  const out = {
    count: o.count,
    gaussians: new Float32Array(gaussians.read(0, o.count * FLOATS_PER_GAUSSIAN * 4)),
    shRest: new Float32Array(shRest.read(0, o.count * SH_REST_FLOATS * 4)),
    shDegree: o.shDegree,
  };
  gaussians.destroy();
  shRest.destroy();
  return out;
}

/** readBuffer (src/readback.ts:4-19): the whole buffer as an ArrayBuffer. */
function readBuffer(device, buffer) {
  return Promise.resolve(buffer.read());
}

/** readTexture (src/readback.ts:22-44): rgba8 bytes, row by row. */
function readTexture(device, texture) {
  return Promise.resolve(new Uint8Array(texture.buffer.read(0, texture.width * texture.height * 4)));
}

/** Raw parameters to the renderer's activated layout (src/trainer.ts:449-461). */
function activate(params) {
  const out = new Float32Array(params.length);
  for (let at = 0; at < params.length; at += FLOATS_PER_GAUSSIAN) {
    for (let k = 0; k < 3; k++) {
      out[at + k] = params[at + k];
      out[at + 4 + k] = Math.exp(params[at + 4 + k]);
      out[at + 12 + k] = 0.5 + SH_C0 * params[at + 12 + k];
    }
    out[at + 3] = 1 / (1 + Math.exp(-params[at + 3]));
    for (let k = 8; k < 12; k++) out[at + k] = params[at + k];
  }
  return out;
}

module.exports = {
  version: native.version,
  initGpu,
  Device,
  Buffer,
  CommandEncoder,
  ComputePassEncoder,
  Texture,
  readBuffer,
  readTexture,
  PrefixSum,
  RadixSort,
  TileRasterizer,
  GaussianRenderer,
  TrainingLoss,
  Trainer,
  DensifyMLP,
  DEFAULT_SETTINGS,
  writeGaussianPly,
  parseGaussianPly,
  meanNeighborDistance2,
  initialGaussians,
  parseColmap,
  trainingCameraData,
  activate,
  FLOATS_PER_GAUSSIAN,
  FLOATS_PER_SPLAT,
  SH_REST_FLOATS,
};
