'use strict';
// Drives the Node classes (diorama_amd/node) the way the reference's checks drive
// its operators, with the reference's own argument lists (an encoder or compute
// pass first, texture views as targets): the prefix sum and radix sort against
// plain JS (src/checks/depthSort.ts:46-104), the loss fixture
// (src/checks/loss.ts:24-89) and the backward fixture (src/checks/backward.ts:49-114),
// plus a standalone TileRasterizer over the renderer's splats.
// Prints one JSON line of results; tests/test_node.py applies the bars.
// Usage: node tests/node_parity.js <cpu|gpu>

const fs = require('fs');
const path = require('path');
const d = require(path.join(__dirname, '..', 'diorama_amd', 'node'));

function fixture(name) {
  return JSON.parse(fs.readFileSync(path.join(__dirname, 'golden', name), 'utf8'));
}

function rng(seed) {
  let s = seed >>> 0;
  return () => {
    s = (s + 0x6d2b79f5) >>> 0;
    let t = s;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

function worstDiff(a, b) {
  let most = 0;
  for (let i = 0; i < a.length; i++) most = Math.max(most, Math.abs(a[i] - b[i]));
  return most;
}

function maxAbs(a) {
  let most = 0;
  for (let i = 0; i < a.length; i++) most = Math.max(most, Math.abs(a[i]));
  return most;
}

function lookAt(eye) {
  const len = Math.hypot(eye[0], eye[1], eye[2]);
  const z = eye.map((v) => v / len);
  const xl = Math.hypot(z[2], z[0]);
  const x = [z[2] / xl, 0, -z[0] / xl];
  const y = [z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]];
  const dot = (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  return [x[0], y[0], z[0], 0, x[1], y[1], z[1], 0, x[2], y[2], z[2], 0, -dot(x, eye), -dot(y, eye), -dot(z, eye), 1];
}

function perspective(fov, aspect) {
  const f = 1 / Math.tan(fov / 2);
  const near = 0.01, far = 100, depth = 1 / (near - far);
  return [f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, far * depth, -1, 0, 0, near * far * depth, 0];
}

function cpuChecks() {
  const out = { version: d.version(), defaults: d.DEFAULT_SETTINGS };
  const errors = [];
  try {
    d.initGpu(4096);
  } catch (e) {
    errors.push(e.message);
  }
  try {
    new d.PrefixSum({ handle: null }, { handle: null }, 16);
  } catch (e) {
    errors.push(e.message);
  }
  try {
    d.parseGaussianPly(new ArrayBuffer(8), { handle: null });
  } catch (e) {
    errors.push(e.message);
  }
  out.errors = errors;
  out.colmap = colmapCheck();
  const raw = new Float32Array(16);
  raw[3] = 0;
  raw[4] = Math.log(2);
  raw[12] = 1;
  out.activated = Array.from(d.activate(raw));
  return out;
}

// A two-photo reconstruction written by hand: camera 1 SIMPLE_PINHOLE f = 500 at
// 640 x 480; photo "b.jpg" at the identity pose moved by t = (0, 0, 2), photo
// "a.jpg" turned 90 degrees about y; three points.
function colmapCheck() {
  const bytes = (n, fill) => {
    const v = new DataView(new ArrayBuffer(n));
    fill(v);
    return v.buffer;
  };
  const cams = bytes(8 + 24 + 24, (v) => {
    v.setBigUint64(0, 1n, true);
    v.setInt32(8, 1, true); v.setInt32(12, 0, true);
    v.setBigUint64(16, 640n, true); v.setBigUint64(24, 480n, true);
    v.setFloat64(32, 500, true); v.setFloat64(40, 320, true); v.setFloat64(48, 240, true);
  });
  const image = (v, at, id, q, t, name) => {
    v.setInt32(at, id, true);
    q.forEach((x, k) => v.setFloat64(at + 4 + 8 * k, x, true));
    t.forEach((x, k) => v.setFloat64(at + 36 + 8 * k, x, true));
    v.setInt32(at + 60, 1, true);
    for (let k = 0; k < name.length; k++) v.setUint8(at + 64 + k, name.charCodeAt(k));
    v.setUint8(at + 64 + name.length, 0);
    v.setBigUint64(at + 65 + name.length, 0n, true);
    return at + 73 + name.length;
  };
  const s = Math.SQRT1_2;
  const imgs = bytes(8 + 2 * 78, (v) => {
    v.setBigUint64(0, 2n, true);
    const at = image(v, 8, 1, [1, 0, 0, 0], [0, 0, 2], 'b.jpg');
    image(v, at, 2, [s, 0, s, 0], [0, 0, 0], 'a.jpg');
  });
  const pts = bytes(8 + 3 * 51, (v) => {
    v.setBigUint64(0, 3n, true);
    for (let i = 0; i < 3; i++) {
      const at = 8 + 51 * i;
      v.setBigUint64(at, BigInt(i), true);
      [i, 2 * i, -i].forEach((x, k) => v.setFloat64(at + 8 + 8 * k, x, true));
      [10 * i, 20 * i, 30 * i].forEach((x, k) => v.setUint8(at + 32 + k, x));
      v.setFloat64(at + 35, 0.1, true);
      v.setBigUint64(at + 43, 0n, true);
    }
  });
  const r = d.parseColmap(cams, imgs, pts);
  const cam = r.cameras.find((c) => c.name === 'b.jpg');
  return { names: r.cameras.map((c) => c.name), position: cam.position, forwardA: r.cameras[0].forward,
    positions: Array.from(r.positions), colors: Array.from(r.colors),
    view: d.trainingCameraData(cam, [640, 480]).view, fy: cam.fy };
}

async function gpuChecks() {
  const out = {};
  const device = d.initGpu(0);

  // Prefix sum over a ragged count (src/checks/prefixSum.ts).
  const random = rng(3);
  const n = 1000003;
  const values = new Uint32Array(n);
  for (let i = 0; i < n; i++) values[i] = Math.floor(random() * 100);
  const data = device.upload(values);
  const scan = new d.PrefixSum(device, data, n);
  let encoder = device.createCommandEncoder();
  let pass = encoder.beginComputePass();
  scan.encode(pass);
  pass.end();
  device.queue.submit([encoder.finish()]);
  const scanned = new Uint32Array(data.read());
  let running = 0;
  let scanBad = 0;
  for (let i = 0; i < n; i++) {
    if (scanned[i] !== running) scanBad++;
    running += values[i];
  }
  const total = new Uint32Array(scan.total.read(0, 4))[0];
  out.scan = { mismatches: scanBad, total, expected: running };
  scan.destroy();
  data.destroy();

  // Radix sort, 32-bit keys and 64-bit (low, high) keys with 16 high bits
  // (src/checks/radixSort.ts); stable, values carried along.
  out.sort = {};
  for (const highBits of [0, 16]) {
    const m = 300001;
    const sort = new d.RadixSort(device, m, { highBits });
    const words = highBits ? 2 : 1;
    const keys = new Uint32Array(m * words);
    const vals = new Uint32Array(m);
    for (let i = 0; i < m; i++) {
      keys[i * words] = Math.floor(random() * 4294967296) >>> 0;
      if (highBits) keys[i * words + 1] = Math.floor(random() * 64);
      vals[i] = i;
    }
    const high = (i) => (highBits ? keys[i * words + 1] : 0);
    device.queue.writeBuffer(sort.keys, 0, keys);
    device.queue.writeBuffer(sort.values, 0, vals);
    encoder = device.createCommandEncoder();
    pass = encoder.beginComputePass();
    sort.encode(pass);
    pass.end();
    device.queue.submit([encoder.finish()]);
    const sk = new Uint32Array(sort.keys.read(0, m * 4 * words));
    const sv = new Uint32Array(sort.values.read(0, m * 4));
    const order = Array.from(vals).sort((a, b) => high(a) - high(b) || keys[a * words] - keys[b * words] || a - b);
    let bad = 0;
    for (let i = 0; i < m; i++) {
      const j = order[i];
      if (sv[i] !== j || sk[i * words] !== keys[j * words] || (highBits && sk[i * words + 1] !== keys[j * words + 1])) bad++;
    }
    out.sort['high' + highBits] = bad;
    sort.destroy();
  }

  // Loss fixture.
  const lf = fixture('loss.json');
  const [lw, lh] = lf.size;
  const rendered = new Float32Array(lw * lh * 4);
  const target = new Uint32Array(lw * lh);
  for (let p = 0; p < lw * lh; p++) {
    for (let c = 0; c < 3; c++) rendered[p * 4 + c] = lf.render[p * 3 + c];
    target[p] = (lf.target[p * 3] | (lf.target[p * 3 + 1] << 8) | (lf.target[p * 3 + 2] << 16) | (255 << 24)) >>> 0;
  }
  const loss = new d.TrainingLoss(device, lw, lh);
  const rb = device.upload(rendered);
  const tb = device.createTexture({ size: [lw, lh], format: 'rgba8unorm' });
  device.queue.writeTexture({ texture: tb }, target, { bytesPerRow: lw * 4 }, [lw, lh]);
  encoder = device.createCommandEncoder();
  loss.encode(encoder, rb, tb.createView());
  device.queue.submit([encoder.finish()]);
  const value = await loss.read();
  const lg = new Float32Array(loss.pixelGrads.read(0, lw * lh * 16));
  let gmax = 0;
  let gerr = 0;
  for (let p = 0; p < lw * lh; p++) {
    for (let c = 0; c < 3; c++) {
      gmax = Math.max(gmax, Math.abs(lf.gradient[p * 3 + c]));
      gerr = Math.max(gerr, Math.abs(lg[p * 4 + c] - lf.gradient[p * 3 + c]));
    }
  }
  out.loss = { error: Math.abs(value - lf.loss), gradientError: gerr / gmax };

  // Backward fixture through GaussianRenderer, with the afterPreprocess hook.
  const bf = fixture('backward.json');
  const count = bf.gaussians.length;
  const g = new Float32Array(count * 16);
  bf.gaussians.forEach((item, i) => {
    g.set(item.position, i * 16);
    g[i * 16 + 3] = item.opacity;
    g.set(item.scale, i * 16 + 4);
    g.set(item.rotation, i * 16 + 8);
    g.set(item.color, i * 16 + 12);
  });
  const rest = new Float32Array([].concat.apply([], bf.shRest));
  const renderer = new d.GaussianRenderer(device);
  renderer.setGaussians(g, { rest, degree: bf.shDegree });
  const [w, h] = bf.viewport;
  let hookSaw = -1;
  let hookPass = false;
  const frame = device.createTexture({ size: [w, h], format: 'rgba8unorm' });
  const camera = { view: bf.view, proj: bf.proj, viewport: bf.viewport };
  encoder = device.createCommandEncoder();
  renderer.encode(encoder, frame.createView(), camera, bf.background, (p) => {
    hookPass = p instanceof d.ComputePassEncoder;
    hookSaw = renderer.splatBuffer.size / (d.FLOATS_PER_SPLAT * 4);
  });
  device.queue.submit([encoder.finish()]);
  await renderer.afterSubmit();
  const frameBytes = await d.readTexture(device, frame);
  const color = new Float32Array(renderer.pixelBuffer.read(0, w * h * 16));
  let imageErr = 0;
  for (let p = 0; p < w * h; p++) {
    for (let c = 0; c < 3; c++) imageErr = Math.max(imageErr, Math.abs(color[p * 4 + c] - bf.image[p * 3 + c]));
  }
  const pg = new Float32Array(w * h * 4);
  for (let p = 0; p < w * h; p++) for (let c = 0; c < 3; c++) pg[p * 4 + c] = bf.pixelGrads[p * 3 + c];
  const pgb = device.upload(pg);
  encoder = device.createCommandEncoder();
  renderer.encodeBackward(encoder, pgb);
  device.queue.submit([encoder.finish()]);
  const grads = renderer.gradients;
  const gg = new Float32Array(grads.gaussians.read(0, count * 64));
  const shg = new Float32Array(grads.sh.read(0, count * 180));
  const groups = { position: [0, 3], scale: [4, 3], rotation: [8, 4], opacity: [3, 1], color: [12, 3] };
  const gradientErrors = {};
  for (const name of Object.keys(groups)) {
    const [off, size] = groups[name];
    let ref = 0;
    let err = 0;
    bf.gradients.forEach((item, i) => {
      const want = [].concat(item[name]);
      for (let k = 0; k < size; k++) {
        ref = Math.max(ref, Math.abs(want[k]));
        err = Math.max(err, Math.abs(gg[i * 16 + off + k] - want[k]));
      }
    });
    gradientErrors[name] = err / ref;
  }
  let shRef = 0;
  let shErr = 0;
  bf.gradients.forEach((item, i) => {
    for (let k = 0; k < 45; k++) {
      shRef = Math.max(shRef, Math.abs(item.sh[k]));
      shErr = Math.max(shErr, Math.abs(shg[i * 45 + k] - item.sh[k]));
    }
  });
  gradientErrors.sh = shErr / shRef;
  out.render = { count: renderer.count, hookSaw, hookPass, imageError: imageErr, gradientErrors,
    textureMatches: frameBytes.length === w * h * 4 && frameBytes[3] === 255 };

  // A standalone TileRasterizer over the renderer's splats (src/tileRasterizer.ts:59,127,200):
  // the same pair lists, so the same image; its splat gradients equal the renderer's
  // up to the order of float atomics.
  const cameraBuffer = device.createBuffer({ size: 144, usage: 0 });
  const splatGrads = device.createBuffer({ size: count * 36, usage: 0 });
  const tiles = new d.TileRasterizer(device, cameraBuffer, renderer.splatBuffer, renderer.count);
  let beforePass = false;
  encoder = device.createCommandEncoder();
  tiles.encode(encoder, null, [w, h], bf.background, (p) => {
    beforePass = p instanceof d.ComputePassEncoder;
  });
  tiles.encodeBackward(encoder, [w, h], pgb, splatGrads);
  device.queue.submit([encoder.finish()]);
  await tiles.afterSubmit();
  const own = new Float32Array(tiles.pixelColor.read(0, w * h * 16));
  const rendererColor = new Float32Array(renderer.pixelBuffer.read(0, w * h * 16));
  const sgA = new Float32Array(splatGrads.read(0, count * 36));
  const sgB = new Float32Array(grads.splats.read(0, count * 36));
  out.tiles = { beforePass, pairs: tiles.lastPairCount, rendererPairs: renderer.tiles.lastPairCount,
    imageDiff: worstDiff(own, rendererColor), gradientDiff: worstDiff(sgA, sgB) / Math.max(1e-30, maxAbs(sgB)) };
  tiles.destroy();

  // The argument lists of the pre-drop-in forms now fail loudly instead of binding
  // an encoder or camera to the wrong parameter.
  const refused = [];
  const expectTypeError = (name, f) => {
    try {
      f();
      refused.push(name + ': accepted');
    } catch (e) {
      refused.push(e instanceof TypeError ? name : name + ': ' + e.message);
    }
  };
  expectTypeError('renderer.encode(target, camera, bg)', () => renderer.encode(null, camera, bf.background));
  expectTypeError('renderer.encodeBackward(grads)', () => renderer.encodeBackward(pgb));
  expectTypeError('loss.encode(rendered, target)', () => loss.encode(rb, tb.createView()));
  expectTypeError('scan.encode()', () => new d.PrefixSum(device, pgb, 4).encode());
  expectTypeError('new TileRasterizer(device, splats, count)', () => new d.TileRasterizer(device, renderer.splatBuffer, count));
  out.refused = refused;
  loss.destroy();
  rb.destroy();
  tb.destroy();
  frame.destroy();
  cameraBuffer.destroy();
  splatGrads.destroy();
  pgb.destroy();
  renderer.destroy();

  // The MLP with only output biases: every sample's logits are those biases.
  const net = new d.DensifyMLP(device, 3);
  const biasOnly = new Float32Array(5252);
  biasOnly.set([0.5, -1.25, 2.0, 3.75], 5248);
  net.weights = biasOnly;
  const feats = device.upload(new Float32Array(12 * 5).map((_, k) => k * 0.1));
  const logits = device.createBuffer(5 * 16);
  net.forward(feats, 5, logits);
  out.mlp = { logits: Array.from(new Float32Array(logits.read())), weightsBack: net.weights[5251] };
  net.destroy();
  feats.destroy();
  logits.destroy();

  // A batch trainer over views of two sizes: one step.
  // The reference's TrainingView holds its photo as an rgba8unorm texture; one view
  // here passes a Uint32Array of rgba8 words instead.
  const views = [[64, 48], [48, 64], [64, 48]].map(([vw, vh], k) => {
    const a = (2 * Math.PI * k) / 3;
    const eye = [4 * Math.sin(a), 0, 4 * Math.cos(a)];
    const words = new Uint32Array(vw * vh).fill(0xff808080);
    let image = words;
    if (k > 0) {
      image = device.createTexture({ size: [vw, vh], format: 'rgba8unorm' });
      device.queue.writeTexture({ texture: image }, words, { bytesPerRow: vw * 4 }, [vw, vh]);
    }
    return { name: 'view' + k, camera: { view: lookAt(eye), proj: perspective(Math.PI / 3, vw / vh), viewport: [vw, vh] },
      image };
  });
  const raw = { params: new Float32Array(16 * 64), sh: new Float32Array(45 * 64) };
  for (let i = 0; i < 64; i++) {
    raw.params.set([random() - 0.5, random() - 0.5, random() - 0.5, 0, -3, -3, -3, 0, 1, 0, 0, 0, 0, 0, 0, 0], i * 16);
  }
  const trainer = new d.Trainer(device, raw, views, Object.assign({}, d.DEFAULT_SETTINGS, { extent: 1 }), { batch: 2 });
  await trainer.step();
  out.batch = { size: trainer.batchSize, loss: await trainer.lastLoss(), iteration: trainer.iteration };

  // The .ply round trip of src/checks/training.ts:122-136 on the trained Gaussians.
  const file = await trainer.writePly().arrayBuffer();
  const scene = d.parseGaussianPly(file, device);
  const back = await trainer.exportRaw();
  const expected = d.activate(back.params);
  const worst = (a, b) => a.reduce((most, value, k) => Math.max(most, Math.abs(value - b[k])), 0);
  out.ply = { bytes: file.byteLength, count: scene.count, degree: scene.shDegree,
    paramsError: worst(scene.gaussians, expected), shError: worst(scene.shRest, back.sh) };
  trainer.destroy();

  // Starting Gaussians from points: a line 0, 1, 3, 7 (src/loadDataset.ts:87-149).
  const line = new Float32Array([0, 0, 0, 1, 0, 0, 3, 0, 0, 7, 0, 0]);
  out.neighbors = Array.from(await d.meanNeighborDistance2(device, line));
  const start = await d.initialGaussians(device, line, new Uint8Array([255, 0, 128, 0, 0, 0, 0, 0, 0, 0, 0, 0]));
  out.initial = Array.from(start.params.subarray(0, 16));
  device.destroy();
  return out;
}

(async () => {
  const mode = process.argv[2] || 'cpu';
  const result = mode === 'gpu' ? await gpuChecks() : cpuChecks();
  process.stdout.write(JSON.stringify(result) + '\n');
})().catch((e) => {
  process.stderr.write(String(e && e.stack ? e.stack : e) + '\n');
  process.exit(1);
});
