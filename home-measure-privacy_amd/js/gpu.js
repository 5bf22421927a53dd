'use strict';
// Drop-in for the renderer of trainer/gpu.js on an AMD MI355X: the same named
// exports and the same renderer object (createRenderer, gpu.js:704-980), backed by
// libgsr through the N-API addon gsr.node instead of WebGPU.  Node 12 compatible
// (CommonJS; no optional chaining, no nullish coalescing, no structuredClone).
//
// Layouts are the reference's: Gaussians {mean[3], logScale[3], rotation[4] wxyz,
// opacityLogit, color[3]} packed as 16 floats (gpu.js:688-694); cameras {view[12],
// fx, fy, cx, cy, width, height}, sent as the 80-byte uniform (gpu.js:696-701).
// Optional SH: gaussians[i].sh = 48 floats ([16][3], coefficient 0 = colour).
// The device-free helpers (packing, .ply / .splat, densify, ...) live in
// formats.js and are re-exported here.
//
// As in the reference, the renderer's functions are async: device work is
// queued in order, and readbacks (read(), gpu.js:716-726) and the tile-list
// entry count wait for the device on a worker thread of the addon, so the event
// loop keeps running meanwhile.  Extension: gradientsBatch() renders many views
// in one batched step, and trainScene({batchViews}) trains on several views per
// step that way.

const path = require('path');
const addon = require(path.join(__dirname, 'gsr.node'));

const F = require(path.join(__dirname, 'formats.js'));
const { packGaussians, unpackGaussians, cameraUniform, sceneExtent, shuffled, densify } = F;

const GSR_E_CAPACITY = -4;

let sharedCtx = null;
function defaultCtx() {
  if (!sharedCtx) sharedCtx = addon.create(0);
  return sharedCtx;
}

/** tileLists (gpu.js:450-481) computed on the GPU: {list, ranges}, same order. */
function tileLists(splats, n, width, height) {
  const ctx = defaultCtx();
  const cam = cameraUniform({ view: new Array(12).fill(0), width, height }, n);
  const tiles = Math.ceil(width / 16) * Math.ceil(height / 16);
  const sBuf = addon.alloc(ctx, n * 48);
  const rBuf = addon.alloc(ctx, tiles * 8);
  try {
    addon.upload(ctx, sBuf, splats.subarray(0, n * 12));
    let entries = 0;
    try {
      entries = addon.binSort(ctx, cam, sBuf, n, null, 0, rBuf);
    } catch (e) {
      if (e.code !== GSR_E_CAPACITY) throw e;
      entries = e.entries;
    }
    const list = new Uint32Array(Math.max(1, entries));
    const ranges = new Uint32Array(tiles * 2);
    if (entries > 0) {
      const oBuf = addon.alloc(ctx, entries * 4);
      try {
        addon.binSort(ctx, cam, sBuf, n, oBuf, entries, rBuf);
        addon.download(ctx, oBuf, list);
      } finally {
        addon.free(ctx, oBuf);
      }
    }
    addon.download(ctx, rBuf, ranges);
    return { list, ranges };
  } finally {
    addon.free(ctx, sBuf);
    addon.free(ctx, rBuf);
  }
}

/** The 80-byte camera uniforms of several views, concatenated (batched calls). */
function camerasUniform(cams, count) {
  const out = new Uint8Array(cams.length * 80);
  cams.forEach((c, v) => out.set(new Uint8Array(cameraUniform(c, count)), v * 80));
  return out;
}

/** The views' RGB images (any numeric arrays) divided by `divisor`, concatenated as float32. */
function concatTargets(images, pixels, divisor) {
  const total = pixels.reduce((a, b) => a + b, 0);
  const out = new Float32Array(total * 3);
  let off = 0;
  images.forEach((img, v) => {
    const m = pixels[v] * 3;
    for (let i = 0; i < m; i++) out[off + i] = img[i] / divisor;
    off += m;
  });
  return out;
}

function shOf(gaussians) {
  if (!gaussians.length || !gaussians[0].sh) return null;
  const out = new Float32Array(gaussians.length * 48);
  gaussians.forEach((g, i) => out.set(g.sh, i * 48));
  return out;
}

/**
 * createRenderer (gpu.js:704): the renderer object, on device `device`.
 * `gpuTiles` is accepted for compatibility; tile lists always follow tileLists.
 */
async function createRenderer({ gpuTiles = true, device = 0 } = {}) { // eslint-disable-line no-unused-vars
  const ctx = addon.create(device);
  const buffer = (bytes) => addon.alloc(ctx, Math.max(16, bytes));
  const release = (bufs) => { for (const b of bufs) if (b !== null && b !== undefined) addon.free(ctx, b); };

  // Buffers reused between training steps, grown by a quarter when too small
  // (scratch(), gpu.js:764-772).
  const pool = new Map();
  function scratch(name, bytes) {
    const have = pool.get(name);
    if (have && have.bytes >= bytes) return have.ptr;
    if (have) addon.free(ctx, have.ptr);
    const size = Math.max(16, Math.ceil(bytes * 1.25 / 4) * 4);
    const ptr = addon.alloc(ctx, size);
    pool.set(name, { ptr, bytes: size });
    return ptr;
  }
  function pooled(name) { return pool.get(name).bytes; }
  // read() of gpu.js:716-726: a copy of device memory once the queued work is done.
  async function readF32(ptr, count) {
    return new Float32Array(await addon.read(ctx, ptr, count * 4));
  }

  // project -> tile lists -> blend (forward(), gpu.js:783-812).  With `reuse`
  // the buffers come from the pool (training) and stay there.
  async function forward(gBuf, shBuf, shDegree, n, cam, reuse = false) {
    const get = reuse ? scratch : (name, bytes) => buffer(bytes);
    const camBuf = cameraUniform(cam, n);
    const sBuf = get('splats', n * 48);
    addon.project(ctx, camBuf, gBuf, shBuf, shDegree, n, sBuf);
    const tiles = Math.ceil(cam.width / 16) * Math.ceil(cam.height / 16);
    const rBuf = get('ranges', tiles * 8);
    let oBuf = reuse ? scratch('order', 4) : null;
    let entries;
    try {
      entries = await addon.binSortAsync(ctx, camBuf, sBuf, n, oBuf, reuse ? pooled('order') / 4 : 0, rBuf);
    } catch (e) {
      if (e.code !== GSR_E_CAPACITY) throw e;
      entries = e.entries;
      oBuf = get('order', entries * 4);
      if (entries > 0) await addon.binSortAsync(ctx, camBuf, sBuf, n, oBuf, entries, rBuf);
    }
    if (oBuf === null) oBuf = buffer(4);
    const pixels = cam.width * cam.height;
    const iBuf = get('image', pixels * 12);
    addon.forward(ctx, camBuf, sBuf, oBuf, rBuf, iBuf, null);
    return { camBuf, sBuf, oBuf, rBuf, iBuf, pixels, entries };
  }

  function upload(gaussians) {
    const gBuf = buffer(gaussians.length * 64);
    addon.upload(ctx, gBuf, packGaussians(gaussians));
    const sh = shOf(gaussians);
    let shBuf = null;
    if (sh) {
      shBuf = buffer(sh.byteLength);
      addon.upload(ctx, shBuf, sh);
    }
    return { gBuf, shBuf, shDegree: sh ? 3 : 0 };
  }

  /** render (gpu.js:815-820): RGB floats, width * height * 3. */
  async function render(gaussians, cam) {
    const u = upload(gaussians);
    const f = await forward(u.gBuf, u.shBuf, u.shDegree, gaussians.length, cam);
    const out = await readF32(f.iBuf, f.pixels * 3);
    release([u.gBuf, u.shBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf]);
    return out;
  }

  /** gradients (gpu.js:826-839): {image, grads (16 per Gaussian), shape (10 per Gaussian)}. */
  async function gradients(gaussians, cam, target, ssim = 0) {
    const n = gaussians.length;
    const u = upload(gaussians);
    const f = await forward(u.gBuf, u.shBuf, u.shDegree, n, cam);
    const tBuf = buffer(f.pixels * 12);
    addon.upload(ctx, tBuf, Float32Array.from(target));
    const lg = buffer(f.pixels * 12);
    addon.lossGrad(ctx, f.camBuf, f.iBuf, tBuf, ssim, lg);
    const gradBuf = buffer(n * 64);
    addon.backward(ctx, f.camBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf);
    const shapeBuf = buffer(n * 40);
    const shGrad = u.shBuf ? buffer(n * 192) : null;
    addon.shape(ctx, f.camBuf, u.gBuf, u.shBuf, u.shDegree, n, gradBuf, shapeBuf, shGrad);
    const image = await readF32(f.iBuf, f.pixels * 3);
    const grads = await readF32(gradBuf, n * 16);
    const shape = await readF32(shapeBuf, n * 10);
    const out = { image, grads, shape };
    if (shGrad) out.shGrad = await readF32(shGrad, n * 48);
    release([u.gBuf, u.shBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, tBuf, lg, gradBuf, shapeBuf, shGrad]);
    return out;
  }

  /** lossGradientOf (gpu.js:962-977): dLoss/dPixel of an image against its photo. */
  async function lossGradientOf(image, target, width, height, lambda) {
    const pixels = width * height;
    const cam = cameraUniform({ view: new Array(12).fill(0), width, height }, 0);
    const iBuf = buffer(pixels * 12);
    const tBuf = buffer(pixels * 12);
    const out = buffer(pixels * 12);
    addon.upload(ctx, iBuf, Float32Array.from(image));
    addon.upload(ctx, tBuf, Float32Array.from(target));
    addon.lossGrad(ctx, cam, iBuf, tBuf, lambda, out);
    const g = await readF32(out, pixels * 3);
    release([iBuf, tBuf, out]);
    return g;
  }

  /**
   * gradientsBatch (extension): gradients() of many views of any resolution in
   * one batched step against one shared set of Gaussians.  targets[v] are the
   * views' RGB floats.  Returns {images[v], grads (16 per Gaussian, colour and
   * opacity summed over views), shape (10 per Gaussian, summed), shGrad (48 per
   * Gaussian, with SH), stats (2 per Gaussian: STATS, gpu.js:483-497, added for
   * every view), entries[v]}.
   */
  async function gradientsBatch(gaussians, cams, targets, ssim = 0) {
    const n = gaussians.length;
    const u = upload(gaussians);
    const camsBuf = camerasUniform(cams, n);
    const pixels = cams.map((c) => c.width * c.height);
    const total = pixels.reduce((a, b) => a + b, 0);
    const bt = addon.batchCreate(ctx, camsBuf, n, u.shDegree);
    const tBuf = buffer(total * 12), iBuf = buffer(total * 12);
    const gradBuf = buffer(n * 64), shapeBuf = buffer(n * 40), statsBuf = buffer(n * 8);
    const shGrad = u.shBuf ? buffer(n * 192) : null;
    try {
      addon.upload(ctx, tBuf, concatTargets(targets, pixels, 1));
      addon.batchStep(bt, u.gBuf, u.shBuf, tBuf, ssim, gradBuf, shapeBuf, shGrad, statsBuf, iBuf, false);
      const all = await readF32(iBuf, total * 3);
      const images = [];
      let off = 0;
      for (const p of pixels) {
        images.push(all.slice(off, off + p * 3));
        off += p * 3;
      }
      const out = { images, grads: await readF32(gradBuf, n * 16), shape: await readF32(shapeBuf, n * 10),
                    stats: await readF32(statsBuf, n * 2), entries: Array.from(addon.batchEntries(bt)) };
      if (shGrad) out.shGrad = await readF32(shGrad, n * 48);
      return out;
    } finally {
      addon.batchDestroy(bt);
      release([u.gBuf, u.shBuf, tBuf, iBuf, gradBuf, shapeBuf, statsBuf, shGrad]);
    }
  }

  /** train (gpu.js:846-878): fixed view cycle, L1, Adam on the 14 numbers. */
  async function train(gaussians, views, steps, lr = { color: 0.05, opacity: 0.05, mean: 0.002, scale: 0.01, rotation: 0.005 }) {
    const n = gaussians.length;
    const gBuf = buffer(n * 64);
    addon.upload(ctx, gBuf, packGaussians(gaussians));
    const moments = buffer(n * 28 * 4);
    const targets = views.map((v) => {
      const b = buffer(v.target.length * 4);
      addon.upload(ctx, b, Float32Array.from(v.target));
      return b;
    });
    const losses = [];
    const stepData = new ArrayBuffer(32);
    for (let t = 1; t <= steps; t++) {
      const view = views[(t - 1) % views.length];
      const f = await forward(gBuf, null, 0, n, view.cam);
      const lg = buffer(f.pixels * 12);
      addon.lossGrad(ctx, f.camBuf, f.iBuf, targets[(t - 1) % views.length], 0, lg);
      const gradBuf = buffer(n * 64);
      addon.backward(ctx, f.camBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf);
      const shapeBuf = buffer(n * 40);
      addon.shape(ctx, f.camBuf, gBuf, null, 0, n, gradBuf, shapeBuf, null);
      new Float32Array(stepData).set([t, 0, lr.color, lr.opacity, lr.mean, lr.scale, lr.rotation, 0]);
      new Uint32Array(stepData)[1] = n;
      addon.adam(ctx, stepData, gBuf, gradBuf, shapeBuf, moments);
      const image = await readF32(f.iBuf, f.pixels * 3);
      let loss = 0;
      for (let i = 0; i < image.length; i++) loss += Math.abs(image[i] - view.target[i]);
      losses.push(loss / image.length);
      release([f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf, shapeBuf]);
    }
    const data = await readF32(gBuf, n * 16);
    release([gBuf, moments].concat(targets));
    return { gaussians: unpackGaussians(data), losses };
  }

  /**
   * trainScene (gpu.js:890-959): one view per step, every view once per round in
   * a new order; densify every `densifyEvery` steps in [densifyFrom,
   * densifyUntil]; `onProgress({step, loss, count})` every 50 steps (awaited);
   * `shouldStop()` ends early.  Views are {cam, image} with image the photo's RGB
   * bytes.  `random` (extension) replaces Math.random for the shuffles and splits.
   * `batchViews` (extension, default 1): views per step, taken in the same
   * shuffled order and rendered in one batched step; Adam then steps on their
   * summed gradients and STATS counts every view.
   */
  async function trainScene(start, views, { steps = 3000, densifyEvery = 100, densifyFrom = 300, densifyUntil = 2000,
                                            maxSplats = 400000, ssim = 0, prune = false, opacityReset = false,
                                            onProgress, shouldStop, lr, random = Math.random, batchViews = 1 } = {}) {
    lr = Object.assign({ color: 0.025, opacity: 0.05, mean: 0.002, scale: 0.005, rotation: 0.001 }, lr);
    let gaussians = start;
    let n = gaussians.length;
    const extent = sceneExtent(views);
    const resetAt = opacityReset ? Math.round((densifyFrom + densifyUntil) / 2 / densifyEvery) * densifyEvery : -1;
    const lowOpacity = Math.log(0.01 / 0.99);
    let gBuf = null, moments = null, statsBuf = null;
    const upload = () => {
      release([gBuf, moments, statsBuf]);
      n = gaussians.length;
      gBuf = buffer(n * 64);
      addon.upload(ctx, gBuf, packGaussians(gaussians));
      moments = buffer(n * 28 * 4);
      statsBuf = buffer(n * 8);
    };
    upload();
    const stepData = new ArrayBuffer(32);
    let lossSum = 0, lossCount = 0;
    let order = [];
    const k = Math.max(1, Math.min(views.length, batchViews | 0));
    // Batched mode: one batch per sequence of view resolutions met, cameras swapped per step.
    const batches = new Map();
    const dropBatches = () => {
      for (const b of batches.values()) addon.batchDestroy(b);
      batches.clear();
    };
    // Mean absolute error of rendered RGB floats against photos' bytes.
    const l1 = (image, photos) => {
      let loss = 0, off = 0;
      for (const photo of photos) {
        for (let i = 0; i < photo.length; i++) loss += Math.abs(image[off + i] - photo[i] / 255);
        off += photo.length;
      }
      return loss / image.length;
    };
    try {
      for (let t = 1; t <= steps; t++) {
        if (shouldStop && shouldStop()) break;
        const chosen = [];
        while (chosen.length < k) {
          if (!order.length) order = shuffled(views.length, random);
          chosen.push(views[order.pop()]);
        }
        new Float32Array(stepData).set([t, 0, lr.color, lr.opacity, lr.mean, lr.scale, lr.rotation, 0]);
        new Uint32Array(stepData)[1] = n;
        if (k === 1) {
          const view = chosen[0];
          const f = await forward(gBuf, null, 0, n, view.cam, true);
          const tBuf = scratch('target', f.pixels * 12);
          addon.upload(ctx, tBuf, Float32Array.from(view.image, (v) => v / 255));
          const gradBuf = scratch('grads', n * 64);
          addon.clear(ctx, gradBuf, n * 64);
          const lg = scratch('lossGrad', f.pixels * 12);
          addon.lossGrad(ctx, f.camBuf, f.iBuf, tBuf, ssim, lg);
          addon.backward(ctx, f.camBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf);
          const shapeBuf = scratch('shape', n * 40);
          addon.shape(ctx, f.camBuf, gBuf, null, 0, n, gradBuf, shapeBuf, null);
          addon.stats(ctx, stepData, gradBuf, f.sBuf, statsBuf);
          addon.adam(ctx, stepData, gBuf, gradBuf, shapeBuf, moments);
          if (t % 25 === 0) {
            lossSum += l1(await readF32(f.iBuf, f.pixels * 3), [view.image]);
            lossCount++;
          }
        } else {
          const camsBuf = camerasUniform(chosen.map((v) => v.cam), n);
          const key = chosen.map((v) => v.cam.width + 'x' + v.cam.height).join(',');
          let bt = batches.get(key);
          if (!bt) {
            bt = addon.batchCreate(ctx, camsBuf, n, 0);
            batches.set(key, bt);
          } else {
            addon.batchSetCameras(bt, camsBuf);
          }
          const pixels = chosen.map((v) => v.cam.width * v.cam.height);
          const total = pixels.reduce((a, b) => a + b, 0);
          const tBuf = scratch('targets', total * 12);
          addon.upload(ctx, tBuf, concatTargets(chosen.map((v) => v.image), pixels, 255));
          const gradBuf = scratch('grads', n * 64), shapeBuf = scratch('shape', n * 40);
          const iBuf = scratch('images', total * 12);
          addon.batchStep(bt, gBuf, null, tBuf, ssim, gradBuf, shapeBuf, null, statsBuf, iBuf, true);
          addon.adam(ctx, stepData, gBuf, gradBuf, shapeBuf, moments);
          if (t % 25 === 0) {
            lossSum += l1(await readF32(iBuf, total * 3), chosen.map((v) => v.image));
            lossCount++;
          }
        }
        if (t >= densifyFrom && t <= densifyUntil && t % densifyEvery === 0) {
          // per-pixel gradients: the 3DGS threshold (NDC units) scaled by half the width
          const gradThreshold = 2e-4 / (0.5 * views[0].cam.width);
          const data = await readF32(gBuf, n * 16);
          const stats = await readF32(statsBuf, n * 2);
          gaussians = densify(unpackGaussians(data), stats, { gradThreshold, maxCount: maxSplats,
            sizeThreshold: 0.01 * extent, maxSize: prune ? 0.1 * extent : Infinity }, random);
          if (t === resetAt) for (const g of gaussians) g.opacityLogit = Math.min(g.opacityLogit, lowOpacity);
          upload();
          dropBatches();
        }
        if (t % 50 === 0) {
          if (onProgress) await onProgress({ step: t, loss: lossCount ? lossSum / lossCount : NaN, count: n });
          lossSum = 0;
          lossCount = 0;
          await new Promise((r) => setTimeout(r));
        }
      }
      return unpackGaussians(await readF32(gBuf, n * 16));
    } finally {
      dropBatches();
      release([gBuf, moments, statsBuf]);
      for (const entry of pool.values()) addon.free(ctx, entry.ptr);
      pool.clear();
    }
  }

  return { device: ctx, render, gradients, train, trainScene, lossGradientOf, gradientsBatch };
}

module.exports = {
  createRenderer, tileLists, packGaussians, unpackGaussians, cameraUniform, sceneExtent, shuffled, densify,
  initFromPoints: F.initFromPoints, neighbourSizes: F.neighbourSizes, toPly: F.toPly, fromPly: F.fromPly,
  toSplat: F.toSplat,
};
