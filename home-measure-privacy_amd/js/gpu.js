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

const path = require('path');
const addon = require(path.join(__dirname, 'gsr.node'));

const GSR_E_CAPACITY = -4;

/** packGaussians (gpu.js:688-694): 16 floats per Gaussian. */
function packGaussians(gaussians) {
  const out = new Float32Array(gaussians.length * 16);
  for (let i = 0; i < gaussians.length; i++) {
    const g = gaussians[i];
    const o = i * 16;
    out[o] = g.mean[0]; out[o + 1] = g.mean[1]; out[o + 2] = g.mean[2];
    out[o + 4] = g.logScale[0]; out[o + 5] = g.logScale[1]; out[o + 6] = g.logScale[2];
    out[o + 7] = g.opacityLogit;
    for (let k = 0; k < 4; k++) out[o + 8 + k] = g.rotation[k];
    out[o + 12] = g.color[0]; out[o + 13] = g.color[1]; out[o + 14] = g.color[2];
  }
  return out;
}

/** unpackGaussians (gpu.js:602-609). */
function unpackGaussians(data) {
  const out = [];
  for (let i = 0; i < data.length / 16; i++) {
    const o = i * 16;
    out.push({
      mean: [data[o], data[o + 1], data[o + 2]],
      logScale: [data[o + 4], data[o + 5], data[o + 6]],
      opacityLogit: data[o + 7],
      rotation: [data[o + 8], data[o + 9], data[o + 10], data[o + 11]],
      color: [data[o + 12], data[o + 13], data[o + 14]],
    });
  }
  return out;
}

/** The 80-byte camera uniform (cameraUniform, gpu.js:696-701). */
function cameraUniform(cam, count) {
  const buf = new ArrayBuffer(80);
  const f = new Float32Array(buf);
  for (let k = 0; k < 12; k++) f[k] = cam.view[k];
  f[12] = cam.fx || 0; f[13] = cam.fy || 0; f[14] = cam.cx || 0; f[15] = cam.cy || 0;
  f[16] = cam.width; f[17] = cam.height;
  const u = new Uint32Array(buf);
  u[18] = count; u[19] = cam.width * cam.height;
  return buf;
}

/** sceneExtent (gpu.js:564-568): 1.1 x the largest camera distance from their mean. */
function sceneExtent(views) {
  const centres = views.map((v) => {
    const m = v.cam.view;
    return [0, 1, 2].map((j) => -(m[j] * m[3] + m[4 + j] * m[7] + m[8 + j] * m[11]));
  });
  const mean = [0, 1, 2].map((j) => centres.reduce((a, c) => a + c[j], 0) / centres.length);
  let far = 1e-3;
  for (const c of centres) far = Math.max(far, Math.hypot(c[0] - mean[0], c[1] - mean[1], c[2] - mean[2]));
  return 1.1 * far;
}

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

  // project -> tile lists -> blend (forward(), gpu.js:783-812).
  function forward(gBuf, shBuf, shDegree, n, cam) {
    const camBuf = cameraUniform(cam, n);
    const sBuf = buffer(n * 48);
    addon.project(ctx, camBuf, gBuf, shBuf, shDegree, n, sBuf);
    const tiles = Math.ceil(cam.width / 16) * Math.ceil(cam.height / 16);
    const rBuf = buffer(tiles * 8);
    let entries;
    try {
      entries = addon.binSort(ctx, camBuf, sBuf, n, null, 0, rBuf);
    } catch (e) {
      if (e.code !== GSR_E_CAPACITY) throw e;
      entries = e.entries;
    }
    const oBuf = buffer(entries * 4);
    if (entries > 0) addon.binSort(ctx, camBuf, sBuf, n, oBuf, entries, rBuf);
    const pixels = cam.width * cam.height;
    const iBuf = buffer(pixels * 12);
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
    const f = forward(u.gBuf, u.shBuf, u.shDegree, gaussians.length, cam);
    const out = new Float32Array(f.pixels * 3);
    addon.download(ctx, f.iBuf, out);
    release([u.gBuf, u.shBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf]);
    return out;
  }

  /** gradients (gpu.js:826-839): {image, grads (16 per Gaussian), shape (10 per Gaussian)}. */
  async function gradients(gaussians, cam, target, ssim = 0) {
    const n = gaussians.length;
    const u = upload(gaussians);
    const f = forward(u.gBuf, u.shBuf, u.shDegree, n, cam);
    const tBuf = buffer(f.pixels * 12);
    addon.upload(ctx, tBuf, Float32Array.from(target));
    const lg = buffer(f.pixels * 12);
    addon.lossGrad(ctx, f.camBuf, f.iBuf, tBuf, ssim, lg);
    const gradBuf = buffer(n * 64);
    addon.backward(ctx, f.camBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf);
    const shapeBuf = buffer(n * 40);
    const shGrad = u.shBuf ? buffer(n * 192) : null;
    addon.shape(ctx, f.camBuf, u.gBuf, u.shBuf, u.shDegree, n, gradBuf, shapeBuf, shGrad);
    const image = new Float32Array(f.pixels * 3);
    const grads = new Float32Array(n * 16);
    const shape = new Float32Array(n * 10);
    addon.download(ctx, f.iBuf, image);
    addon.download(ctx, gradBuf, grads);
    addon.download(ctx, shapeBuf, shape);
    const out = { image, grads, shape };
    if (shGrad) {
      out.shGrad = new Float32Array(n * 48);
      addon.download(ctx, shGrad, out.shGrad);
    }
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
    const g = new Float32Array(pixels * 3);
    addon.download(ctx, out, g);
    release([iBuf, tBuf, out]);
    return g;
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
      const f = forward(gBuf, null, 0, n, view.cam);
      const lg = buffer(f.pixels * 12);
      addon.lossGrad(ctx, f.camBuf, f.iBuf, targets[(t - 1) % views.length], 0, lg);
      const gradBuf = buffer(n * 64);
      addon.backward(ctx, f.camBuf, f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf);
      const shapeBuf = buffer(n * 40);
      addon.shape(ctx, f.camBuf, gBuf, null, 0, n, gradBuf, shapeBuf, null);
      new Float32Array(stepData).set([t, 0, lr.color, lr.opacity, lr.mean, lr.scale, lr.rotation, 0]);
      new Uint32Array(stepData)[1] = n;
      addon.adam(ctx, stepData, gBuf, gradBuf, shapeBuf, moments);
      const image = new Float32Array(f.pixels * 3);
      addon.download(ctx, f.iBuf, image);
      let loss = 0;
      for (let i = 0; i < image.length; i++) loss += Math.abs(image[i] - view.target[i]);
      losses.push(loss / image.length);
      release([f.sBuf, f.oBuf, f.rBuf, f.iBuf, lg, gradBuf, shapeBuf]);
    }
    const data = new Float32Array(n * 16);
    addon.download(ctx, gBuf, data);
    release([gBuf, moments].concat(targets));
    return { gaussians: unpackGaussians(data), losses };
  }

  return { device: ctx, render, gradients, train, lossGradientOf };
}

module.exports = { createRenderer, packGaussians, unpackGaussians, tileLists, sceneExtent, cameraUniform };
