'use strict';
// Drives the batched extension of the JavaScript shim: gradientsBatch() on a
// scene written by test_node_batch.py, and the addon's busy guard while a read()
// is pending.  Writes the results as raw float32 files.
const fs = require('fs');
const path = require('path');
const gpu = require(path.join(__dirname, '..', 'home-measure-privacy_amd', 'js', 'gpu.js'));
const addon = require(path.join(__dirname, '..', 'home-measure-privacy_amd', 'js', 'gsr.node'));

function f32(file) {
  const b = fs.readFileSync(file);
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
}

async function main() {
  const dir = process.argv[2];
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'));
  const gaussians = gpu.unpackGaussians(f32(path.join(dir, 'gs.f32')));
  const sh = f32(path.join(dir, 'sh.f32'));
  gaussians.forEach((g, i) => { g.sh = sh.slice(i * 48, i * 48 + 48); });
  const targets = meta.cams.map((c, v) => f32(path.join(dir, `target${v}.f32`)));
  const r = await gpu.createRenderer();
  const out = await r.gradientsBatch(gaussians, meta.cams, targets, meta.ssim);
  const write = (name, arr) => fs.writeFileSync(path.join(dir, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
  out.images.forEach((img, v) => write(`image${v}.f32`, img));
  write('grads.f32', out.grads);
  write('shape.f32', out.shape);
  write('shgrad.f32', out.shGrad);
  write('stats.f32', out.stats);
  // While a read() is pending its context refuses other calls.
  const ctx = addon.create(0);
  const buf = addon.alloc(ctx, 1 << 20);
  const pending = addon.read(ctx, buf, 1 << 20);
  let busy = false;
  try {
    addon.alloc(ctx, 16);
  } catch (e) {
    busy = /busy/.test(e.message);
  }
  const data = await pending;
  addon.free(ctx, buf);  // accepted again once the read has settled
  // The batch knobs read back as eight values, the depth slices last.
  const camsBuf = new Uint8Array(meta.cams.length * 80);
  meta.cams.forEach((c, v) => camsBuf.set(new Uint8Array(gpu.cameraUniform(c, 100000)), v * 80));
  const bt = addon.batchCreate(ctx, camsBuf, 100000, 0);
  const tuningDefault = Array.from(addon.batchTuning(bt));
  addon.batchSetTuning(bt, new Uint32Array([1, 2048, 0, 0, 1024, 256, 3, 4]));
  const tuningSet = Array.from(addon.batchTuning(bt));
  addon.batchDestroy(bt);
  fs.writeFileSync(path.join(dir, 'done.json'), JSON.stringify({ ok: true, busy, readBytes: data.byteLength,
    entries: out.entries, tuningDefault, tuningSet }));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
