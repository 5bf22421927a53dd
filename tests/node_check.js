'use strict';
// Drives the JavaScript shim the way a trainer does: reads a scene written by
// test_node_addon.py, calls createRenderer().gradients / render / lossGradientOf
// and tileLists, and writes the results as raw float32 / uint32 files.
const fs = require('fs');
const path = require('path');
const gpu = require(path.join(__dirname, '..', 'home-measure-privacy_amd', 'js', 'gpu.js'));

async function main() {
  const dir = process.argv[2];
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'));
  const raw = fs.readFileSync(path.join(dir, 'gs.f32'));
  const gs = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
  const tr = fs.readFileSync(path.join(dir, 'target.f32'));
  const target = new Float32Array(tr.buffer, tr.byteOffset, tr.byteLength / 4);
  const gaussians = gpu.unpackGaussians(gs);
  const r = await gpu.createRenderer();
  const out = await r.gradients(gaussians, meta.cam, target, meta.ssim);
  const image = await r.render(gaussians, meta.cam);
  const lg = await r.lossGradientOf(image, target, meta.cam.width, meta.cam.height, 0.2);
  fs.writeFileSync(path.join(dir, 'image.f32'), Buffer.from(out.image.buffer));
  fs.writeFileSync(path.join(dir, 'render.f32'), Buffer.from(image.buffer));
  fs.writeFileSync(path.join(dir, 'grads.f32'), Buffer.from(out.grads.buffer));
  fs.writeFileSync(path.join(dir, 'shape.f32'), Buffer.from(out.shape.buffer));
  fs.writeFileSync(path.join(dir, 'lossgrad.f32'), Buffer.from(lg.buffer));
  const sp = fs.readFileSync(path.join(dir, 'splats.f32'));
  const splats = new Float32Array(sp.buffer, sp.byteOffset, sp.byteLength / 4);
  const tl = gpu.tileLists(splats, splats.length / 12, meta.cam.width, meta.cam.height);
  fs.writeFileSync(path.join(dir, 'list.u32'), Buffer.from(tl.list.buffer));
  fs.writeFileSync(path.join(dir, 'ranges.u32'), Buffer.from(tl.ranges.buffer));
  let threw = false;
  try {
    await r.render(gaussians, { view: meta.cam.view, fx: 1, fy: 1, cx: 0, cy: 0, width: 0.5, height: 3 });
  } catch (e) {
    threw = /gsr error/.test(e.message) && typeof e.code === 'number';
  }
  fs.writeFileSync(path.join(dir, 'done.json'), JSON.stringify({ ok: true, threwOnBadCamera: threw }));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
