'use strict';
// Runs createRenderer().trainScene on a job written by test_train_scene.py:
// start Gaussians (packed float32), views with RGB byte images, options and the
// random sequence that replaces Math.random.  Writes the trained Gaussians
// (packed float32) and the progress reports.
const fs = require('fs');
const path = require('path');
const gpu = require(path.join(__dirname, '..', 'home-measure-privacy_amd', 'js', 'gpu.js'));

async function main() {
  const dir = process.argv[2];
  const job = JSON.parse(fs.readFileSync(path.join(dir, 'job.json'), 'utf8'));
  const raw = fs.readFileSync(path.join(dir, 'start.f32'));
  const start = gpu.unpackGaussians(new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4));
  const views = job.cams.map((cam, k) => ({ cam, image: new Uint8Array(fs.readFileSync(path.join(dir, `view${k}.u8`))) }));
  let next = 0;
  const random = () => {
    if (next >= job.random.length) throw new Error('random sequence exhausted');
    return job.random[next++];
  };
  const progress = [];
  const r = await gpu.createRenderer();
  const t0 = Date.now();
  const out = await r.trainScene(start, views, Object.assign({}, job.options, {
    random,
    onProgress: (p) => { progress.push(p); },
  }));
  const seconds = (Date.now() - t0) / 1000;
  fs.writeFileSync(path.join(dir, 'trained.f32'), Buffer.from(gpu.packGaussians(out).buffer));
  fs.writeFileSync(path.join(dir, 'done.json'), JSON.stringify({ progress, count: out.length, seconds,
    randomUsed: next }));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
