'use strict';
// Runs the host-side helpers of home-measure-privacy_amd/js/formats.js on a job
// written by test_formats.py and prints the results as JSON.  Needs no GPU.
const fs = require('fs');
const path = require('path');
const F = require(path.join(__dirname, '..', 'home-measure-privacy_amd', 'js', 'formats.js'));

function sequence(values) {
  let k = 0;
  return () => {
    if (k >= values.length) throw new Error('random sequence exhausted');
    return values[k++];
  };
}

const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const gaussians = job.gaussians;
for (const k of job.poison || []) gaussians[k].mean[0] = NaN;
const out = {};
out.splat = F.toSplat(gaussians).toString('base64');
const clean = gaussians.filter((g, i) => (job.poison || []).indexOf(i) < 0);
out.ply = F.toPly(clean).toString('base64');
out.plyBack = F.fromPly(F.toPly(clean)).map((g) => Object.assign({}, g, g.sh ? { sh: Array.from(g.sh) } : {}));
const fixture = fs.readFileSync(job.fixture);
out.fixture = F.fromPly(fixture);
out.fixtureAgain = F.toPly(out.fixture).toString('base64');
out.densify = F.densify(clean, job.stats, job.densifyOptions, sequence(job.random));
out.sizes = Array.from(F.neighbourSizes(job.xyz, job.xyz.length / 3));
out.init = F.initFromPoints({ xyz: job.xyz, rgb: job.rgb, count: job.xyz.length / 3 });
out.initFixed = F.initFromPoints({ xyz: job.xyz, rgb: job.rgb, count: job.xyz.length / 3 }, 0.02);
out.shuffled = F.shuffled(job.shuffleN, sequence(job.shuffleRandom));
out.extent = F.sceneExtent(job.views.map((v) => ({ cam: { view: v } })));
process.stdout.write(JSON.stringify(out));
