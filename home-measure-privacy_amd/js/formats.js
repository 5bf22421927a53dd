'use strict';
// Host-side helpers of trainer/gpu.js that need no device: record packing, the
// camera uniform, scene extent, view shuffling, starting sizes, densification and
// the .ply / .splat formats.  Pure JavaScript (Node 12: CommonJS, no optional
// chaining, no nullish coalescing, no structuredClone), so it loads without the
// addon.  gpu.js re-exports everything here.
//
// Gaussians are the reference's objects {mean[3], logScale[3], rotation[4] wxyz,
// opacityLogit, color[3]} (gpu.js:512-518).  Extension: an optional `sh` of 48
// floats ([16][3], coefficient 0 in colour units) travels through .ply files as
// f_dc / f_rest_*.

const SH0 = 0.28209479177387814;  // Y_0^0, the DC basis value (gpu.js:499, 991)

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

/** shuffled (neighbours.js:56-63): 0..n-1 in random order, Fisher-Yates from the top. */
function shuffled(n, random = Math.random) {
  const order = [];
  for (let i = 0; i < n; i++) order.push(i);
  for (let i = n - 1; i > 0; i--) {
    const j = Math.floor(random() * (i + 1));
    const t = order[i]; order[i] = order[j]; order[j] = t;
  }
  return order;
}

/**
 * neighbourSizes (neighbours.js:13-53): per point, the root mean square distance
 * to its (up to) 3 nearest other points among those in the 27 cubes of edge 0.1
 * around its own cube, clamped to [min, max]; `fallback` when there are none.
 * Cubes are keyed exactly (the reference packs 16 bits an axis, which only
 * differs beyond 3 km from the origin).
 */
function neighbourSizes(xyz, count, { min = 0.005, max = 0.1, fallback = 0.06 } = {}) {
  const CELL = 0.1;
  const cellOf = (p, a) => Math.floor(xyz[p * 3 + a] / CELL);
  const cubes = new Map();
  const keyOf = (i, j, k) => i + ',' + j + ',' + k;
  for (let p = 0; p < count; p++) {
    const key = keyOf(cellOf(p, 0), cellOf(p, 1), cellOf(p, 2));
    const members = cubes.get(key);
    if (members) members.push(p);
    else cubes.set(key, [p]);
  }
  const sizes = new Float32Array(count);
  for (let p = 0; p < count; p++) {
    const ci = cellOf(p, 0), cj = cellOf(p, 1), ck = cellOf(p, 2);
    const x = xyz[p * 3], y = xyz[p * 3 + 1], z = xyz[p * 3 + 2];
    let d0 = Infinity, d1 = Infinity, d2 = Infinity;  // ascending
    for (let di = -1; di <= 1; di++) {
      for (let dj = -1; dj <= 1; dj++) {
        for (let dk = -1; dk <= 1; dk++) {
          const members = cubes.get(keyOf(ci + di, cj + dj, ck + dk));
          if (!members) continue;
          for (const q of members) {
            if (q === p) continue;
            const dx = xyz[q * 3] - x, dy = xyz[q * 3 + 1] - y, dz = xyz[q * 3 + 2] - z;
            const d = dx * dx + dy * dy + dz * dz;
            if (d >= d2) continue;
            if (d < d0) { d2 = d1; d1 = d0; d0 = d; } else if (d < d1) { d2 = d1; d1 = d; } else d2 = d;
          }
        }
      }
    }
    let sum = 0, k = 0;
    for (const d of [d0, d1, d2]) if (Number.isFinite(d)) { sum += d; k++; }
    const size = k ? Math.sqrt(sum / k) : fallback;
    sizes[p] = Math.min(max, Math.max(min, size));
  }
  return sizes;
}

/** initFromPoints (gpu.js:507-521): one faint isotropic Gaussian per point. */
function initFromPoints({ xyz, rgb, count }, size) {
  let sizes;
  if (size) { sizes = new Float32Array(count); sizes.fill(size); } else sizes = neighbourSizes(xyz, count);
  const opacityLogit = Math.log(0.1 / (1 - 0.1));
  const out = [];
  for (let i = 0; i < count; i++) {
    const s = Math.log(sizes[i]);
    out.push({
      mean: [xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]],
      logScale: [s, s, s],
      rotation: [1, 0, 0, 0],
      opacityLogit,
      color: [rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]],
    });
  }
  return out;
}

function cloneGaussian(g) {
  const c = {};
  for (const k of Object.keys(g)) {
    const v = g[k];
    c[k] = Array.isArray(v) ? v.slice() : (ArrayBuffer.isView(v) ? v.slice() : v);
  }
  return c;
}

/**
 * densify (gpu.js:575-599): drops Gaussians with opacity below minOpacity or a
 * world size above maxSize; Gaussians whose mean screen gradient (stats) reaches
 * gradThreshold are cloned when at most sizeThreshold large, else split into two
 * children scattered by their scale and shrunk by 1.6.  At most maxCount
 * Gaussians result.  `random` (extension) replaces Math.random for the scatter.
 */
function densify(gaussians, stats, { gradThreshold = 2e-4, sizeThreshold = 0.1, minOpacity = 0.005, maxCount = 400000,
                                     maxSize = Infinity } = {}, random = Math.random) {
  const out = [];
  let room = maxCount - gaussians.length;
  const shrink = Math.log(1.6);
  for (let i = 0; i < gaussians.length; i++) {
    const g = gaussians[i];
    if (1 / (1 + Math.exp(-g.opacityLogit)) < minOpacity) continue;
    const size = Math.max(Math.exp(g.logScale[0]), Math.exp(g.logScale[1]), Math.exp(g.logScale[2]));
    if (size > maxSize) continue;
    const avg = stats[i * 2 + 1] ? stats[i * 2] / stats[i * 2 + 1] : 0;
    if (avg < gradThreshold || room <= 0) { out.push(g); continue; }
    room--;
    if (size <= sizeThreshold) {
      out.push(g, cloneGaussian(g));
      continue;
    }
    for (let k = 0; k < 2; k++) {
      const c = cloneGaussian(g);
      c.mean = [0, 1, 2].map((j) => g.mean[j] + (random() - 0.5) * Math.exp(g.logScale[j]));
      c.logScale = g.logScale.map((l) => l - shrink);
      out.push(c);
    }
  }
  return out;
}

// ---- .ply (gpu.js:524-533 writes, gpu.js:983-999 reads) -------------------------

function restCount(gaussians) {
  return gaussians.length && gaussians[0].sh ? 45 : 0;
}

/**
 * toPly (gpu.js:524-533): binary little-endian 3DGS .ply, float properties
 * x y z nx ny nz f_dc_0-2 opacity scale_0-2 rot_0-3.  Returns a Buffer (the
 * reference returns a Blob of the same bytes).  Extension: Gaussians with `sh`
 * also write f_rest_0-44 after f_dc (channel-major, as 3DGS stores them) and take
 * f_dc from sh[0..2].
 */
function toPly(gaussians) {
  const rest = restCount(gaussians);
  const props = ['x', 'y', 'z', 'nx', 'ny', 'nz', 'f_dc_0', 'f_dc_1', 'f_dc_2'];
  for (let k = 0; k < rest; k++) props.push('f_rest_' + k);
  props.push('opacity', 'scale_0', 'scale_1', 'scale_2', 'rot_0', 'rot_1', 'rot_2', 'rot_3');
  const lines = ['ply', 'format binary_little_endian 1.0', 'element vertex ' + gaussians.length];
  for (const p of props) lines.push('property float ' + p);
  lines.push('end_header', '');
  const header = Buffer.from(lines.join('\n'), 'latin1');
  const P = props.length;
  const body = new Float32Array(gaussians.length * P);
  for (let i = 0; i < gaussians.length; i++) {
    const g = gaussians[i];
    const o = i * P;
    body[o] = g.mean[0]; body[o + 1] = g.mean[1]; body[o + 2] = g.mean[2];
    const dc = rest ? [g.sh[0], g.sh[1], g.sh[2]] : g.color;
    for (let c = 0; c < 3; c++) body[o + 6 + c] = (dc[c] - 0.5) / SH0;
    for (let c = 0; c < 3 && rest; c++)
      for (let k = 1; k < 16; k++) body[o + 9 + c * 15 + k - 1] = g.sh[k * 3 + c];
    const t = o + 9 + rest;
    body[t] = g.opacityLogit;
    body[t + 1] = g.logScale[0]; body[t + 2] = g.logScale[1]; body[t + 3] = g.logScale[2];
    for (let k = 0; k < 4; k++) body[t + 4 + k] = g.rotation[k];
  }
  return Buffer.concat([header, Buffer.from(body.buffer)]);
}

/**
 * fromPly (gpu.js:983-999): Gaussians from a binary 3DGS .ply whose header fits
 * the first 4096 bytes and whose properties are floats.  colour = clamp(f_dc *
 * SH0 + 0.5, 0, 1).  Extension: with f_rest_* present, `sh` holds the 16 x 3
 * coefficients (sh[0..2] = f_dc * SH0 + 0.5 unclamped, missing bands zero).  A
 * missing required property throws (the reference reads past the record).
 */
function fromPly(buffer) {
  const bytes = Buffer.isBuffer(buffer) ? buffer
    : (ArrayBuffer.isView(buffer) ? Buffer.from(buffer.buffer, buffer.byteOffset, buffer.byteLength)
      : Buffer.from(buffer));
  const head = bytes.slice(0, 4096).toString('latin1');
  const marker = 'end_header\n';
  const at = head.indexOf(marker);
  if (at < 0) throw new Error('fromPly: no end_header in the first 4096 bytes');
  const headerEnd = at + marker.length;
  const header = head.slice(0, headerEnd);
  const m = /element vertex (\d+)/.exec(header);
  if (!m) throw new Error('fromPly: no vertex element');
  const count = Number(m[1]);
  const props = [];
  const re = /property float (\w+)/g;
  let pm;
  while ((pm = re.exec(header)) !== null) props.push(pm[1]);
  const P = props.length;
  const need = count * P * 4;
  if (bytes.length < headerEnd + need) throw new Error('fromPly: file shorter than its header says');
  const data = new Float32Array(need / 4);
  new Uint8Array(data.buffer).set(bytes.slice(headerEnd, headerEnd + need));
  const col = (name) => {
    const k = props.indexOf(name);
    if (k < 0) throw new Error('fromPly: no property ' + name);
    return k;
  };
  const ix = ['x', 'y', 'z', 'scale_0', 'scale_1', 'scale_2', 'rot_0', 'rot_1', 'rot_2', 'rot_3', 'opacity', 'f_dc_0',
    'f_dc_1', 'f_dc_2'].map(col);
  let rest = 0;
  while (props.indexOf('f_rest_' + rest) >= 0) rest++;
  const per = Math.floor(rest / 3);
  const restAt = rest ? col('f_rest_0') : -1;
  const out = [];
  for (let i = 0; i < count; i++) {
    const v = (k) => data[i * P + ix[k]];
    const dc = [v(11) * SH0 + 0.5, v(12) * SH0 + 0.5, v(13) * SH0 + 0.5];
    const g = {
      mean: [v(0), v(1), v(2)],
      logScale: [v(3), v(4), v(5)],
      rotation: [v(6), v(7), v(8), v(9)],
      opacityLogit: v(10),
      color: dc.map((c) => Math.min(1, Math.max(0, c))),
    };
    if (per > 0) {
      const sh = new Float32Array(48);
      sh[0] = dc[0]; sh[1] = dc[1]; sh[2] = dc[2];
      for (let c = 0; c < 3; c++)
        for (let k = 1; k <= Math.min(per, 15); k++) sh[k * 3 + c] = data[i * P + restAt + c * per + k - 1];
      g.sh = sh;
    }
    out.push(g);
  }
  return out;
}

/**
 * toSplat (gpu.js:539-560): 32 bytes a splat (position, scale as floats; colour,
 * opacity, normalised rotation as bytes), largest and most opaque first, splats
 * with a non-finite value left out.  Returns a Buffer.
 */
function toSplat(all) {
  const finite = (g) => [].concat(g.mean, g.logScale, [g.opacityLogit], g.rotation, g.color).every(Number.isFinite);
  const gaussians = all.filter(finite);
  const weight = (g) => Math.exp(g.logScale[0] + g.logScale[1] + g.logScale[2]) / (1 + Math.exp(-g.opacityLogit));
  const order = gaussians.map((g, i) => [weight(g), i]);
  order.sort((a, b) => b[0] - a[0] || a[1] - b[1]);  // heaviest first, ties in input order
  const out = new ArrayBuffer(gaussians.length * 32);
  const f = new Float32Array(out);
  const u = new Uint8Array(out);
  const byte = (v) => Math.max(0, Math.min(255, Math.round(v)));
  order.forEach((entry, k) => {
    const g = gaussians[entry[1]];
    for (let j = 0; j < 3; j++) {
      f[k * 8 + j] = g.mean[j];
      f[k * 8 + 3 + j] = Math.exp(g.logScale[j]);
    }
    const q = g.rotation;
    const len = Math.hypot(q[0], q[1], q[2], q[3]) || 1;
    const o = k * 32 + 24;
    for (let c = 0; c < 3; c++) u[o + c] = byte(g.color[c] * 255);
    u[o + 3] = byte(255 / (1 + Math.exp(-g.opacityLogit)));
    for (let j = 0; j < 4; j++) u[o + 4 + j] = byte((q[j] / len) * 128 + 128);
  });
  return Buffer.from(out);
}

module.exports = {
  SH0, packGaussians, unpackGaussians, cameraUniform, sceneExtent, shuffled, neighbourSizes, initFromPoints,
  densify, toPly, fromPly, toSplat,
};
