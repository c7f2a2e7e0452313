'use strict';
// splat_oracle.js -- the JavaScript CPU path: Winding's splat projection, key and
// stable counting sort, restated in plain JS (doubles) on the CPU.
//
// TEST INFRASTRUCTURE ONLY: the parity tests cross-check it against
// oracle/splat_oracle.c bit for bit, and the benchmark times it on the host's
// cores as the CPU baseline. The product never loads it.
//
// Same canonical definition as splat_oracle.c: the reference's formulas in
// double from the f32 inputs, in the reference's expression order; JavaScript
// has no fused multiply-add, and Math.sqrt and division are IEEE.
//   sigma      put()            src/scene/splats.js:350-368
//   footprint  footprint()      src/render/splats.js:102-125
//   splat      makeKeys()       src/render/splats.js:155-186
//   sort       DrawList idiom   src/render/drawlist.js:158-203, one 65536-bucket pass, stable
//   prepare    SplatPass.prepare src/render/splats.js:559-587
// Node 12 syntax (no ?? or ?.), CommonJS.

const C13 = Math.fround(1.3);
const C03 = Math.fround(0.3);
const C2PI = Math.fround(6.2831853);
const BUCKETS = 65536;

/** The view uniform as a plain object from 37 floats (the wd_view layout). */
function viewFromFloats(f) {
  return {
    modelView: f.subarray(0, 16), projection: f.subarray(16, 32),
    viewport: [f[32], f[33]], depthMin: f[34], depthScale: f[35], cull: f[36],
  };
}

function sigma(dv, o, out) {
  const sx = dv.getFloat32(o + 12, true), sy = dv.getFloat32(o + 16, true), sz = dv.getFloat32(o + 20, true);
  const b = (k) => dv.getUint8(o + 28 + k);
  const qw = (b(0) - 128) / 128, qx = (b(1) - 128) / 128, qy = (b(2) - 128) / 128, qz = (b(3) - 128) / 128;
  const length = Math.sqrt(qw * qw + qx * qx + qy * qy + qz * qz) || 1;
  const w = qw / length, u = qx / length, v = qy / length, t = qz / length;
  const r00 = 1 - 2 * (v * v + t * t), r01 = 2 * (u * v - w * t), r02 = 2 * (u * t + w * v);
  const r10 = 2 * (u * v + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (v * t - w * u);
  const r20 = 2 * (u * t - w * v), r21 = 2 * (v * t + w * u), r22 = 1 - 2 * (u * u + v * v);
  const m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  const m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  const m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  out[0] = m00 * m00 + m01 * m01 + m02 * m02;
  out[1] = m00 * m10 + m01 * m11 + m02 * m12;
  out[2] = m00 * m20 + m01 * m21 + m02 * m22;
  out[3] = m10 * m10 + m11 * m11 + m12 * m12;
  out[4] = m10 * m20 + m11 * m21 + m12 * m22;
  out[5] = m20 * m20 + m21 * m21 + m22 * m22;
  return out;
}

function footprint(view, vx, vy, vz, sg, cov) {
  const p = view.projection, m = view.modelView;
  const p00 = p[0], p11 = p[5];
  const fx = p00 * view.viewport[0] * 0.5;
  const fy = p11 * view.viewport[1] * 0.5;
  let j00 = fx, j11 = fy, j02 = 0, j12 = 0;
  if (p[15] === 0) {
    const d = -vz;
    const lx = C13 / p00, ly = C13 / p11;
    let cx = vx / d, cy = vy / d;
    if (cx < -lx) cx = -lx;
    if (cx > lx) cx = lx;
    if (cy < -ly) cy = -ly;
    if (cy > ly) cy = ly;
    const x = cx * d, y = cy * d;
    j00 = fx / d;
    j11 = fy / d;
    j02 = fx * x / (d * d);
    j12 = fy * y / (d * d);
  }
  const t00 = j00 * m[0] + j02 * m[2], t01 = j00 * m[4] + j02 * m[6], t02 = j00 * m[8] + j02 * m[10];
  const t10 = j11 * m[1] + j12 * m[2], t11 = j11 * m[5] + j12 * m[6], t12 = j11 * m[9] + j12 * m[10];
  const s00 = sg[0], s01 = sg[1], s02 = sg[2], s11 = sg[3], s12 = sg[4], s22 = sg[5];
  const u00 = t00 * s00 + t01 * s01 + t02 * s02;
  const u01 = t00 * s01 + t01 * s11 + t02 * s12;
  const u02 = t00 * s02 + t01 * s12 + t02 * s22;
  const u10 = t10 * s00 + t11 * s01 + t12 * s02;
  const u11 = t10 * s01 + t11 * s11 + t12 * s12;
  const u12 = t10 * s02 + t11 * s12 + t12 * s22;
  cov[0] = u00 * t00 + u01 * t01 + u02 * t02;
  cov[1] = u00 * t10 + u01 * t11 + u02 * t12;
  cov[2] = u10 * t10 + u11 * t11 + u12 * t12;
  return cov;
}

const SG = new Float64Array(6);
const FP = new Float64Array(3);

/** makeKeys for splat i: returns the key (0..65535) of a survivor, or -1 when culled. cov gets the footprint. */
function splat(dv, i, view, cov) {
  const o = i * 32;
  const x = dv.getFloat32(o, true), y = dv.getFloat32(o + 4, true), z = dv.getFloat32(o + 8, true);
  const m = view.modelView, P = view.projection;
  const v0 = m[0] * x + m[4] * y + m[8] * z + m[12];
  const v1 = m[1] * x + m[5] * y + m[9] * z + m[13];
  const v2 = m[2] * x + m[6] * y + m[10] * z + m[14];
  const v3 = m[3] * x + m[7] * y + m[11] * z + m[15];
  const c0 = P[0] * v0 + P[4] * v1 + P[8] * v2 + P[12] * v3;
  const c1 = P[1] * v0 + P[5] * v1 + P[9] * v2 + P[13] * v3;
  const c2 = P[2] * v0 + P[6] * v1 + P[10] * v2 + P[14] * v3;
  const w = P[3] * v0 + P[7] * v1 + P[11] * v2 + P[15] * v3;
  if (v2 >= 0 || c2 > w || Math.abs(c0) > C13 * w || Math.abs(c1) > C13 * w) return -1;
  const a = dv.getUint8(o + 27);
  if (a <= 1) return -1;
  // Sigma as put() stores it: each term rounded to f32 (a Float32Array, src/scene/splats.js:318,361-367)
  const sg = sigma(dv, o, SG);
  for (let k = 0; k < 6; k++) sg[k] = Math.fround(sg[k]);
  const f = footprint(view, v0, v1, v2, sg, cov || FP);
  if (view.cull > 0) {
    const det = (f[0] + C03) * (f[2] + C03) - f[1] * f[1];
    const area = C2PI * Math.sqrt(det > 0 ? det : 0);
    const alpha = a / 255;
    if (alpha * area < view.cull) return -1;
  }
  let t = (-v2 - view.depthMin) * view.depthScale;
  if (t < 0) t = 0;
  if (t > 65535) t = 65535;
  return 65535 - Math.floor(t);
}

/** Keys of every splat (Int32Array, -1 culled) and, when cov is given (n x 3), the footprints. */
function keys(bytes, view, cov) {
  const n = bytes.length / 32;
  const dv = new DataView(bytes.buffer, bytes.byteOffset, bytes.byteLength);
  const out = new Int32Array(n);
  const c = new Float64Array(3);
  for (let i = 0; i < n; i++) {
    const k = splat(dv, i, view, c);
    out[i] = k;
    if (cov && k >= 0) { cov[3 * i] = c[0]; cov[3 * i + 1] = c[1]; cov[3 * i + 2] = c[2]; }
  }
  return out;
}

/** Stable counting sort of the survivors (key >= 0), ties by index. */
function sort(k) {
  const counts = new Uint32Array(BUCKETS);
  const n = k.length;
  for (let i = 0; i < n; i++) if (k[i] >= 0) counts[k[i]]++;
  let total = 0;
  for (let b = 0; b < BUCKETS; b++) {
    const c = counts[b];
    counts[b] = total;
    total += c;
  }
  const order = new Uint32Array(total);
  for (let i = 0; i < n; i++) if (k[i] >= 0) order[counts[k[i]]++] = i;
  return order;
}

function run(bytes, view, cov) {
  return sort(keys(bytes, view, cov));
}

/** SplatPass.prepare's uniform (37 floats), or null when the box is wholly behind. */
function prepare(viewM, world, projection, width, height, min, max, cull) {
  const b = world || new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]);
  const mv = new Float32Array(16);
  for (let c = 0; c < 4; c++) {
    const b0 = b[c * 4], b1 = b[c * 4 + 1], b2 = b[c * 4 + 2], b3 = b[c * 4 + 3];
    for (let r = 0; r < 4; r++) mv[c * 4 + r] = b0 * viewM[r] + b1 * viewM[4 + r] + b2 * viewM[8 + r] + b3 * viewM[12 + r];
  }
  let near = Infinity, far = -Infinity;
  for (let k = 0; k < 8; k++) {
    const x = k & 1 ? max[0] : min[0];
    const y = k & 2 ? max[1] : min[1];
    const z = k & 4 ? max[2] : min[2];
    const depth = -(mv[2] * x + mv[6] * y + mv[10] * z + mv[14]);
    if (depth < near) near = depth;
    if (depth > far) far = depth;
  }
  if (far <= 0) return null;
  near = Math.max(near, 0);
  const f = new Float32Array(37);
  f.set(mv, 0);
  f.set(projection, 16);
  f[32] = width;
  f[33] = height;
  f[34] = near;
  f[35] = (BUCKETS - 1) / Math.max(far - near, 1e-6);
  f[36] = cull;
  return f;
}

module.exports = { viewFromFloats, sigma, footprint, splat, keys, sort, run, prepare, BUCKETS };
