'use strict';
// The JavaScript API end to end on the GPU: Camera -> SplatScene -> SplatPass.prepare -> sort ->
// read, checked against the JavaScript CPU path (oracle/splat_oracle.js) in the same process.
const assert = require('assert');
const path = require('path');

const root = path.join(__dirname, '..', '..');
const wd = require(path.join(root, 'winding_amd', 'js', 'index.js'));
const oracle = require(path.join(root, 'oracle', 'splat_oracle.js'));

let passed = 0;
function test(name, fn) {
  fn();
  passed++;
  console.log('  ok  ' + name);
}

function record(pos, scale, rgba, quat) {
  const b = new Uint8Array(32);
  const v = new DataView(b.buffer);
  pos.concat(scale).forEach((x, i) => v.setFloat32(i * 4, x, true));
  b.set(rgba, 24);
  b.set(quat, 28);
  return b;
}

const gpu = wd.Gpu.open(0);

function sortOnce(bytes, cam, width, height, cull) {
  const splats = new wd.Splats(gpu, wd.parseSplats(bytes));
  const scene = new wd.SplatScene();
  const entity = scene.addSplats({ splats });
  const pass = new wd.SplatPass(gpu);
  pass.cull = cull;
  pass.prepare(scene, cam, null, width, height);
  pass.sort();
  const got = pass.read(entity, { cov2d: true });
  const uniform = wd.cloudUniform(new Float32Array(37), cam.view, new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]),
    cam.projection, width, height, splats.min, splats.max, cull);
  pass.destroy();
  splats.destroy();
  return { got, uniform };
}

test('65,536 seeded splats: order bit-exact with the JS CPU path, footprints within 1e-5', () => {
  const bytes = gpu.native.synth(0, 0, 65536);
  const cam = new wd.Camera({ fovY: Math.PI / 3, near: 0.1 });
  cam.position.set([0, 0, 10]);
  cam.update(1280 / 720);
  const { got, uniform } = sortOnce(bytes, cam, 1280, 720, 0.5);
  const view = oracle.viewFromFloats(uniform);
  const cov = new Float64Array(65536 * 3);
  const want = oracle.run(bytes, view, cov);
  assert.strictEqual(got.count, want.length);
  assert.deepStrictEqual(Array.from(got.order), Array.from(want));
  let worst = 0;
  for (const i of want) {
    const s = Math.max(Math.abs(cov[3 * i]), Math.abs(cov[3 * i + 2]));
    for (let k = 0; k < 3; k++) worst = Math.max(worst, Math.abs(got.cov2d[4 * i + k] - cov[3 * i + k]) / s);
  }
  assert.ok(worst <= 1e-5, 'cov2d error ' + worst);
  console.log('      V = ' + got.count + ', cov2d max rel err ' + worst.toExponential(2));
});

test('two splats on the view axis are drawn back to front from either side (test/gpu.test.js:1804-1870)', () => {
  const red = record([0, 0, 1], [0.4, 0.4, 0.4], [255, 0, 0, 255], [255, 128, 128, 128]);
  const blue = record([0, 0, -1], [0.4, 0.4, 0.4], [0, 0, 255, 255], [255, 128, 128, 128]);
  for (const [first, second, ri, bi] of [[red, blue, 0, 1], [blue, red, 1, 0]]) {
    const bytes = new Uint8Array(64);
    bytes.set(first, 0);
    bytes.set(second, 32);
    const cam = new wd.Camera({ fovY: 0.8, near: 0.1 });
    cam.position.set([0, 0, 5]);
    cam.update(160 / 120);
    assert.deepStrictEqual(Array.from(sortOnce(bytes, cam, 160, 120, 0.5).got.order), [bi, ri]);
    cam.position.set([0, 0, -5]);
    cam.update(160 / 120);
    assert.deepStrictEqual(Array.from(sortOnce(bytes, cam, 160, 120, 0.5).got.order), [ri, bi]);
  }
});

test('a faint speck is left out at the default cull and kept at 0 (test/gpu.test.js:1914-1962)', () => {
  const bytes = new Uint8Array(64);
  bytes.set(record([0, 0, 0], [0.0005, 0.0005, 0.0005], [255, 255, 255, 50], [255, 128, 128, 128]), 0);
  bytes.set(record([1, 0, 0], [0.2, 0.2, 0.2], [255, 255, 255, 255], [255, 128, 128, 128]), 32);
  const cam = new wd.Camera({ fovY: 0.8, near: 0.1 });
  cam.position.set([0, 0, 5]);
  cam.update(160 / 120);
  assert.deepStrictEqual(Array.from(sortOnce(bytes, cam, 160, 120, 0.5).got.order), [1]);
  assert.deepStrictEqual(Array.from(sortOnce(bytes, cam, 160, 120, 0).got.order).sort(), [0, 1]);
});

test('a NaN position is refused at upload, by index', () => {
  const bytes = new Uint8Array(64);
  bytes.set(record([0, 0, 0], [1, 1, 1], [0, 0, 0, 255], [255, 128, 128, 128]), 0);
  bytes.set(record([NaN, 0, 0], [1, 1, 1], [0, 0, 0, 255], [255, 128, 128, 128]), 32);
  assert.throws(() => new wd.Splats(gpu, wd.parseSplats(bytes)), /splat 1 has a position that is not a number/);
});

test('three clouds sorted in one batch match each cloud sorted alone', () => {
  const cam = new wd.Camera({ fovY: Math.PI / 3, near: 0.1 });
  cam.position.set([2, 1, 9]);
  cam.update(1280 / 720);
  const sizes = [5000, 70001, 4096];
  const clouds = sizes.map((n, k) => new wd.Splats(gpu, wd.parseSplats(gpu.native.synth(k % 2, 50 + k, n))));
  const scene = new wd.SplatScene();
  const entities = clouds.map((splats) => scene.addSplats({ splats }));
  const pass = new wd.SplatPass(gpu);
  pass.prepare(scene, cam, null, 1280, 720);
  assert.strictEqual(pass.sort(), 3);   // one batch launch
  const batched = entities.map((e) => pass.read(e));
  clouds.forEach((splats, k) => {
    const one = sortOnce(gpu.native.synth(k % 2, 50 + k, sizes[k]), cam, 1280, 720, 0.5).got;
    assert.strictEqual(batched[k].count, one.count);
    assert.deepStrictEqual(Array.from(batched[k].order), Array.from(one.order));
  });
  // the autotuner keeps the fastest configuration; the sorts after it order the same
  const report = pass.tune(2);
  assert.ok(['phased', 'persistent'].includes(report.schedule));
  assert.strictEqual(report.tile, 256 * report.keysPerThread);
  assert.strictEqual(report.timed.length, 6);
  assert.ok(report.timed.every((t) => t.ms > 0));
  assert.throws(() => gpu.native.batchConfigure(pass._batch, 'phased', 3, 16), /super-tiles of 3/);
  assert.throws(() => gpu.native.batchConfigure(pass._batch, 'sideways', 1, 16), /'phased' or 'persistent'/);
  gpu.native.batchConfigure(pass._batch, 'persistent', 8, 8);
  pass.prepare(scene, cam, null, 1280, 720);
  pass._items.forEach((i) => { i.sort = true; });
  assert.strictEqual(pass.sort(), 3);
  entities.forEach((e, k) => assert.deepStrictEqual(Array.from(pass.read(e).order), Array.from(batched[k].order)));
  pass.destroy();
  clouds.forEach((c) => c.destroy());
});

test('learned sort: a predictor trained on the exact sort, then 4 views, each equal to the exact order', () => {
  const n = 200000;
  const splats = new wd.Splats(gpu, wd.parseSplats(gpu.native.synth(0, 3, n)));
  const world = new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]);
  const uniformAt = (a, r, y) => {
    const cam = new wd.Camera({ fovY: Math.PI / 3, near: 0.1 });
    cam.position.set([r * Math.sin(a), y, r * Math.cos(a)]);
    cam.update(1280 / 720);
    return wd.cloudUniform(new Float32Array(37), cam.view, world, cam.projection, 1280, 720, splats.min, splats.max, 0.5).slice();
  };
  const train = [];
  for (let k = 0; k < 256; k++) train.push(uniformAt(2 * Math.PI * k / 256, 9 + 3 * ((k * 37) % 11) / 11, 3 * Math.sin(k)));
  const data = wd.trainingSet(gpu, splats, train);
  const predictor = new wd.Predictor(gpu, 1);
  predictor.train(data.poses, data.targets, { epochs: 150, batch: 64, lr: 2e-3, seed: 2 });
  const bytes = predictor.save();
  const copy = new wd.Predictor(gpu, 7);
  copy.load(bytes);
  const sorter = new wd.LearnedSort(gpu, copy, 8);
  const views = [0.3, 1.9, 3.4, 5.0].map((a) => uniformAt(a, 10, 1));
  sorter.sort(splats, views);
  const exact = gpu.native.stateCreate(gpu.handle, n);
  views.forEach((u, k) => {
    const got = sorter.read(k);
    gpu.native.sort(gpu.handle, splats.handle, u, exact);
    const want = gpu.native.stateRead(gpu.handle, exact, 0);
    assert.strictEqual(got.count, want.count);
    assert.deepStrictEqual(Array.from(got.order), Array.from(want.order));
    assert.strictEqual(got.fallback, 0);
  });
  gpu.native.destroy(exact);
  sorter.destroy();
  copy.destroy();
  predictor.destroy();
  splats.destroy();
});

test('a capture with harmonics changes colour with the view, in its own space (test/gpu.test.js:1968-2035)', () => {
  // A grey splat whose only harmonic is red's third coefficient, the -x term, at 0.6 (half 0x38cd).
  const bytes = record([0, 0, 0], [0.4, 0.4, 0.4], [128, 128, 128, 255], [255, 128, 128, 128]);
  const sh = new Uint32Array(wd.shWords(1));
  sh[3] = 0x38cd;  // halves: coefficient k, channel c at 3k + c; red of k = 2 is half 6, word 3's low half
  const splats = new wd.Splats(gpu, { count: 1, degree: 1, records: bytes, sh });
  const colourFrom = (x, turned) => {
    const scene = new wd.SplatScene();
    const entity = scene.addSplats({ splats, matrix: turned ? [-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1] : undefined });
    const cam = new wd.Camera({ fovY: 0.8, near: 0.1 });
    cam.position.set([x, 0, 0]);
    cam.update(160 / 120);
    const pass = new wd.SplatPass(gpu);
    pass.cull = 0;
    pass.prepare(scene, cam, null, 160, 120);
    pass.sort();
    const got = pass.read(entity, { cov2d: true });
    pass.destroy();
    const bits = new Uint32Array(got.cov2d.buffer, got.cov2d.byteOffset, 4)[3];
    return [bits & 255, (bits >> 8) & 255, (bits >> 16) & 255];
  };
  const plusX = colourFrom(5, false), minusX = colourFrom(-5, false), turned = colourFrom(-5, true);
  splats.destroy();
  assert.ok(plusX[0] > plusX[1] + 40 && minusX[0] < minusX[1] - 40, plusX + ' / ' + minusX);
  assert.ok(Math.abs(plusX[1] - minusX[1]) <= 2 && Math.abs(turned[0] - plusX[0]) <= 2, turned + ' / ' + plusX);
  console.log('      from +x ' + plusX + ', from -x ' + minusX + ', turned ' + turned);
});

test('the draw\'s ellipse set-up beside cov2d: a round splat seen head on (src/render/splats.js:263-291)', () => {
  const bytes = record([0, 0, 0], [0.4, 0.4, 0.4], [200, 100, 50, 255], [255, 128, 128, 128]);
  const splats = new wd.Splats(gpu, wd.parseSplats(bytes));
  const scene = new wd.SplatScene();
  const entity = scene.addSplats({ splats });
  const cam = new wd.Camera({ fovY: 0.8, near: 0.1 });
  cam.position.set([0, 0, 5]);
  cam.update(160 / 120);
  const pass = new wd.SplatPass(gpu);
  pass.cull = 0;
  pass.ellipses = true;
  pass.prepare(scene, cam, null, 160, 120);
  pass.sort();
  const got = pass.read(entity, { cov2d: true, ellipses: true });
  pass.destroy();
  // beside a second cloud the two share one batch launch, and the ellipses come out the same
  const other = new wd.Splats(gpu, wd.parseSplats(record([0.5, 0, -1], [0.2, 0.3, 0.2], [10, 20, 30, 200], [255, 128, 128, 128])));
  const scene2 = new wd.SplatScene();
  const first = scene2.addSplats({ splats });
  scene2.addSplats({ splats: other });
  const pass2 = new wd.SplatPass(gpu);
  pass2.cull = 0;
  pass2.ellipses = true;
  pass2.prepare(scene2, cam, null, 160, 120);
  assert.strictEqual(pass2.sort(), 2);
  const batched = pass2.read(first, { ellipses: true });
  pass2.destroy();
  other.destroy();
  splats.destroy();
  assert.deepStrictEqual(new Uint32Array(batched.ellipses.buffer, batched.ellipses.byteOffset, 8),
    new Uint32Array(got.ellipses.buffer, got.ellipses.byteOffset, 8));
  assert.strictEqual(got.count, 1);
  const [xx, xy, yy] = got.cov2d;
  const e = got.ellipses;
  assert.ok(Math.abs(xy) < 1e-6 * xx && Math.abs(xx - yy) < 1e-6 * xx, xx + ' ' + xy + ' ' + yy);
  // alpha 255 reaches log(255) > 4, so e = 2; a circle's axes run along x, then down y
  const r = Math.sqrt(2 * (xx + 0.3)) * 2;
  assert.ok(Math.abs(e[3] - 5) < 1e-5 && Math.abs(e[0]) < 1e-6 && Math.abs(e[1]) < 1e-6, 'clip ' + Array.from(e.subarray(0, 4)));
  const close = (a, b) => Math.abs(a - b) <= 1e-5 * r;
  assert.ok(close(Math.hypot(e[4], e[5]), r) && close(Math.hypot(e[6], e[7]), r), Array.from(e.subarray(4)));
  assert.ok(Math.abs(e[4] * e[6] + e[5] * e[7]) <= 1e-5 * r * r, 'axes at right angles');
  console.log('      clip w ' + e[3] + ', axes ' + Array.from(e.subarray(4)).map((x) => x.toFixed(3)));
});

test('an emitter\'s particles sort far to near, whatever slots they were born in (test/gpu.test.js:2666-2760)', () => {
  const depths = [3, 40, 0.5, 12, 200, 7];
  const bytes = new Uint8Array(32 * (depths.length + 2));
  const f = new Float32Array(bytes.buffer);
  depths.forEach((d, k) => {
    const s = 8 * (k + 2);  // the ring starts two slots in
    f.set([0.1 * k, 0, 100 - d, 0.5, 0, 0, 0, 2], s);
  });
  const native = gpu.native;
  const pool = native.poolCreate(gpu.handle, bytes);
  const state = native.stateCreate(gpu.handle, depths.length);
  native.particlesSort(gpu.handle, pool, 2, depths.length, new Float32Array([0, 0, 100, 0, 0, 0, -1, 0]), state);
  const got = native.stateRead(gpu.handle, state, 0);
  native.destroy(state);
  native.destroy(pool);
  const want = depths.map((d, k) => [d, k + 2]).sort((a, b) => b[0] - a[0]).map((x) => x[1]);
  assert.deepStrictEqual(Array.from(got.order), want);
  assert.strictEqual(got.count, depths.length);
});

test('captures decoded on the GPU: a .ply, a .spz and a .sog through parseSplats / readSplats and new Splats', () => {
  const zlib = require('zlib');
  const SH_C0 = 0.28209479177387814;
  const props = ['x', 'y', 'z', 'f_dc_0', 'f_dc_1', 'f_dc_2', 'opacity', 'scale_0', 'scale_1', 'scale_2', 'rot_0', 'rot_1', 'rot_2', 'rot_3'];
  const rows = [
    { x: -1, y: 0, z: 4, f_dc_0: (1 - 0.5) / SH_C0, f_dc_1: 0, f_dc_2: (0 - 0.5) / SH_C0, opacity: 0, scale_0: Math.log(2), scale_1: Math.log(3), scale_2: 0, rot_0: 2 },
    { x: 5, y: -2, z: 0, rot_0: 1, opacity: 2 },
  ];
  const head = Buffer.from('ply\nformat binary_little_endian 1.0\nelement vertex 2\n' + props.map((p) => 'property float ' + p + '\n').join('') + 'end_header\n');
  const body = new Float32Array(rows.length * props.length);
  rows.forEach((row, i) => props.forEach((p, k) => { body[i * props.length + k] = row[p] || 0; }));
  const ply = new Uint8Array(Buffer.concat([head, Buffer.from(body.buffer)]));
  const splats = new wd.Splats(gpu, wd.parseSplats(ply));
  assert.strictEqual(splats.count, 2);
  assert.strictEqual(splats.degree, 0);
  assert.deepStrictEqual(Array.from(splats.min), [-1, -2, 0]);
  assert.deepStrictEqual(Array.from(splats.max), [5, 0, 4]);
  const scene = new wd.SplatScene();
  const entity = scene.addSplats({ splats });
  const cam = new wd.Camera({ fovY: 0.8, near: 0.1 });
  cam.position.set([0, 0, 20]);
  cam.update(160 / 120);
  const pass = new wd.SplatPass(gpu);
  pass.cull = 0;
  pass.prepare(scene, cam, null, 160, 120);
  pass.sort();
  const got = pass.read(entity, { cov2d: true });
  pass.destroy();
  splats.destroy();
  assert.strictEqual(got.count, 2);
  assert.deepStrictEqual(Array.from(got.order), [1, 0]);  // z 0 is farther from the eye at z 20 than z 4
  const bits = new Uint32Array(got.cov2d.buffer, got.cov2d.byteOffset, 8)[3];
  assert.deepStrictEqual([bits & 255, (bits >> 8) & 255, (bits >> 16) & 255, bits >>> 24], [255, 128, 0, 128]);
  // a .spz of version 2: one splat, gzipped; the library gunzips it
  const spz = new Uint8Array(16 + 9 + 1 + 3 + 3 + 3);
  const sv = new DataView(spz.buffer);
  sv.setUint32(0, 0x5053474e, true); sv.setUint32(4, 2, true); sv.setUint32(8, 1, true); spz[13] = 12;
  spz.set([0, 0x10, 0, 0, 0xf0, 0xff, 0, 0xf8, 0xff], 16);  // x = 1, y = -1 and z = -0.5, turned to 1 and 0.5
  spz[25] = 200; spz.set([128, 128, 128], 26); spz.set([100, 100, 100], 29); spz.set([128, 128, 128], 32);
  const s = new wd.Splats(gpu, { format: 'spz', bytes: new Uint8Array(zlib.gzipSync(Buffer.from(spz))) });
  assert.strictEqual(s.count, 1);
  assert.deepStrictEqual(Array.from(s.min), [1, 1, 0.5]);
  s.destroy();
});

test('SplatPass batches a cloud with harmonics with a plain one, as their own sorts would order and shade them', () => {
  const bytes = gpu.native.synth(0, 3, 5000);
  const sh = new Uint32Array(wd.shWords(2) * 5000);
  for (let k = 0; k < sh.length; k++) sh[k] = ((k * 2654435761) >>> 0) & 0x33ff33ff;  // halves below 0.25 in size
  const shaded = new wd.Splats(gpu, { count: 5000, degree: 2, records: bytes, sh });
  const plain = new wd.Splats(gpu, wd.parseSplats(gpu.native.synth(0, 4, 7000)));
  const scene = new wd.SplatScene();
  const a = scene.addSplats({ splats: shaded });
  const b = scene.addSplats({ splats: plain, matrix: [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 2, 0, 0, 1] });
  const cam = new wd.Camera({ fovY: 1.0, near: 0.1 });
  cam.position.set([0, 2, 12]);
  cam.update(640 / 480);
  const pass = new wd.SplatPass(gpu);
  pass.prepare(scene, cam, null, 640, 480);
  assert.strictEqual(pass.sort(), 2);
  const got = [pass.read(a, { cov2d: true }), pass.read(b, { cov2d: true })];
  // each cloud alone, with its own uniform, through the single sort
  [[shaded, a], [plain, b]].forEach(([splats, entity], k) => {
    const state = pass._states.get(scene.splats.get(entity));
    const alone = gpu.native.stateCreate(gpu.handle, splats.count);
    gpu.native.sort(gpu.handle, splats.handle, state.uniform, alone);
    const want = gpu.native.stateRead(gpu.handle, alone, splats.count);
    gpu.native.destroy(alone);
    assert.strictEqual(got[k].count, want.count);
    assert.deepStrictEqual(Array.from(got[k].order), Array.from(want.order));
    assert.deepStrictEqual(new Uint32Array(got[k].cov2d.buffer, got[k].cov2d.byteOffset, splats.count * 4),
      new Uint32Array(want.cov2d.buffer, want.cov2d.byteOffset, splats.count * 4));
  });
  pass.destroy();
  shaded.destroy();
  plain.destroy();
});

/** A zip of the given files, stored (method 0), enough for readSplats' unzip. */
// CRC-32 bit by bit (no table), for the zips these tests write
function crc32(data) {
  let c = 0xffffffff;
  for (const b of data) {
    c ^= b;
    for (let k = 0; k < 8; k++) c = (c >>> 1) ^ (0xedb88320 & -(c & 1));
  }
  return (c ^ 0xffffffff) >>> 0;
}

function storedZip(files) {
  const parts = [], directory = [];
  let offset = 0;
  for (const [name, data] of Object.entries(files)) {
    const n = Buffer.from(name);
    const crc = crc32(data);
    const local = Buffer.alloc(30 + n.length);
    local.writeUInt32LE(0x04034b50, 0); local.writeUInt32LE(crc, 14); local.writeUInt32LE(data.length, 18);
    local.writeUInt32LE(data.length, 22); local.writeUInt16LE(n.length, 26); n.copy(local, 30);
    const central = Buffer.alloc(46 + n.length);
    central.writeUInt32LE(0x02014b50, 0); central.writeUInt32LE(crc, 16); central.writeUInt32LE(data.length, 20);
    central.writeUInt32LE(data.length, 24); central.writeUInt16LE(n.length, 28); central.writeUInt32LE(offset, 42);
    n.copy(central, 46);
    parts.push(local, Buffer.from(data));
    directory.push(central);
    offset += local.length + data.length;
  }
  const size = directory.reduce((a, d) => a + d.length, 0);
  const end = Buffer.alloc(22);
  end.writeUInt32LE(0x06054b50, 0); end.writeUInt16LE(directory.length, 10); end.writeUInt32LE(size, 12); end.writeUInt32LE(offset, 16);
  return new Uint8Array(Buffer.concat([...parts, ...directory, end]));
}

(async () => {
  // a one-splat .sog without harmonics: its images are their pixels as JSON, which decodeImage reads
  const image = (rgba) => Buffer.from(JSON.stringify({ width: 1, height: 1, rgba }));
  const meta = {
    version: 2, count: 1, means: { mins: [-1, -1, -1], maxs: [1, 1, 1], files: ['l.webp', 'u.webp'] },
    scales: { codebook: Array.from({ length: 256 }, (_, k) => -6 + k * 0.05), files: ['s.webp'] },
    quats: { files: ['q.webp'] }, sh0: { codebook: Array.from({ length: 256 }, (_, k) => -2 + k / 64), files: ['c.webp'] },
  };
  const bytes = storedZip({
    'meta.json': Buffer.from(JSON.stringify(meta)),
    'l.webp': image([0, 0xff, 0x80, 255]), 'u.webp': image([0x80, 0xff, 0x80, 255]),
    'q.webp': image([128, 128, 128, 0]), 's.webp': image([100, 100, 100, 255]), 'c.webp': image([160, 128, 96, 200]),
  });
  const decodeImage = async (file) => JSON.parse(Buffer.from(file).toString());
  const data = await wd.readSplats(bytes, { decodeImage });
  const s = new wd.Splats(gpu, data);
  assert.strictEqual(s.count, 1);
  assert.strictEqual(s.degree, 0);
  // x: stored 0x8000 of 65535 between -1 and 1 is just above 0 before exp(|n|) - 1
  assert.ok(Math.abs(s.min[0] - (Math.exp(-1 + 2 * 32768 / 65535) - 1)) < 1e-6 && s.min[1] > 1.7, Array.from(s.min));
  s.destroy();
  await assert.rejects(wd.readSplats(bytes), /a .sog needs decodeImage/);
  passed++;
  console.log('  ok  a .sog through readSplats (zip, meta.json, decodeImage) and new Splats');
  gpu.destroy();
  console.log(passed + ' checks passed');
})().catch((error) => {
  console.error(error);
  process.exit(1);
});
