'use strict';
// The JavaScript API without a GPU, in the reference's test style (a plain script, node:assert):
// parseSplats' refusals, the per-cloud uniform against the addon's C version and the JS oracle,
// and SplatPass' bookkeeping (sort-skip, TAA, eviction, refusals) over a fake addon, the way the
// reference's Node suites fake the rhi (test/render.test.js:589-600).
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

let seed = 12345;
function rand() {
  seed = (seed * 1103515245 + 12345) >>> 0;
  return seed / 4294967296;
}

test('parseSplats refuses what is not .splat records, by what is wrong with it', () => {
  assert.throws(() => wd.parseSplats(new Uint8Array(33)), /not a whole number of 32-byte/);
  assert.throws(() => wd.parseSplats(new Uint8Array(0)), /0 bytes is not a whole number/);
  const zip = new Uint8Array(64); zip.set([0x50, 0x4b, 0x03, 0x04]);
  assert.throws(() => wd.parseSplats(zip), /this is a zip/);
  const gz = new Uint8Array(64); gz.set([0x1f, 0x8b]);
  assert.throws(() => wd.parseSplats(gz), /this is gzipped/);
  const ply = new TextEncoder().encode('ply\nformat binary_little_endian 1.0\n');
  assert.strictEqual(wd.parseSplats(ply).format, 'ply');  // decoded on the GPU by new Splats
  const v4 = new Uint8Array(64); new DataView(v4.buffer).setUint32(0, 0x5053474e, true); new DataView(v4.buffer).setUint32(4, 4, true);
  assert.throws(() => wd.parseSplats(v4), /version 4, compressed with zstd/);
  const ok = wd.parseSplats(record([1, 2, 3], [1, 1, 1], [1, 2, 3, 4], [255, 128, 128, 128]));
  assert.strictEqual(ok.count, 1);
  assert.strictEqual(ok.degree, 0);
});

// A zip as a .sog packs it, written here from the format (local headers, data, central directory, end
// record): each entry stored or deflated, an optional trailing comment, CRC-32 computed bit by bit.
function zipOf(files, options) {
  const zlib = require('zlib');
  const opt = options || {};
  const crc = (data) => {
    let c = 0xffffffff;
    for (const b of data) {
      c ^= b;
      for (let k = 0; k < 8; k++) c = (c >>> 1) ^ (0xedb88320 & -(c & 1));
    }
    return (c ^ 0xffffffff) >>> 0;
  };
  const parts = [], directory = [];
  let offset = 0;
  for (const [name, raw] of Object.entries(files)) {
    const data = Buffer.from(raw);
    const packed = opt.deflate ? zlib.deflateRawSync(data) : data;
    const n = Buffer.from(name);
    const local = Buffer.alloc(30 + n.length);
    local.writeUInt32LE(0x04034b50, 0); local.writeUInt16LE(opt.deflate ? 8 : 0, 8); local.writeUInt32LE(crc(data), 14);
    local.writeUInt32LE(packed.length, 18); local.writeUInt32LE(data.length, 22); local.writeUInt16LE(n.length, 26);
    n.copy(local, 30);
    const central = Buffer.alloc(46 + n.length);
    central.writeUInt32LE(0x02014b50, 0); central.writeUInt16LE(opt.method !== undefined ? opt.method : (opt.deflate ? 8 : 0), 10);
    central.writeUInt32LE((crc(data) ^ (opt.badCrc ? 1 : 0)) >>> 0, 16); central.writeUInt32LE(opt.zip64 ? 0xffffffff : packed.length, 20);
    central.writeUInt32LE(data.length, 24); central.writeUInt16LE(n.length, 28); central.writeUInt32LE(offset, 42);
    n.copy(central, 46);
    parts.push(local, packed);
    directory.push(central);
    offset += local.length + packed.length;
  }
  const size = directory.reduce((a, d) => a + d.length, 0);
  const comment = Buffer.from(opt.comment || '');
  const end = Buffer.alloc(22);
  end.writeUInt32LE(0x06054b50, 0); end.writeUInt16LE(directory.length, 8); end.writeUInt16LE(directory.length, 10);
  end.writeUInt32LE(size, 12); end.writeUInt32LE(offset, 16); end.writeUInt16LE(comment.length, 20);
  return new Uint8Array(Buffer.concat([...parts, ...directory, end, comment]));
}

function sogFiles(meta) {
  const px = (rgba) => Buffer.from(JSON.stringify({ width: 1, height: 1, rgba }));
  return {
    'meta.json': JSON.stringify(meta), 'l.webp': px([0, 0, 0, 255]), 'u.webp': px([0, 0, 0, 255]),
    'q.webp': px([128, 128, 128, 0]), 's.webp': px([100, 100, 100, 255]), 'c.webp': px([1, 2, 3, 4]),
  };
}

const SOG_META = {
  version: 2, count: 1, means: { mins: [-1, -1, -1], maxs: [1, 1, 1], files: ['l.webp', 'u.webp'] },
  scales: { codebook: [0, 1], files: ['s.webp'] }, quats: { files: ['q.webp'] }, sh0: { codebook: [0, 1], files: ['c.webp'] },
};

async function zipChecks() {
  const decodeImage = async (file) => JSON.parse(Buffer.from(file).toString());
  const read = (bytes) => wd.readSplats(bytes, { decodeImage });
  // stored and deflated entries read alike, behind a trailing comment too
  for (const opt of [{}, { deflate: true }, { deflate: true, comment: 'made by a test' }]) {
    const got = await read(zipOf(sogFiles(SOG_META), opt));
    assert.strictEqual(got.format, 'sog');
    assert.strictEqual(got.sog.count, 1);
    assert.deepStrictEqual(Array.from(got.sog.images.sh0.rgba), [1, 2, 3, 4]);
    assert.deepStrictEqual(Array.from(got.sog.mins), [-1, -1, -1]);
  }
  const good = zipOf(sogFiles(SOG_META));
  await assert.rejects(read(good.subarray(0, good.length - 30)), /a zip with no end record: cut short\?/);
  const damaged = Uint8Array.from(good);
  const dir = new DataView(damaged.buffer).getUint32(damaged.length - 6, true);
  damaged[dir] ^= 0xff;
  await assert.rejects(read(damaged), /a zip whose directory is damaged/);
  await assert.rejects(read(zipOf(sogFiles(SOG_META), { zip64: true })), /a zip64 \.sog is not read/);
  await assert.rejects(read(zipOf(sogFiles(SOG_META), { method: 12 })), /meta\.json is compressed with method 12; stored and deflate are read/);
  await assert.rejects(read(zipOf(sogFiles(SOG_META), { badCrc: true })), /meta\.json in the zip is damaged/);
  const cut = zipOf({ 'meta.json': JSON.stringify(SOG_META) });
  const cv = new DataView(cut.buffer);
  cv.setUint32(cut.length - 22 - 46 - 9 + 20, 1 << 20, true);  // the directory claims more bytes than the file holds
  await assert.rejects(read(cut), /the zip is cut short in meta\.json/);
  const files = sogFiles(SOG_META);
  delete files['q.webp'];
  await assert.rejects(read(zipOf(files)), /the \.sog has no q\.webp/);
  await assert.rejects(read(zipOf(sogFiles(Object.assign({}, SOG_META, { version: undefined })))), /a \.sog of version 1; version 2 is read/);
  await assert.rejects(read(zipOf(sogFiles(Object.assign({}, SOG_META, { count: 0 })))), /a \.sog of 0 splats/);
  await assert.rejects(read(zipOf(sogFiles(Object.assign({}, SOG_META, { count: 2 })))), /l\.webp has 1 pixels, for 2 splats/);
  await assert.rejects(wd.readSplats(good), /a \.sog needs decodeImage/);
}

test('lookAt puts the target on -Z; reverse-Z maps near to 1', () => {
  const view = wd.mat4LookAt(new Float32Array(16), [0, 0, 5], [0, 0, 0], [0, 1, 0]);
  assert.ok(Math.abs(view[14] + 5) < 1e-6);
  const p = wd.mat4PerspectiveReverseZInfinite(new Float32Array(16), Math.PI / 3, 16 / 9, 0.1);
  const z = p[10] * -0.1 + p[14], w = p[11] * -0.1 + p[15];
  assert.ok(Math.abs(z / w - 1) < 1e-6);
});

test('the per-cloud uniform is the JS oracle\'s and the C library\'s, bit for bit', () => {
  const addon = wd.native();
  for (let t = 0; t < 300; t++) {
    const eye = [rand() * 40 - 20, rand() * 40 - 20, rand() * 40 - 20];
    const view = wd.mat4LookAt(new Float32Array(16), eye, [0, 0, 0], [0, 1, 0]);
    const proj = wd.mat4PerspectiveReverseZInfinite(new Float32Array(16), 0.3 + rand() * 1.7, 0.5 + rand() * 2, 0.1);
    const world = new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, rand() * 4 - 2, rand() * 4 - 2, rand() * 4 - 2, 1]);
    const min = new Float32Array([rand() * -5, rand() * -5, rand() * -5]);
    const max = new Float32Array([min[0] + rand() * 5, min[1] + rand() * 5, min[2] + rand() * 5]);
    const cull = [0, 0.5, 1][t % 3];
    const js = wd.cloudUniform(new Float32Array(37), view, world, proj, 1280, 720, min, max, cull);
    const c = addon.viewPrepare(view, world, proj, 1280, 720, min, max, cull);
    const o = oracle.prepare(view, world, proj, 1280, 720, min, max, cull);
    assert.strictEqual(js === null, c === null);
    assert.strictEqual(js === null, o === null);
    if (js !== null) {
      assert.deepStrictEqual(Array.from(js), Array.from(c));
      assert.deepStrictEqual(Array.from(js), Array.from(o));
    }
  }
});

// A fake addon: records what SplatPass asks of the GPU.
function fakeGpu() {
  const calls = { sorts: 0, states: 0, destroyed: 0 };
  const addon = {
    open: () => ({ ctx: true }),
    close() {},
    sync() {},
    cloudCreate: (ctx, records) => ({ handle: { cloud: true }, count: records.length / 32, min: new Float32Array([-1, -1, -1]), max: new Float32Array([1, 1, 1]) }),
    stateCreate: () => { calls.states++; return { state: true }; },
    destroy: () => { calls.destroyed++; },
    sort: () => { calls.sorts++; },
    stateRead: () => ({ order: new Uint32Array(0), count: 0, recomputed: 0 }),
    viewPrepare: () => null,
  };
  return { gpu: new wd.Gpu(0, addon), calls };
}

test('a frame where nothing moved sorts nothing; a move, a new cull or a new cloud sorts again', () => {
  const { gpu, calls } = fakeGpu();
  const splats = new wd.Splats(gpu, wd.parseSplats(new Uint8Array(64)));
  const scene = new wd.SplatScene();
  scene.addSplats({ splats });
  const cam = new wd.Camera({ fovY: 0.8 }).update(160 / 120);
  const pass = new wd.SplatPass(gpu);
  assert.strictEqual(pass.prepare(scene, cam, null, 160, 120), 1);
  assert.strictEqual(pass.sorted, 1);
  assert.strictEqual(pass.sort(), 1);
  pass.prepare(scene, cam, null, 160, 120);
  assert.strictEqual(pass.sorted, 0);
  assert.strictEqual(pass.sort(), 0);
  pass.cull = 0;
  pass.prepare(scene, cam, null, 160, 120);
  assert.strictEqual(pass.sorted, 1);
  pass.sort();
  cam.position[0] = 0.5;
  cam.update(160 / 120);
  pass.prepare(scene, cam, null, 160, 120);
  assert.strictEqual(pass.sorted, 1);
  assert.strictEqual(calls.sorts, 2);
  assert.strictEqual(calls.states, 1);
});

test('under TAA the sort is decided by the unjittered projection', () => {
  const { gpu } = fakeGpu();
  const splats = new wd.Splats(gpu, wd.parseSplats(new Uint8Array(64)));
  const scene = new wd.SplatScene();
  scene.addSplats({ splats });
  const cam = new wd.Camera().update(1);
  const pass = new wd.SplatPass(gpu);
  pass.unjittered = new Float32Array(cam.projection);
  pass.prepare(scene, cam, null, 100, 100);
  pass.sort();
  for (let i = 0; i < 5; i++) {
    cam.projection[8] = (i + 1) * 1e-3;   // the jitter moves, the camera does not
    pass.prepare(scene, cam, null, 100, 100);
    assert.strictEqual(pass.sorted, 0);
  }
});

test('clouds are ordered far first; removed or long-unseen clouds give their buffers back', () => {
  const { gpu, calls } = fakeGpu();
  const near = new wd.Splats(gpu, wd.parseSplats(new Uint8Array(64)));
  const far = new wd.Splats(gpu, wd.parseSplats(new Uint8Array(64)));
  const scene = new wd.SplatScene();
  const a = scene.addSplats({ splats: near, position: [0, 0, 2] });
  const b = scene.addSplats({ splats: far, position: [0, 0, -8] });
  const cam = new wd.Camera().update(1);
  const pass = new wd.SplatPass(gpu);
  pass.prepare(scene, cam, null, 100, 100);
  assert.deepStrictEqual(pass.clouds().map((c) => c.entity), [b, a]);
  scene.remove(a);
  pass.prepare(scene, cam, null, 100, 100);
  assert.strictEqual(calls.destroyed, 1);
  const other = new wd.SplatScene();
  for (let f = 0; f < wd.KEEP_FRAMES + 1; f++) pass.prepare(other, cam, null, 100, 100);
  assert.strictEqual(calls.destroyed, 2);
});

test('refusals: a negative cull, unloaded splats, a cloud wholly behind is skipped', () => {
  const { gpu } = fakeGpu();
  const splats = new wd.Splats(gpu, wd.parseSplats(new Uint8Array(64)));
  const scene = new wd.SplatScene();
  scene.addSplats({ splats });
  const cam = new wd.Camera().update(1);
  const pass = new wd.SplatPass(gpu);
  pass.cull = -1;
  assert.throws(() => pass.prepare(scene, cam, null, 100, 100), /splatCull must be 0 or more, got -1/);
  pass.cull = 0.5;
  cam.target.set([0, 0, 20]);   // looking away: the box is behind
  cam.update(1);
  assert.strictEqual(pass.prepare(scene, cam, null, 100, 100), 0);
  splats.destroy();
  assert.throws(() => pass.prepare(scene, cam, null, 100, 100), /these splats were unloaded/);
  assert.throws(() => scene.addSplats({ splats: {} }), /what new Splats returned/);
});

zipChecks().then(() => {
  passed++;
  console.log('  ok  a .sog\'s zip: stored, deflated, a comment; every refusal by its message');
  console.log(passed + ' checks passed');
}).catch((error) => {
  console.error(error);
  process.exit(1);
});
