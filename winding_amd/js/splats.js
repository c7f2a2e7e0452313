'use strict';
// Gaussian splats on an MI355X: the splat module's JavaScript API over the HIP path.
//
// The shapes follow the reference (Winding v1.9.0):
//   parseSplats(bytes)             src/scene/splats.js:60-76,373-388 (the .splat path)
//   new Splats(gpu, data) / destroy src/render/splats.js:389-416
//   SplatPass.prepare(...)          src/render/splats.js:546-618
//   SplatPass sort                  _encodeSort, src/render/splats.js:676-697
// What runs where differs: the records go to the GPU as they are (32 bytes a splat) and the
// covariance put() worked out on the CPU is worked out there, in the projection kernel; the
// sort is stable (key, then splat index); the survivors' order, count and footprints can be
// read back into typed arrays, or left on the GPU for whatever draws them.
const path = require('path');

const SPLAT_BYTES = 32;
/** Coefficients a colour channel has past the constant one, by degree (src/scene/splats.js:43-47). */
const SH_COEFFICIENTS = [0, 3, 8, 15];
/** 32-bit words a splat's harmonics take: three channels of half floats, packed. */
const shWords = (degree) => Math.ceil(SH_COEFFICIENTS[degree] * 3 / 2);
const BUCKETS = 65536;
/** Opacity x pixels a splat must add to be drawn (src/render/splats.js:78). */
const SPLAT_CULL = 0.5;
/** modelView, projection, viewport, depth range, cull: what a sort depends on. */
const CLOUD_FLOATS = 37;
/** Frames a node's sort buffers outlive the last frame that drew it (src/render/splats.js:70). */
const KEEP_FRAMES = 120;

let nativeAddon = null;
/** The N-API addon; throws when it is not built (there is no CPU fallback). */
function native() {
  if (nativeAddon === null) {
    const file = path.join(__dirname, '..', 'winding_amd.node');
    try {
      nativeAddon = require(file);
    } catch (error) {
      throw new Error('winding_amd: the addon ' + file + ' is not built (python -m winding_amd.build): ' + error.message);
    }
  }
  return nativeAddon;
}

/** The GPU: one HIP device and its stream, the reference's rhi for this path. */
class Gpu {
  constructor(index, addon) {
    this.native = addon || native();
    this.handle = this.native.open(index === undefined ? 0 : index);
  }

  static open(index) {
    return new Gpu(index);
  }

  sync() {
    this.native.sync(this.handle);
  }

  destroy() {
    if (this.handle !== null) this.native.close(this.handle);
    this.handle = null;
  }
}

const isZip = (d) => d[0] === 0x50 && d[1] === 0x4b && d[2] === 0x03 && d[3] === 0x04;

/**
 * A capture's splats from the bytes of a .splat or .ply file (parseSplats, src/scene/splats.js:58-80).
 * A .splat gives { count, degree: 0, records, min: null, max: null }: the bytes themselves, 32 a
 * splat. A .ply gives { format: 'ply', bytes }: it is decoded on the GPU by new Splats. Either way
 * min and max are filled in when the capture is uploaded, where the GPU checks every position.
 */
function parseSplats(bytes) {
  const data = bytes instanceof Uint8Array ? bytes : new Uint8Array(bytes);
  if (data.length >= 8 && data[0] === 0x4e && data[1] === 0x47 && data[2] === 0x53 && data[3] === 0x50
    && new DataView(data.buffer, data.byteOffset, 8).getUint32(4, true) >= 4) {
    throw new Error('splats: a .spz of version 4, compressed with zstd, which browsers cannot unpack; '
      + 'save it as version 3, or as a .ply or .splat');
  }
  const isPly = data.length >= 4 && data[0] === 0x70 && data[1] === 0x6c && data[2] === 0x79 && (data[3] === 0x0a || data[3] === 0x0d);
  if (isPly) return { format: 'ply', bytes: data, count: null, degree: null, min: null, max: null };
  if (isZip(data)) throw new Error('splats: this is a zip: unpack it with readSplats, which reads a .sog');
  if (data[0] === 0x1f && data[1] === 0x8b) throw new Error('splats: this is gzipped: unpack it with readSplats, which reads a .spz');
  if (data.length === 0 || data.length % SPLAT_BYTES !== 0) {
    throw new Error('splats: not a .ply, and ' + data.length + ' bytes is not a whole number of 32-byte .splat records');
  }
  return { count: data.length / SPLAT_BYTES, degree: 0, records: data, min: null, max: null };
}

// ---- zip, as a .sog packs its meta.json and images -------------------------------------------
// Written from the format's specification (PKWARE's APPNOTE.TXT, sections 4.3.7, 4.3.12 and 4.3.16):
// the end-of-central-directory record is found by scanning back over at most a 65,535-byte comment,
// the central directory is walked entry by entry, and each entry's data is found behind its local
// header. Entries stored (method 0) or deflated (method 8) are read, with their size and CRC-32 checked;
// zip64 archives are not. The refusals carry the reference's messages (src/scene/splats.js:97-125).

/** Little-endian fields of the three records read: name -> [offset in the record, bytes]. */
const ZIP_RECORDS = {
  end: { magic: 0x06054b50, bytes: 22, fields: { entries: [10, 2], dirOffset: [16, 4] } },
  central: {
    magic: 0x02014b50, bytes: 46,
    fields: { method: [10, 2], crc: [16, 4], packed: [20, 4], size: [24, 4], nameBytes: [28, 2],
      extraBytes: [30, 2], commentBytes: [32, 2], localOffset: [42, 4] },
  },
  local: { magic: 0x04034b50, bytes: 30, fields: { nameBytes: [26, 2], extraBytes: [28, 2] } },
};

let crcTable = null;
/** CRC-32 (the IEEE polynomial, reflected: 0xEDB88320), byte at a time through a 256-entry table. */
function crc32(bytes) {
  if (crcTable === null) {
    crcTable = new Int32Array(256);
    for (let b = 0; b < 256; b++) {
      let r = b;
      for (let k = 0; k < 8; k++) r = (r & 1) ? (0xedb88320 ^ (r >>> 1)) : (r >>> 1);
      crcTable[b] = r;
    }
  }
  let c = -1;
  for (let i = 0; i < bytes.length; i++) c = crcTable[(c ^ bytes[i]) & 0xff] ^ (c >>> 8);
  return (c ^ -1) >>> 0;
}

class ZipArchive {
  constructor(bytes) {
    this.bytes = bytes;
    this.view = new DataView(bytes.buffer, bytes.byteOffset, bytes.byteLength);
    this.entries = new Map();
    const end = this._findEnd();
    if (end === null) throw new Error('splats: a zip with no end record: cut short?');
    let at = end.dirOffset;
    for (let k = 0; k < end.entries; k++) {
      const e = this._record('central', at);
      if (e === null) throw new Error('splats: a zip whose directory is damaged');
      const name = Buffer.from(bytes.subarray(at + 46, at + 46 + e.nameBytes)).toString('utf8');
      if (e.packed === 0xffffffff || e.size === 0xffffffff || e.localOffset === 0xffffffff) {
        throw new Error('splats: a zip64 .sog is not read');
      }
      this.entries.set(name, e);
      at += 46 + e.nameBytes + e.extraBytes + e.commentBytes;
    }
  }

  /** The record of `kind` at `at`, its fields read, or null when its magic is not there. */
  _record(kind, at) {
    const r = ZIP_RECORDS[kind];
    if (at < 0 || at + r.bytes > this.bytes.length || this.view.getUint32(at, true) !== r.magic) return null;
    const out = {};
    for (const name of Object.keys(r.fields)) {
      const [offset, size] = r.fields[name];
      out[name] = size === 4 ? this.view.getUint32(at + offset, true) : this.view.getUint16(at + offset, true);
    }
    return out;
  }

  /** The end record: the last 22 bytes, or earlier by the length of a trailing comment. */
  _findEnd() {
    const last = this.bytes.length - ZIP_RECORDS.end.bytes;
    for (let at = last; at >= 0 && at >= last - 0xffff; at--) {
      const r = this._record('end', at);
      if (r !== null) return r;
    }
    return null;
  }

  has(name) {
    return this.entries.has(name);
  }

  /** An entry's bytes, inflated when deflated, checked against the directory's size and CRC-32. */
  read(name) {
    const e = this.entries.get(name);
    const local = this._record('local', e.localOffset);
    const start = local === null ? -1 : e.localOffset + 30 + local.nameBytes + local.extraBytes;
    if (start < 0 || start + e.packed > this.bytes.length) throw new Error('splats: the zip is cut short in ' + name);
    if (e.method !== 0 && e.method !== 8) {
      throw new Error('splats: ' + name + ' is compressed with method ' + e.method + '; stored and deflate are read');
    }
    const packed = this.bytes.subarray(start, start + e.packed);
    const data = e.method === 0 ? packed : new Uint8Array(require('zlib').inflateRawSync(packed));
    if (data.length !== e.size || crc32(data) !== e.crc) throw new Error('splats: ' + name + ' in the zip is damaged');
    return data;
  }
}

/**
 * As parseSplats, for every format the reference's readSplats takes (src/scene/splats.js:84-92): a
 * .spz is handed to the GPU as it is (the library gunzips it); a .sog is unzipped here, its
 * meta.json read and its images decoded by decodeImage(bytes) -> { width, height, rgba }, and the
 * splats decoded on the GPU by new Splats.
 */
async function readSplats(bytes, options) {
  const data = bytes instanceof Uint8Array ? bytes : new Uint8Array(bytes);
  if (isZip(data)) return readSog(data, options && options.decodeImage);
  if (data[0] === 0x1f && data[1] === 0x8b) return { format: 'spz', bytes: data, count: null, degree: null, min: null, max: null };
  return parseSplats(data);
}

/**
 * A .sog's host part: its meta.json and its images, decoded by the caller's decodeImage(bytes) ->
 * { width, height, rgba }, in the shape the GPU decoder takes (wd_sog). What meta.json fixes and what
 * is refused follow the reference's readSog (src/scene/splats.js:128-160), with its messages; the
 * splats themselves are decoded on the GPU by new Splats.
 */
async function readSog(data, decodeImage) {
  if (typeof decodeImage !== 'function') throw new Error('splats: a .sog needs decodeImage, for its WebP images');
  const zip = new ZipArchive(data);
  const bytesOf = (name) => {
    if (!zip.has(name)) throw new Error('splats: the .sog has no ' + name);
    return zip.read(name);
  };
  const meta = JSON.parse(Buffer.from(bytesOf('meta.json')).toString('utf8'));
  if (meta.version !== 2) {
    const v = meta.version === undefined ? 1 : meta.version;
    throw new Error('splats: a .sog of version ' + v + '; version 2 is read (re-save it with splat-transform)');
  }
  const count = meta.count;
  if (!(Number.isInteger(count) && count > 0)) throw new Error('splats: a .sog of ' + count + ' splats');
  const bands = meta.shN ? meta.shN.bands : 0;
  const degree = SH_COEFFICIENTS[bands] > 0 ? bands : 0;
  // [wd_sog image, file in the zip, a pixel a splat?]; the harmonic palette is as wide as its degree makes it
  const wanted = [
    ['means_l', meta.means.files[0], true], ['means_u', meta.means.files[1], true], ['quats', meta.quats.files[0], true],
    ['scales', meta.scales.files[0], true], ['sh0', meta.sh0.files[0], true],
  ];
  if (degree > 0) wanted.push(['centroids', meta.shN.files[0], false], ['labels', meta.shN.files[1], true]);
  const decoded = await Promise.all(wanted.map(async ([, file, perSplat]) => {
    const img = await decodeImage(bytesOf(file));
    const pixels = img.width * img.height;
    if (perSplat && pixels < count) throw new Error('splats: ' + file + ' has ' + pixels + ' pixels, for ' + count + ' splats');
    return { width: img.width, height: img.height, rgba: Uint8Array.from(img.rgba) };
  }));
  const images = {};
  wanted.forEach(([key], k) => { images[key] = decoded[k]; });
  const sog = {
    count, bands: degree, images,
    mins: Float64Array.from(meta.means.mins), maxs: Float64Array.from(meta.means.maxs),
    scaleCodebook: Float64Array.from(meta.scales.codebook), colourCodebook: Float64Array.from(meta.sh0.codebook),
  };
  if (degree > 0) {
    sog.shCount = meta.shN.count === undefined ? 0xffffffff : meta.shN.count;
    sog.shCodebook = Float64Array.from(meta.shN.codebook);
  }
  return { format: 'sog', sog, count: null, degree: null, min: null, max: null };
}

/**
 * A capture on the GPU. `count` splats; `min` and `max`, the corners of the box around
 * their centres. Takes what parseSplats or readSplats give (a .ply, .spz or .sog is decoded on
 * the GPU), or the reference's parsed data { count, degree, centers, covariances, sh }. Refused
 * by name when a position is not a number or the capture is larger than one sort takes.
 */
class Splats {
  constructor(gpu, data, label) {
    const native = gpu.native;
    let made;
    let degree = data.degree || 0;
    if (data.format === 'ply') made = native.decodePly(gpu.handle, data.bytes);
    else if (data.format === 'spz') made = native.decodeSpz(gpu.handle, data.bytes);
    else if (data.format === 'sog') made = native.decodeSog(gpu.handle, data.sog);
    if (made !== undefined) {
      degree = made.degree;
    } else {
      if (!(degree >= 0 && degree <= 3)) throw new Error('splats: harmonics of degree ' + degree + '; this path takes 0 to 3');
      if (degree > 0 && !(data.sh instanceof Uint32Array)) throw new Error('splats: harmonics of degree ' + degree + ' need data.sh (Uint32Array)');
      made = data.centers instanceof Float32Array
        ? native.cloudCreateSplats(gpu.handle, data.centers, data.covariances)
        : native.cloudCreate(gpu.handle, data.records);
      if (degree > 0) {
        try {
          native.cloudSH(gpu.handle, made.handle, data.sh, degree);
        } catch (error) {
          native.destroy(made.handle);
          throw error;
        }
      }
    }
    this.gpu = gpu;
    this.label = label === undefined ? 'splats' : label;
    this.handle = made.handle;
    this.count = made.count;
    this.degree = degree;
    /** Given as centres and covariances (decoded, or the reference's parsed data), not .splat records. */
    this.sigma = data.format !== undefined || data.centers instanceof Float32Array;
    this.min = made.min;
    this.max = made.max;
    this.unloaded = false;
  }

  destroy() {
    if (!this.unloaded) this.gpu.native.destroy(this.handle);
    this.unloaded = true;
  }
}

/** The scene records prepare iterates: entity -> { splats, world }, as scene.splats holds them. */
class SplatScene {
  constructor() {
    this.splats = new Map();
    this.hidden = null;
    this._next = 1;
  }

  addSplats(options) {
    const o = options || {};
    const splats = o.splats;
    if (!splats || typeof splats.count !== 'number' || !splats.handle) {
      throw new Error('addSplats: splats must be what new Splats returned');
    }
    const world = new Float32Array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]);
    if (o.matrix) world.set(o.matrix);
    if (o.position) { world[12] = o.position[0]; world[13] = o.position[1]; world[14] = o.position[2]; }
    const entity = this._next++;
    this.splats.set(entity, { splats, world });
    return entity;
  }

  splatsOf(entity) {
    const r = this.splats.get(entity);
    return r === undefined ? null : r;
  }

  remove(entity) {
    this.splats.delete(entity);
  }
}

function sameFloats(a, b, n) {
  for (let i = 0; i < n; i++) if (a[i] !== b[i]) return false;
  return true;
}

/**
 * The per-cloud uniform prepare() writes (src/render/splats.js:559-587): modelView = view x world,
 * the depth the box's eight corners span, depthMin and depthScale; the first 37 floats of Cloud.
 * Null when the cloud is wholly behind the camera.
 */
function cloudUniform(out, view, world, projection, width, height, min, max, cull) {
  const mv = out.subarray(0, 16);
  for (let c = 0; c < 4; c++) {
    const b0 = world[c * 4], b1 = world[c * 4 + 1], b2 = world[c * 4 + 2], b3 = world[c * 4 + 3];
    for (let r = 0; r < 4; r++) mv[c * 4 + r] = b0 * view[r] + b1 * view[4 + r] + b2 * view[8 + r] + b3 * view[12 + r];
  }
  let near = Infinity;
  let far = -Infinity;
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
  out.set(projection, 16);
  out[32] = width;
  out[33] = height;
  out[34] = near;
  out[35] = (BUCKETS - 1) / Math.max(far - near, 1e-6);
  out[36] = cull;
  return out;
}

/** The reverse-Z infinite perspective over an affine model-view, the uniform's floats 0-31. */
function reverseZ(u) {
  return u[17] === 0 && u[18] === 0 && u[19] === 0 && u[20] === 0 && u[22] === 0 && u[23] === 0 && u[26] === 0
    && u[27] === -1 && u[28] === 0 && u[29] === 0 && u[31] === 0 && u[3] === 0 && u[7] === 0 && u[11] === 0 && u[15] === 1;
}

class SplatPass {
  constructor(gpu) {
    this.gpu = gpu;
    this.native = gpu.native;
    /** Each placed cloud's sort buffers, by its scene record. */
    this._states = new Map();
    this._items = [];
    this._frame = 0;
    /** Splats in the clouds the last frame drew, before culling. */
    this.count = 0;
    /** Clouds the last prepare marked for sorting; the rest kept the order they had. */
    this.sorted = 0;
    /** The least a splat may add to the picture, opacity x pixels, and be drawn. */
    this.cull = SPLAT_CULL;
    /**
     * Whether each sort also makes the draw's ellipse set-up beside cov2d (what the splat vertex
     * shader works out per instance, src/render/splats.js:263-291); read(entity, { ellipses: true }).
     */
    this.ellipses = false;
    /** The projection without TAA's jitter, when jittered: what the sort is decided by. */
    this.unjittered = null;
    this._key = new Float32Array(CLOUD_FLOATS);
  }

  /**
   * This frame's clouds: each one whose camera, transform or canvas moved has its uniform
   * written, to be sorted again by sort(). Returns how many clouds there are.
   */
  prepare(scene, camera, environment, width, height) {
    const cull = this.cull;
    if (!(cull >= 0 && Number.isFinite(cull))) throw new Error('splatCull must be 0 or more, got ' + cull);
    this._frame++;
    const items = this._items;
    items.length = 0;
    this.count = 0;
    this.sorted = 0;
    for (const [entity, record] of scene.splats) {
      if (scene.hidden && scene.hidden[entity] === 1) continue;
      const splats = record.splats;
      if (splats.unloaded) throw new Error('addSplats: these splats were unloaded; remove the node first');
      const state = this._stateFor(record, splats, scene, entity);
      state.lastFrame = this._frame;
      const f = cloudUniform(state.uniform, camera.view, record.world, camera.projection, width, height, splats.min, splats.max, cull);
      if (f === null) continue;   // wholly behind the camera
      const key = this._key;
      key.set(f.subarray(0, CLOUD_FLOATS));
      if (this.unjittered !== null) key.set(this.unjittered, 16);
      const ellipses = this.ellipses === true;
      if (state.ellipses !== ellipses) this.native.stateEllipses(this.gpu.handle, state.handle, ellipses);
      const sort = !sameFloats(state.last, key, CLOUD_FLOATS) || state.cull !== cull || state.ellipses !== ellipses;
      state.cull = cull;
      state.ellipses = ellipses;
      if (sort) {
        state.next.set(key);
        this.sorted++;
      }
      const m = f;
      const cx = (splats.min[0] + splats.max[0]) / 2, cy = (splats.min[1] + splats.max[1]) / 2, cz = (splats.min[2] + splats.max[2]) / 2;
      items.push({ state, entity, sort, count: splats.count, depth: -(m[2] * cx + m[6] * cy + m[10] * cz + m[14]) });
      this.count += splats.count;
    }
    items.sort((a, b) => b.depth - a.depth);
    for (const [record, state] of this._states) {
      const removed = state.scene === scene && scene.splats.get(state.entity) !== record;
      if (removed || state.lastFrame < this._frame - KEEP_FRAMES) {
        this._destroyState(state);
        this._states.delete(record);
      }
    }
    return items.length;
  }

  /**
   * Sorts on the GPU every cloud prepare marked; returns how many. Enqueued, not waited for.
   * One cloud is one sort; more are one batch, every marked cloud in one graph launch (what
   * _encodeSort's loop of per-cloud dispatches becomes, src/render/splats.js:676-697).
   */
  sort() {
    const marked = this._items.filter((i) => i.sort);
    for (const item of marked) item.state.last.set(item.state.next);
    // A cloud with harmonics or of centres and covariances, or a state that makes the draw's
    // ellipses, has a sort of its own under a projection other than the reverse-Z perspective
    // (the only one a batch shades, takes them and makes ellipses under); the rest share one launch.
    const own = (i) => (i.state.ellipses || i.state.splats.degree > 0 || i.state.splats.sigma) && !reverseZ(i.state.uniform);
    const plain = marked.filter((i) => !own(i));
    for (const item of marked) {
      if (own(item)) {
        const state = item.state;
        this.native.sort(this.gpu.handle, state.splats.handle, state.uniform, state.handle);
      }
    }
    if (plain.length === 1) {
      const state = plain[0].state;
      this.native.sort(this.gpu.handle, state.splats.handle, state.uniform, state.handle);
    } else if (plain.length > 1) {
      if (this._batch === undefined || this._batchJobs < plain.length) {
        if (this._batch !== undefined) this.native.destroy(this._batch);
        this._batchJobs = Math.max(8, plain.length);
        this._batch = this.native.batchCreate(this.gpu.handle, this._batchJobs);
      }
      this.native.batchSet(this._batch, plain.map((i) => i.state.splats.handle),
        plain.map((i) => i.state.uniform), plain.map((i) => i.state.handle));
      this.native.batchRun(this.gpu.handle, this._batch);
    }
    for (const item of marked) item.state.sortedOnce = true;
    return marked.length;
  }

  /**
   * Times the batch's configurations on the clouds of the last sort (schedule, work item, tile)
   * and keeps the fastest for the sorts after it; returns the tuner's report, or null when the
   * last sort was not a batch. No counterpart in the reference, whose sort has one form.
   */
  tune(reps = 5) {
    if (this._batch === undefined) return null;
    return this.native.batchTune(this.gpu.handle, this._batch, reps);
  }

  /** The clouds of the last prepare, far first: [{ entity, count }]. */
  clouds() {
    return this._items.map((i) => ({ entity: i.entity, count: i.count }));
  }

  /**
   * Waits for entity's last sort and reads it back: { order (Uint32Array, back to front),
   * count, recomputed, cov2d (Float32Array, 4 a splat by index: xx, xy, yy, rgba bits),
   * ellipses (Float32Array, 8 a splat by index: clip x y z w, then along and across, scaled by
   * the reach, in pixels; zeros where nothing is drawn) }.
   */
  read(entity, options) {
    const item = this._items.find((i) => i.entity === entity);
    if (item === undefined || !item.state.sortedOnce) throw new Error('splats: entity ' + entity + ' has not been sorted');
    const wantCov = options && options.cov2d;
    const out = this.native.stateRead(this.gpu.handle, item.state.handle, wantCov ? item.count : 0);
    if (options && options.ellipses) out.ellipses = this.native.stateReadEllipses(this.gpu.handle, item.state.handle, item.count);
    return out;
  }

  _stateFor(record, splats, scene, entity) {
    let state = this._states.get(record);
    if (state !== undefined && state.splats === splats) return state;
    if (state !== undefined) this._destroyState(state);
    const last = new Float32Array(CLOUD_FLOATS);
    last.fill(NaN);   // NaN matches nothing, so the first frame sorts
    state = {
      splats, scene, entity, lastFrame: 0, last, next: new Float32Array(CLOUD_FLOATS),
      uniform: new Float32Array(CLOUD_FLOATS), cull: undefined, sortedOnce: false, ellipses: false,
      handle: this.native.stateCreate(this.gpu.handle, Math.max(1, splats.count)),
    };
    this._states.set(record, state);
    return state;
  }

  _destroyState(state) {
    if (state.handle !== null) this.native.destroy(state.handle);
    state.handle = null;
  }

  destroy() {
    if (this._batch !== undefined) this.native.destroy(this._batch);
    this._batch = undefined;
    for (const state of this._states.values()) this._destroyState(state);
    this._states.clear();
    this._items.length = 0;
  }
}

module.exports = {
  SPLAT_BYTES, BUCKETS, SPLAT_CULL, CLOUD_FLOATS, KEEP_FRAMES, SH_COEFFICIENTS, shWords,
  native, Gpu, parseSplats, readSplats, Splats, SplatScene, SplatPass, cloudUniform,
};
