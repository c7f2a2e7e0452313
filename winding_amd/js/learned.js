'use strict';
// The learned depth-CDF predictor and the learned sort, from JavaScript. Neither has a counterpart
// in the reference: they sort views of one cloud with one read of its records, in bins the
// predictor places, and give the exact order whatever it predicts (a failed check re-sorts the
// view on the exact path, on the GPU, in the same enqueue).
const { native, CLOUD_FLOATS } = require('./splats.js');

const POSE_FLOATS = 12;
const BINS = 256;

/** 12 -> 128 -> 128 -> 256 on the matrix cores; its softmax is 256 equal-mass bin widths in key space. */
class Predictor {
  constructor(gpu, seed) {
    this.gpu = gpu;
    this.native = gpu.native || native();
    this.handle = this.native.mlpCreate(gpu.handle, seed === undefined ? 0 : seed);
  }

  /** Adam on the GPU; poses n x 12, targets n x 256 (Float32Array). Returns the last epoch's mean loss. */
  train(poses, targets, options) {
    const o = Object.assign({ epochs: 40, batch: 256, lr: 2e-3, seed: 0 }, options || {});
    return this.native.mlpTrain(this.gpu.handle, this.handle, poses, targets, o.epochs, o.batch, o.lr, o.seed);
  }

  /** n x 256 predicted widths. */
  predict(poses) {
    return this.native.mlpPredict(this.gpu.handle, this.handle, poses);
  }

  save() {
    return this.native.mlpSave(this.handle);
  }

  load(bytes) {
    this.native.mlpLoad(this.handle, bytes);
  }

  destroy() {
    if (this.handle !== null) this.native.destroy(this.handle);
    this.handle = null;
  }
}

/** The predictor's 12 inputs for a 37-float cloud uniform. */
function poseFeatures(uniform) {
  return native().poseFeatures(uniform);
}

/**
 * Training data from the exact sort itself: for each uniform, one exact sort of `splats` and the
 * 256 equal-mass widths of the keys it produced. Returns { poses, targets } (Float32Arrays).
 */
function trainingSet(gpu, splats, uniforms) {
  const nat = gpu.native || native();
  const state = nat.stateCreate(gpu.handle, Math.max(1, splats.count));
  const poses = new Float32Array(uniforms.length * POSE_FLOATS);
  const targets = new Float32Array(uniforms.length * BINS);
  try {
    uniforms.forEach((u, k) => {
      nat.sort(gpu.handle, splats.handle, u, state);
      const edges = nat.stateQuantiles(gpu.handle, state);
      poses.set(nat.poseFeatures(u), k * POSE_FLOATS);
      for (let b = 0; b < BINS; b++) targets[k * BINS + b] = (edges[b + 1] - edges[b]) / 65536;
    });
  } finally {
    nat.destroy(state);
  }
  return { poses, targets };
}

/** Up to maxViews (1..8) views of one cloud a run, each into its own sort state. */
class LearnedSort {
  constructor(gpu, predictor, maxViews) {
    this.gpu = gpu;
    this.native = gpu.native || native();
    this.predictor = predictor;
    this.maxViews = maxViews || 8;
    this.handle = this.native.learnedCreate(gpu.handle, this.maxViews);
    this.states = [];
    this.capacity = 0;
  }

  /** Sorts `splats` for each uniform (Float32Array(37) each). Enqueued, not waited for. */
  sort(splats, uniforms, cap) {
    if (uniforms.length > this.maxViews) throw new Error('learned: ' + uniforms.length + ' views; this one takes 1 to ' + this.maxViews);
    for (const u of uniforms)
      if (!(u instanceof Float32Array) || u.length < CLOUD_FLOATS) throw new TypeError('learned: a uniform is 37 floats');
    if (this.capacity < splats.count) {
      for (const s of this.states) this.native.destroy(s);
      this.states = [];
      this.capacity = Math.max(1, splats.count);
    }
    while (this.states.length < uniforms.length) this.states.push(this.native.stateCreate(this.gpu.handle, this.capacity));
    this.native.learnedRun(this.gpu.handle, this.handle, this.predictor.handle, splats.handle, uniforms,
      this.states.slice(0, uniforms.length), cap || 0);
  }

  /** Waits and reads view k of the last run: { order, count, recomputed, fallback, cov2d? }. */
  read(k, options) {
    const wantCov = options && options.cov2d ? this.capacity : 0;
    return this.native.stateRead(this.gpu.handle, this.states[k], wantCov);
  }

  destroy() {
    for (const s of this.states) this.native.destroy(s);
    this.states = [];
    if (this.handle !== null) this.native.destroy(this.handle);
    this.handle = null;
  }
}

module.exports = { POSE_FLOATS, BINS, Predictor, LearnedSort, poseFeatures, trainingSet };
