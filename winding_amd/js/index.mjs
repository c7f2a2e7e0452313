// ES-module entry of the splat path's JavaScript API, as the reference's modules are imported.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const api = require('./index.js');

export const {
  SPLAT_BYTES, BUCKETS, SPLAT_CULL, CLOUD_FLOATS, KEEP_FRAMES, SH_COEFFICIENTS, shWords,
  native, Gpu, parseSplats, readSplats, Splats, SplatScene, SplatPass, cloudUniform,
  mat4Create, mat4Multiply, mat4LookAt, mat4PerspectiveReverseZInfinite, mat4OrthographicReverseZ, Camera,
  POSE_FLOATS, BINS, Predictor, LearnedSort, poseFeatures, trainingSet,
} = api;
export default api;
