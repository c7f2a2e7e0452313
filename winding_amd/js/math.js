'use strict';
// The matrix conventions the splat path is fed with: column-major m[col * 4 + row], column vectors
// (src/core/math/mat4.js:1-12). Maths in doubles, stored into Float32Arrays, as the reference does.

function mat4Create() {
  const m = new Float32Array(16);
  m[0] = m[5] = m[10] = m[15] = 1;
  return m;
}

/** out = a * b (src/core/math/mat4.js:96-126). */
function mat4Multiply(out, a, b) {
  const r = new Float64Array(16);
  for (let c = 0; c < 4; c++) {
    const b0 = b[c * 4], b1 = b[c * 4 + 1], b2 = b[c * 4 + 2], b3 = b[c * 4 + 3];
    for (let row = 0; row < 4; row++) r[c * 4 + row] = b0 * a[row] + b1 * a[4 + row] + b2 * a[8 + row] + b3 * a[12 + row];
  }
  out.set(r);
  return out;
}

/** World -> view; right-handed, the camera looks down its own -Z (src/core/math/mat4.js:243-273). */
function mat4LookAt(out, eye, center, up) {
  let zx = eye[0] - center[0], zy = eye[1] - center[1], zz = eye[2] - center[2];
  const zLen = Math.sqrt(zx * zx + zy * zy + zz * zz);
  if (!(zLen > 1e-6)) throw new Error('mat4LookAt: eye and center coincide');
  zx /= zLen; zy /= zLen; zz /= zLen;
  let xx = up[1] * zz - up[2] * zy, xy = up[2] * zx - up[0] * zz, xz = up[0] * zy - up[1] * zx;
  const xLen = Math.sqrt(xx * xx + xy * xy + xz * xz);
  if (!(xLen > 1e-6)) throw new Error('mat4LookAt: up is parallel to the view direction');
  xx /= xLen; xy /= xLen; xz /= xLen;
  const yx = zy * xz - zz * xy, yy = zz * xx - zx * xz, yz = zx * xy - zy * xx;
  out[0] = xx; out[1] = yx; out[2] = zx; out[3] = 0;
  out[4] = xy; out[5] = yy; out[6] = zy; out[7] = 0;
  out[8] = xz; out[9] = yz; out[10] = zz; out[11] = 0;
  out[12] = -(xx * eye[0] + xy * eye[1] + xz * eye[2]);
  out[13] = -(yx * eye[0] + yy * eye[1] + yz * eye[2]);
  out[14] = -(zx * eye[0] + zy * eye[1] + zz * eye[2]);
  out[15] = 1;
  return out;
}

/** Reverse-Z, infinite far plane, WebGPU clip space (src/core/math/mat4.js:455-468). */
function mat4PerspectiveReverseZInfinite(out, fovYRadians, aspect, near) {
  const f = 1 / Math.tan(fovYRadians * 0.5);
  out.fill(0);
  out[0] = f / aspect;
  out[5] = f;
  out[11] = -1;
  out[14] = near;
  return out;
}

/** Reverse-Z orthographic box (src/core/math/mat4.js:408-426). */
function mat4OrthographicReverseZ(out, left, right, bottom, top, near, far) {
  const invWidth = 1 / (right - left), invHeight = 1 / (top - bottom), invDepth = 1 / (far - near);
  out.fill(0);
  out[0] = 2 * invWidth;
  out[5] = 2 * invHeight;
  out[10] = invDepth;
  out[12] = -(right + left) * invWidth;
  out[13] = -(top + bottom) * invHeight;
  out[14] = far * invDepth;
  out[15] = 1;
  return out;
}

/**
 * The part of Camera (src/scene/camera.js:76-115,258-278) a sort reads: position, target, up,
 * fovY, near, orthographic; update(aspect) writes view and projection.
 */
class Camera {
  constructor(options) {
    const o = options || {};
    this.fovY = o.fovY === undefined ? Math.PI / 3 : o.fovY;
    this.near = o.near === undefined ? 0.1 : o.near;
    this.orthographic = !!o.orthographic;
    if (this.orthographic) this.far = o.far === undefined ? 1000 : o.far;
    this.position = new Float32Array([0, 0, 5]);
    this.target = new Float32Array([0, 0, 0]);
    this.up = new Float32Array([0, 1, 0]);
    this.view = mat4Create();
    this.projection = mat4Create();
    this.aspect = 1;
  }

  update(aspect) {
    this.aspect = aspect;
    mat4LookAt(this.view, this.position, this.target, this.up);
    if (this.orthographic) {
      const dx = this.position[0] - this.target[0], dy = this.position[1] - this.target[1], dz = this.position[2] - this.target[2];
      const halfHeight = Math.sqrt(dx * dx + dy * dy + dz * dz) * Math.tan(this.fovY * 0.5);
      const halfWidth = halfHeight * aspect;
      mat4OrthographicReverseZ(this.projection, -halfWidth, halfWidth, -halfHeight, halfHeight, this.near, this.far);
    } else {
      mat4PerspectiveReverseZInfinite(this.projection, this.fovY, aspect, this.near);
    }
    return this;
  }
}

module.exports = { mat4Create, mat4Multiply, mat4LookAt, mat4PerspectiveReverseZInfinite, mat4OrthographicReverseZ, Camera };
