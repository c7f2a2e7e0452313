"""Python binding of libwinding_amd.so (ctypes over include/winding_amd.h).

The host API the parity tests and the benchmark drive; the JavaScript module
under winding_amd/js drives the same library through the N-API addon. Loading
fails loudly when the library is not built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# WINDING_AMD_LIB: another build of the library (say, one made with other keys a thread)
LIB_PATH = os.environ.get("WINDING_AMD_LIB") or os.path.join(ROOT, "libwinding_amd.so")

WD_OK = 0
WD_RECORD_BYTES = 32
WD_BUCKETS = 65536
WD_SPLAT_CULL = 0.5
WD_MAX_SORT = 1 << 28
WD_MAX_NARROW = 1 << 24
WD_NOALLOC_MAX_SPLATS = 1 << 21
WD_COARSE_BINS = 256
WD_POSE_FLOATS = 12
WD_SH_COEFFICIENTS = (0, 3, 8, 15)
WD_E_ARG = -1
WD_E_DATA = -4

# Every symbol include/winding_amd.h declares, in header order.
EXPORTS = [
    "wd_last_error", "wd_open", "wd_close", "wd_sync", "wd_stream", "wd_capacity",
    "wd_cloud_create", "wd_cloud_create_device", "wd_cloud_info", "wd_cloud_records", "wd_cloud_destroy",
    "wd_cloud_sh", "wd_cloud_create_splats", "wd_cloud_read", "wd_decode_ply", "wd_decode_spz", "wd_decode_sog",
    "wd_view_prepare", "wd_state_create", "wd_state_destroy", "wd_state_buffers", "wd_sort", "wd_sort_noalloc",
    "wd_state_read", "wd_state_stats", "wd_state_read_keys", "wd_state_ellipses", "wd_state_ellipse_buffer",
    "wd_state_read_ellipses", "wd_state_keys_per_thread", "wd_pool_create", "wd_pool_device", "wd_pool_destroy",
    "wd_particles_sort",
    "wd_profile", "wd_profile_read", "wd_time_sorts",
    "wd_batch_create", "wd_batch_schedule", "wd_batch_configure", "wd_batch_tune", "wd_batch_set", "wd_batch_run",
    "wd_batch_stuck", "wd_batch_destroy",
    "wd_mlp_create", "wd_mlp_destroy", "wd_mlp_set", "wd_mlp_get", "wd_mlp_save", "wd_mlp_load",
    "wd_mlp_grad", "wd_mlp_train", "wd_mlp_predict", "wd_pose_features",
    "wd_learned_create", "wd_learned_run", "wd_learned_destroy", "wd_state_quantiles",
    "wd_synth_records", "wd_debug_canonical", "wd_debug_selftest", "wd_debug_copy", "wd_debug_batch_force_stuck",
]


class SogImage(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class Sog(C.Structure):
    """wd_sog: a .sog's meta.json numbers and its decoded images."""
    _fields_ = [("count", C.c_uint32), ("bands", C.c_uint32), ("sh_count", C.c_uint32),
                ("mins", C.c_double * 3), ("maxs", C.c_double * 3),
                ("scale_codebook", C.c_void_p), ("scale_codebook_size", C.c_uint32),
                ("colour_codebook", C.c_void_p), ("colour_codebook_size", C.c_uint32),
                ("sh_codebook", C.c_void_p), ("sh_codebook_size", C.c_uint32),
                ("means_l", SogImage), ("means_u", SogImage), ("quats", SogImage), ("scales", SogImage),
                ("sh0", SogImage), ("centroids", SogImage), ("labels", SogImage)]


class ParticleView(C.Structure):
    """wd_particle_view: the particle sort's View uniform (src/render/particles.js:171-174)."""
    _fields_ = [("eye", C.c_float * 4), ("forward", C.c_float * 4)]


class View(C.Structure):
    """wd_view: the first 37 floats of the reference's Cloud uniform (src/render/splats.js:80-92)."""
    _fields_ = [
        ("model_view", C.c_float * 16),
        ("projection", C.c_float * 16),
        ("viewport", C.c_float * 2),
        ("depth_min", C.c_float),
        ("depth_scale", C.c_float),
        ("cull", C.c_float),
    ]

    def as_array(self):
        return np.frombuffer(bytes(self), dtype=np.float32).copy()


class Job(C.Structure):
    """wd_job: one (cloud, view, state) of a batch."""
    _fields_ = [("cloud", C.c_void_p), ("view", View), ("state", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("count", C.c_uint32), ("recomputed", C.c_uint32), ("fallback", C.c_uint32),
                ("overflow_bins", C.c_uint32)]


WD_TUNE_CONFIGS = 6


class Tuned(C.Structure):
    """wd_tuned: one configuration the autotuner timed."""
    _fields_ = [("schedule", C.c_int), ("super_tiles", C.c_uint32), ("keys_per_thread", C.c_uint32),
                ("ms", C.c_double)]


class Tuning(C.Structure):
    """wd_tuning: the autotuner's choice (schedule, work item, keys a thread, tile, digits) and the mean ms a
    run of each configuration it timed."""
    _fields_ = [("schedule", C.c_int), ("super_tiles", C.c_uint32), ("keys_per_thread", C.c_uint32),
                ("tile", C.c_uint32), ("radix_bits", C.c_uint32), ("passes", C.c_uint32),
                ("timed", Tuned * WD_TUNE_CONFIGS)]

    def report(self):
        """The choice and every timed configuration, as plain data."""
        names = ["phased", "persistent"]
        return {"schedule": names[self.schedule], "super_tiles": self.super_tiles,
                "keys_per_thread": self.keys_per_thread, "tile": self.tile,
                "radix": "%d x %d-bit" % (self.passes, self.radix_bits),
                "timed_ms": {"%s/%d/kpt%d" % (names[t.schedule], t.super_tiles, t.keys_per_thread): t.ms
                             for t in self.timed}}


class WindingError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


_lib = None


def lib():
    """The loaded library; raises if it was not built (python -m winding_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("winding_amd: %s is not built; run python -m winding_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
        FP = C.POINTER(C.c_float)
        sig = {
            "wd_last_error": (C.c_char_p, []),
            "wd_open": (I, [I, C.POINTER(P)]),
            "wd_close": (I, [P]),
            "wd_sync": (I, [P]),
            "wd_stream": (I, [P, C.POINTER(P)]),
            "wd_capacity": (I, [P, U64, C.POINTER(U64)]),
            "wd_cloud_create": (I, [P, P, U64, C.POINTER(P), FP, FP]),
            "wd_cloud_create_device": (I, [P, P, U64, C.POINTER(P), FP, FP]),
            "wd_cloud_info": (I, [P, C.POINTER(U64), FP, FP]),
            "wd_cloud_records": (I, [P, C.POINTER(P)]),
            "wd_cloud_sh": (I, [P, P, P, U64, U32]),
            "wd_cloud_create_splats": (I, [P, P, P, U64, C.POINTER(P), P, P]),
            "wd_cloud_read": (I, [P, P, P, P, P, P]),
            "wd_decode_ply": (I, [P, P, U64, C.POINTER(P), P, P]),
            "wd_decode_spz": (I, [P, P, U64, C.POINTER(P), P, P]),
            "wd_decode_sog": (I, [P, C.POINTER(Sog), C.POINTER(P), P, P]),
            "wd_cloud_destroy": (I, [P]),
            "wd_view_prepare": (I, [FP, FP, FP, C.c_float, C.c_float, FP, FP, C.c_float, C.POINTER(View)]),
            "wd_state_create": (I, [P, U64, C.POINTER(P)]),
            "wd_state_destroy": (I, [P]),
            "wd_state_buffers": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
            "wd_sort": (I, [P, P, C.POINTER(View), P]),
            "wd_sort_noalloc": (I, [P, P, U64, C.POINTER(View), P, P, P]),
            "wd_state_read": (I, [P, P, P, U64, P, U64, C.POINTER(Stats)]),
            "wd_state_stats": (I, [P, P, C.POINTER(Stats)]),
            "wd_state_read_keys": (I, [P, P, P, P, U64]),
            "wd_state_ellipses": (I, [P, P, I]),
            "wd_state_keys_per_thread": (I, [P, U32]),
            "wd_state_ellipse_buffer": (I, [P, C.POINTER(P)]),
            "wd_state_read_ellipses": (I, [P, P, P, U64]),
            "wd_pool_create": (I, [P, P, U64, C.POINTER(P)]),
            "wd_pool_device": (I, [P, C.POINTER(P)]),
            "wd_pool_destroy": (I, [P]),
            "wd_particles_sort": (I, [P, P, U64, U64, C.POINTER(ParticleView), P]),
            "wd_profile": (I, [P, I]),
            "wd_profile_read": (I, [P, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(U64), I,
                                    C.POINTER(I)]),
            "wd_time_sorts": (I, [P, P, P, P, U32, P]),
            "wd_batch_create": (I, [P, U32, C.POINTER(P)]),
            "wd_batch_schedule": (I, [P, I, U32]),
            "wd_batch_configure": (I, [P, I, U32, U32]),
            "wd_batch_tune": (I, [P, P, U32, C.POINTER(Tuning)]),
            "wd_batch_set": (I, [P, C.POINTER(Job), U32]),
            "wd_batch_run": (I, [P, P]),
            "wd_batch_stuck": (I, [P, P, C.POINTER(U32)]),
            "wd_batch_destroy": (I, [P]),
            "wd_mlp_create": (I, [P, U64, C.POINTER(P)]),
            "wd_mlp_destroy": (I, [P]),
            "wd_mlp_set": (I, [P, P]),
            "wd_mlp_get": (I, [P, P]),
            "wd_mlp_save": (I, [P, P, U64, C.POINTER(U64)]),
            "wd_mlp_load": (I, [P, P, U64]),
            "wd_mlp_grad": (I, [P, P, P, P, U32, P, C.POINTER(C.c_float)]),
            "wd_mlp_train": (I, [P, P, P, P, U32, U32, U32, C.c_float, U64, C.POINTER(C.c_float)]),
            "wd_mlp_predict": (I, [P, P, P, U32, P]),
            "wd_pose_features": (I, [C.POINTER(View), P]),
            "wd_learned_create": (I, [P, U32, C.POINTER(P)]),
            "wd_learned_run": (I, [P, P, P, P, C.POINTER(View), C.POINTER(P), U32, U32]),
            "wd_learned_destroy": (I, [P]),
            "wd_state_quantiles": (I, [P, P, P]),
            "wd_synth_records": (I, [I, U64, U64, P]),
            "wd_debug_canonical": (I, [P, P, C.POINTER(View), P]),
            "wd_debug_selftest": (I, [P, P, P, U32, P]),
            "wd_debug_copy": (I, [P, U64, I, C.POINTER(C.c_double)]),
            "wd_debug_batch_force_stuck": (I, [P, P]),
        }
        missing = sorted(set(EXPORTS) - set(sig))
        if missing:
            raise ImportError("winding_amd: no ctypes signature for %s" % ", ".join(missing))
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != WD_OK and rc != 1:
        raise WindingError(rc, lib().wd_last_error().decode())
    return rc


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def sort_noalloc(ctx, records, n, view, order, count, cov2d):
    """wd_sort_noalloc: device addresses (ints) of the records (n x 32 bytes), order (u32 x n), count (u32) and
    cov2d (f32x4 x n); allocates nothing, enqueued on the context's stream."""
    check(lib().wd_sort_noalloc(ctx.ptr, records, n, C.byref(view), order, count, cov2d))


def synth_records(kind, seed, n):
    """n seeded .splat records (uint8 array n x 32): kind 0 Gaussian mixture, 1 slab."""
    out = np.empty((n, 32), dtype=np.uint8)
    check(lib().wd_synth_records(kind, seed, n, out.ctypes.data))
    return out


def prepare_view(view, projection, width, height, mn, mx, cull=WD_SPLAT_CULL, world=None):
    """SplatPass.prepare's uniform for one cloud (src/render/splats.js:559-587); None when wholly behind."""
    v = View()
    view = np.ascontiguousarray(view, np.float32)
    projection = np.ascontiguousarray(projection, np.float32)
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    w = None if world is None else _fp(np.ascontiguousarray(world, np.float32))
    rc = check(lib().wd_view_prepare(_fp(view), w, _fp(projection), width, height, _fp(mn), _fp(mx), cull,
                                     C.byref(v)))
    return None if rc == 1 else v


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """mat4LookAt (src/core/math/mat4.js:243-273), f64 maths stored as f32."""
    eye, target, up = (np.asarray(x, np.float64) for x in (eye, target, up))
    z = eye - target
    z = z / np.sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2])
    x = np.array([up[1] * z[2] - up[2] * z[1], up[2] * z[0] - up[0] * z[2], up[0] * z[1] - up[1] * z[0]])
    x = x / np.sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2])
    y = np.array([z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]])
    m = np.zeros(16)
    m[0], m[1], m[2] = x[0], y[0], z[0]
    m[4], m[5], m[6] = x[1], y[1], z[1]
    m[8], m[9], m[10] = x[2], y[2], z[2]
    m[12] = -(x[0] * eye[0] + x[1] * eye[1] + x[2] * eye[2])
    m[13] = -(y[0] * eye[0] + y[1] * eye[1] + y[2] * eye[2])
    m[14] = -(z[0] * eye[0] + z[1] * eye[1] + z[2] * eye[2])
    m[15] = 1
    return m.astype(np.float32)


def perspective(fovy, aspect, near):
    """mat4PerspectiveReverseZInfinite (src/core/math/mat4.js:455-468)."""
    f = 1.0 / np.tan(fovy * 0.5)
    m = np.zeros(16)
    m[0] = f / aspect
    m[5] = f
    m[11] = -1
    m[14] = near
    return m.astype(np.float32)


class Context:
    """A device and its stream (wd_open)."""

    def __init__(self, device=0):
        p = C.c_void_p()
        check(lib().wd_open(device, C.byref(p)))
        self.ptr = p

    def close(self):
        if self.ptr:
            check(lib().wd_close(self.ptr))
            self.ptr = None

    def sync(self):
        check(lib().wd_sync(self.ptr))

    def stream(self):
        s = C.c_void_p()
        check(lib().wd_stream(self.ptr, C.byref(s)))
        return s.value

    def profile(self, on):
        check(lib().wd_profile(self.ptr, 1 if on else 0))

    def profile_read(self):
        names = (C.c_char_p * 16)()
        ms = (C.c_double * 16)()
        launches = (C.c_uint64 * 16)()
        n = C.c_int()
        check(lib().wd_profile_read(self.ptr, names, ms, launches, 16, C.byref(n)))
        return {names[i].decode(): (ms[i], launches[i]) for i in range(n.value) if launches[i]}

    def copy_ms(self, nbytes, reps=20):
        """Mean ms of a plain device copy of nbytes (one read stream, one write stream)."""
        ms = C.c_double()
        check(lib().wd_debug_copy(self.ptr, nbytes, reps, C.byref(ms)))
        return ms.value

    def selftest(self, args, expect):
        args = np.ascontiguousarray(args, np.float64)
        expect = np.ascontiguousarray(expect, np.float64)
        mm = np.zeros(3, np.uint32)
        check(lib().wd_debug_selftest(self.ptr, args.ctypes.data, expect.ctypes.data, len(args) // 2,
                                      mm.ctypes.data))
        return mm

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Cloud:
    """Uploaded .splat records (new Splats(rhi, data), src/render/splats.js:389-409)."""

    def __init__(self, ctx, records, device=False, nbytes=None):
        p = C.c_void_p()
        self.min = np.zeros(3, np.float32)
        self.max = np.zeros(3, np.float32)
        if device:
            check(lib().wd_cloud_create_device(ctx.ptr, records, nbytes, C.byref(p), _fp(self.min), _fp(self.max)))
            self.count = nbytes // 32
        else:
            records = np.ascontiguousarray(records, dtype=np.uint8)
            check(lib().wd_cloud_create(ctx.ptr, records.ctypes.data, records.nbytes, C.byref(p), _fp(self.min),
                                        _fp(self.max)))
            self.count = records.nbytes // 32
        self.ctx = ctx
        self.ptr = p

    @classmethod
    def _wrap(cls, ctx, p, mn, mx):
        self = cls.__new__(cls)
        self.ctx, self.ptr, self.min, self.max = ctx, p, mn, mx
        n = C.c_uint64()
        check(lib().wd_cloud_info(p, C.byref(n), None, None))
        self.count = n.value
        d = C.c_uint32()
        check(lib().wd_cloud_read(ctx.ptr, p, None, None, None, C.byref(d)))
        self.degree = d.value
        return self

    @classmethod
    def from_splats(cls, ctx, centers, covariances):
        """The reference's Splats data (parseSplats' output): centres (n, 4) f32 with the colour bits in the
        fourth, covariances (n, 6) f32 (new Splats(rhi, data), src/render/splats.js:389-416)."""
        centers = np.ascontiguousarray(centers, np.float32).reshape(-1, 4)
        covariances = np.ascontiguousarray(covariances, np.float32).reshape(-1, 6)
        assert len(centers) == len(covariances)
        p = C.c_void_p()
        mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
        check(lib().wd_cloud_create_splats(ctx.ptr, centers.ctypes.data, covariances.ctypes.data, len(centers),
                                           C.byref(p), _fp(mn), _fp(mx)))
        return cls._wrap(ctx, p, mn, mx)

    def read(self):
        """(centres (n, 4) f32, covariances (n, 6) f32, sh words (n, W) u32 or None, degree) of a decoded cloud."""
        centers = np.zeros((self.count, 4), np.float32)
        covs = np.zeros((self.count, 6), np.float32)
        w = (WD_SH_COEFFICIENTS[self.degree] * 3 + 1) // 2
        sh = np.zeros((self.count, w), np.uint32) if self.degree else None
        check(lib().wd_cloud_read(self.ctx.ptr, self.ptr, centers.ctypes.data, covs.ctypes.data,
                                  sh.ctypes.data if sh is not None else None, None))
        return centers, covs, sh, self.degree

    def destroy(self):
        if self.ptr:
            check(lib().wd_cloud_destroy(self.ptr))
            self.ptr = None

    def set_sh(self, words, degree):
        """Spherical harmonics of degree 1-3 in the reference's `sh` layout (u32 words, halves), 0 to remove."""
        words = np.ascontiguousarray(words, dtype=np.uint32) if degree else np.zeros(1, np.uint32)
        check(lib().wd_cloud_sh(self.ctx.ptr, self.ptr, words.ctypes.data, words.size if degree else 0, degree))
        self.degree = degree

    def canonical(self, view):
        out = np.zeros((self.count, 5), np.float64)
        check(lib().wd_debug_canonical(self.ctx.ptr, self.ptr, C.byref(view), out.ctypes.data))
        return out


def _decoded(ctx, fn, data):
    data = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    data = np.ascontiguousarray(data, np.uint8)
    p = C.c_void_p()
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    check(fn(ctx.ptr, data.ctypes.data, data.nbytes, C.byref(p), _fp(mn), _fp(mx)))
    return Cloud._wrap(ctx, p, mn, mx)


def decode_ply(ctx, data):
    """A binary .ply as training writes it, decoded on the GPU (parsePly, src/scene/splats.js:398-468)."""
    return _decoded(ctx, lib().wd_decode_ply, data)


def decode_spz(ctx, data):
    """A .spz of version 1 to 3, gzipped or not, decoded on the GPU (parseSpz, src/scene/splats.js:225-303)."""
    return _decoded(ctx, lib().wd_decode_spz, data)


def _unzip(data):
    """A zip's files by name, stored and deflated entries, not zip64 (unzip, src/scene/splats.js:98-126)."""
    import io
    import zipfile
    try:
        z = zipfile.ZipFile(io.BytesIO(bytes(data)))
    except zipfile.BadZipFile as e:
        raise WindingError(WD_E_DATA, "splats: a zip with no end record: cut short?") from e
    files = {}
    for info in z.infolist():
        if info.compress_type not in (zipfile.ZIP_STORED, zipfile.ZIP_DEFLATED):
            raise WindingError(WD_E_DATA, "splats: %s is compressed with method %d; stored and deflate are read"
                               % (info.filename, info.compress_type))
        files[info.filename] = (lambda name: (lambda: z.read(name)))(info.filename)
    return files


def read_sog(ctx, data, decode_image=None):
    """A .sog (a zip of meta.json and WebP images), unpacked here and decoded on the GPU (readSog,
    src/scene/splats.js:128-203). decode_image(bytes) -> (width, height, rgba uint8), the bytes as stored."""
    import json
    if not callable(decode_image):
        raise WindingError(WD_E_ARG, "splats: a .sog needs decodeImage, for its WebP images")
    files = _unzip(data)

    def file(name):
        if name not in files:
            raise WindingError(WD_E_DATA, "splats: the .sog has no %s" % name)
        return files[name]()

    meta = json.loads(file("meta.json").decode())
    if meta.get("version") != 2:
        raise WindingError(WD_E_DATA, "splats: a .sog of version %s; version 2 is read (re-save it with splat-transform)"
                           % meta.get("version", 1))
    count = meta.get("count")
    if not (isinstance(count, int) and count > 0):
        raise WindingError(WD_E_DATA, "splats: a .sog of %s splats" % count)
    keep = []

    def image(name, check_count=True):
        w, h, rgba = decode_image(file(name))
        rgba = np.ascontiguousarray(rgba, np.uint8).reshape(-1)
        if rgba.size < 4 * w * h:
            raise WindingError(WD_E_DATA, "splats: %s decoded to %d bytes, not %d x %d rgba" % (name, rgba.size, w, h))
        if check_count and w * h < count:
            raise WindingError(WD_E_DATA, "splats: %s has %d pixels, for %d splats" % (name, w * h, count))
        keep.append(rgba)
        return SogImage(rgba.ctypes.data, w, h)

    def book(values):
        a = np.ascontiguousarray(values, np.float64)
        keep.append(a)
        return a.ctypes.data, a.size

    s = Sog()
    s.count = count
    s.means_l, s.means_u = image(meta["means"]["files"][0]), image(meta["means"]["files"][1])
    s.quats = image(meta["quats"]["files"][0])
    s.scales = image(meta["scales"]["files"][0])
    s.sh0 = image(meta["sh0"]["files"][0])
    bands = meta["shN"]["bands"] if meta.get("shN") else 0
    degree = bands if bands in (1, 2, 3) else 0
    for a in range(3):
        s.mins[a] = meta["means"]["mins"][a]
        s.maxs[a] = meta["means"]["maxs"][a]
    s.scale_codebook, s.scale_codebook_size = book(meta["scales"]["codebook"])
    s.colour_codebook, s.colour_codebook_size = book(meta["sh0"]["codebook"])
    if degree:
        s.bands = degree
        s.sh_count = meta["shN"].get("count", 0xFFFFFFFF)
        s.sh_codebook, s.sh_codebook_size = book(meta["shN"]["codebook"])
        s.centroids = image(meta["shN"]["files"][0], check_count=False)
        s.labels = image(meta["shN"]["files"][1])
    p = C.c_void_p()
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    check(lib().wd_decode_sog(ctx.ptr, C.byref(s), C.byref(p), _fp(mn), _fp(mx)))
    return Cloud._wrap(ctx, p, mn, mx)


class Pool:
    """A particle pool on the device (32-byte particles: position, age, velocity, lifetime)."""

    def __init__(self, ctx, particles):
        particles = np.ascontiguousarray(particles, np.uint8)
        assert particles.nbytes % 32 == 0
        self.ctx = ctx
        self.count = particles.nbytes // 32
        p = C.c_void_p()
        check(lib().wd_pool_create(ctx.ptr, particles.ctypes.data, self.count, C.byref(p)))
        self.ptr = p

    def destroy(self):
        if self.ptr:
            check(lib().wd_pool_destroy(self.ptr))
            self.ptr = None


def particle_view(eye, forward):
    v = ParticleView()
    v.eye[:3] = [float(x) for x in eye]
    v.forward[:3] = [float(x) for x in forward]
    return v


class State:
    """Sort buffers for clouds up to `capacity` splats (SplatPass._stateFor, src/render/splats.js:629-674)."""

    def __init__(self, ctx, capacity):
        p = C.c_void_p()
        check(lib().wd_state_create(ctx.ptr, capacity, C.byref(p)))
        self.ctx = ctx
        self.ptr = p
        self.capacity = capacity

    def destroy(self):
        if self.ptr:
            check(lib().wd_state_destroy(self.ptr))
            self.ptr = None

    def sort(self, cloud, view):
        check(lib().wd_sort(self.ctx.ptr, cloud.ptr, C.byref(view), self.ptr))

    def time_sorts(self, cloud, view, reps):
        """Milliseconds of each of `reps` sorts, HIP events around each on the library's stream."""
        ms = np.zeros(reps, np.float64)
        check(lib().wd_time_sorts(self.ctx.ptr, cloud.ptr, C.byref(view), self.ptr, reps, ms.ctypes.data))
        return ms

    def stats(self):
        s = Stats()
        check(lib().wd_state_stats(self.ctx.ptr, self.ptr, C.byref(s)))
        return s

    def quantiles(self):
        """The 257 equal-mass key edges of the last sort (the predictor's training target)."""
        edges = np.zeros(257, np.uint32)
        check(lib().wd_state_quantiles(self.ctx.ptr, self.ptr, edges.ctypes.data))
        return edges

    def read(self, n_cov=0):
        """(order[:V], cov2d n_cov x 4 float32, stats)."""
        s = self.stats()
        order = np.empty(max(s.count, 1), np.uint32)
        cov = np.empty((max(n_cov, 1), 4), np.float32)
        check(lib().wd_state_read(self.ctx.ptr, self.ptr, order.ctypes.data, len(order),
                                  cov.ctypes.data if n_cov else None, n_cov, C.byref(s)))
        return order[:s.count], cov[:n_cov], s

    def buffers(self):
        """Device addresses (order, count, cov2d)."""
        o, c, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().wd_state_buffers(self.ptr, C.byref(o), C.byref(c), C.byref(v)))
        return o.value, c.value, v.value

    def sort_particles(self, pool, view, first=0, count=None):
        """Sorts slots [first, first + count) of `pool` far to near (wd_particles_sort)."""
        count = pool.count - first if count is None else count
        check(lib().wd_particles_sort(self.ctx.ptr, pool.ptr, first, count, C.byref(view), self.ptr))

    def set_keys_per_thread(self, kpt):
        """The tile this state's sorts take: 4, 8 or 16 keys a thread, or 0 (by the cloud's size)."""
        check(lib().wd_state_keys_per_thread(self.ptr, kpt))

    def set_ellipses(self, on):
        """Whether each sort also writes the draw's ellipse set-up (8 floats a splat) beside cov2d."""
        check(lib().wd_state_ellipses(self.ctx.ptr, self.ptr, 1 if on else 0))

    def read_ellipses(self, n):
        """(n, 8) f32: clip x y z w, then (along e, across e) in pixels; zeros where nothing is drawn."""
        out = np.empty((n, 8), np.float32)
        check(lib().wd_state_read_ellipses(self.ctx.ptr, self.ptr, out.ctypes.data, n))
        return out

    def read_keys(self, n):
        keys = np.empty(n, np.uint16)
        mask = np.empty((n + 63) // 64, np.uint64)
        check(lib().wd_state_read_keys(self.ctx.ptr, self.ptr, keys.ctypes.data, mask.ctypes.data, n))
        keep = np.unpackbits(mask.view(np.uint8), bitorder="little")[:n].astype(bool)
        return keys, keep




WD_SCHEDULE_PHASED, WD_SCHEDULE_PERSISTENT = 0, 1


class Batch:
    """Many (cloud, view, state) sorts in one hipGraph launch (wd_batch_*): six phase kernels over
    every job, or one persistent kernel over a ticketed queue."""

    def __init__(self, ctx, max_jobs):
        p = C.c_void_p()
        check(lib().wd_batch_create(ctx.ptr, max_jobs, C.byref(p)))
        self.ctx = ctx
        self.ptr = p

    def set(self, jobs):
        """jobs: [(Cloud, View, State)]."""
        arr = (Job * len(jobs))()
        for i, (cloud, view, state) in enumerate(jobs):
            arr[i].cloud = cloud.ptr.value
            arr[i].view = view
            arr[i].state = state.ptr.value
        check(lib().wd_batch_set(self.ptr, arr, len(jobs)))

    def run(self):
        check(lib().wd_batch_run(self.ctx.ptr, self.ptr))

    def schedule(self, which, super_tiles=1):
        """Phased: super_tiles (1 or 2) tiles a scatter workgroup takes at once. Persistent: tiles an
        item takes one after the other (1 to 64)."""
        check(lib().wd_batch_schedule(self.ptr, which, super_tiles))

    def configure(self, which, super_tiles, keys_per_thread):
        """The schedule and work item as schedule(), and the tile: 8 or 16 keys a thread (2048 or 4096)."""
        check(lib().wd_batch_configure(self.ptr, which, super_tiles, keys_per_thread))

    def tune(self, reps=5):
        """Times the configurations on the set jobs and keeps the fastest: a Tuning."""
        t = Tuning()
        check(lib().wd_batch_tune(self.ctx.ptr, self.ptr, reps, C.byref(t)))
        return t

    def stuck(self):
        s = C.c_uint32()
        check(lib().wd_batch_stuck(self.ctx.ptr, self.ptr, C.byref(s)))
        return s.value

    def force_stuck(self):
        """Raises the live "a wait ran out" flag from the host (a test of the reporting)."""
        check(lib().wd_debug_batch_force_stuck(self.ctx.ptr, self.ptr))

    def destroy(self):
        if self.ptr:
            check(lib().wd_batch_destroy(self.ptr))
            self.ptr = None


WD_MLP_PARAMS = 51200


def pose_features(view):
    """The predictor's 12 inputs for a view."""
    out = np.zeros(WD_POSE_FLOATS, np.float32)
    check(lib().wd_pose_features(C.byref(view), out.ctypes.data))
    return out


class Mlp:
    """The learned depth-CDF predictor (wd_mlp_*): 12 -> 128 -> 128 -> 256, bf16 matrix cores."""

    def __init__(self, ctx, seed=0):
        p = C.c_void_p()
        check(lib().wd_mlp_create(ctx.ptr, seed, C.byref(p)))
        self.ctx = ctx
        self.ptr = p

    def destroy(self):
        if self.ptr:
            check(lib().wd_mlp_destroy(self.ptr))
            self.ptr = None

    def params(self):
        out = np.zeros(WD_MLP_PARAMS, np.float32)
        check(lib().wd_mlp_get(self.ptr, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().wd_mlp_set(self.ptr, p.ctypes.data))

    def save(self):
        size = C.c_uint64()
        check(lib().wd_mlp_save(self.ptr, None, 0, C.byref(size)))
        buf = np.zeros(size.value, np.uint8)
        check(lib().wd_mlp_save(self.ptr, buf.ctypes.data, size.value, C.byref(size)))
        return buf.tobytes()

    def load(self, data):
        buf = np.frombuffer(data, np.uint8).copy()
        check(lib().wd_mlp_load(self.ptr, buf.ctypes.data, len(buf)))

    def grad(self, poses, targets):
        poses = np.ascontiguousarray(poses, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(WD_MLP_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().wd_mlp_grad(self.ctx.ptr, self.ptr, poses.ctypes.data, targets.ctypes.data, len(poses),
                                g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def train(self, poses, targets, epochs, batch=256, lr=1e-3, seed=0):
        poses = np.ascontiguousarray(poses, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        loss = C.c_float()
        check(lib().wd_mlp_train(self.ctx.ptr, self.ptr, poses.ctypes.data, targets.ctypes.data, len(poses), epochs,
                                 batch, lr, seed, C.byref(loss)))
        return loss.value

    def predict(self, poses):
        poses = np.ascontiguousarray(poses, np.float32)
        out = np.zeros((len(poses), 256), np.float32)
        check(lib().wd_mlp_predict(self.ctx.ptr, self.ptr, poses.ctypes.data, len(poses), out.ctypes.data))
        return out


class Learned:
    """The learned sort (wd_learned_*): up to 8 views of one cloud from one read of the records, binned by
    the predictor's depth CDF, exact by construction (a failed check re-sorts on the exact path)."""

    def __init__(self, ctx, max_views=8):
        p = C.c_void_p()
        check(lib().wd_learned_create(ctx.ptr, max_views, C.byref(p)))
        self.ctx = ctx
        self.ptr = p

    def run(self, mlp, cloud, views, states, cap=0):
        k = len(views)
        varr = (View * k)(*views)
        sarr = (C.c_void_p * k)(*[s.ptr.value for s in states])
        assert len(states) == k
        check(lib().wd_learned_run(self.ctx.ptr, self.ptr, mlp.ptr, cloud.ptr, varr, sarr, k, cap))

    def destroy(self):
        if self.ptr:
            check(lib().wd_learned_destroy(self.ptr))
            self.ptr = None


def widths_from_edges(edges):
    """The predictor's target: bin widths in key space, (edges[k+1] - edges[k]) / 65536."""
    e = np.asarray(edges, np.float64)
    return (np.diff(e) / 65536.0).astype(np.float32)


from . import workloads  # noqa: E402  (needs the names above)
