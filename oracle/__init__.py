"""ctypes binding of the CPU oracle (oracle/splat_oracle.c).

TEST INFRASTRUCTURE ONLY. The parity tests, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py load it as the checker; nothing in the product
(winding_amd/) imports it. Build: python -m winding_amd.build.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OView(C.Structure):
    _fields_ = [
        ("model_view", C.c_float * 16),
        ("projection", C.c_float * 16),
        ("viewport", C.c_float * 2),
        ("depth_min", C.c_float),
        ("depth_scale", C.c_float),
        ("cull", C.c_float),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle: %s is not built; run python -m winding_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, U64 = C.c_void_p, C.c_uint64
        L.or_sigma.argtypes = [P, P]
        L.or_splat.restype = C.c_int
        L.or_splat.argtypes = [P, C.POINTER(OView), P, P, P]
        L.or_keys.argtypes = [P, U64, C.POINTER(OView), P, P, P, P]
        L.or_sort.restype = C.c_uint32
        L.or_sort.argtypes = [P, P, U64, P]
        L.or_run.restype = C.c_uint32
        L.or_run.argtypes = [P, U64, C.POINTER(OView), P, P, P, P]
        L.or_bounds.restype = C.c_int64
        L.or_bounds.argtypes = [P, U64, P, P]
        L.or_prepare.restype = C.c_int
        L.or_prepare.argtypes = [P, P, P, C.c_float, C.c_float, P, P, C.c_float, C.POINTER(OView)]
        L.or_look_at.argtypes = [P, P, P, P]
        L.or_perspective.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.or_mat4_multiply.argtypes = [P, P, P]
        L.or_eye.restype = C.c_int
        L.or_eye.argtypes = [P, P]
        L.or_sh_words.restype = C.c_int
        L.or_sh_words.argtypes = [C.c_int]
        L.or_shade.restype = C.c_uint32
        L.or_shade.argtypes = [P, P, C.c_int, P]
        L.or_shade_all.argtypes = [P, U64, P, C.c_int, P, P]
        _lib = L
    return _lib


def to_oview(view):
    """Any object laid out as wd_view (37 floats) -> OView."""
    return OView.from_buffer_copy(bytes(view))


def sigma(record):
    out = np.zeros(6, np.float64)
    lib().or_sigma(np.ascontiguousarray(record, np.uint8).ctypes.data, out.ctypes.data)
    return out


def keys(records, view):
    """Per splat: key (u16, 0 when culled), keep (bool), FP64 cov n x 3, alpha*area (NaN when untested)."""
    records = np.ascontiguousarray(records, np.uint8).reshape(-1, 32)
    n = len(records)
    k = np.zeros(n, np.uint16)
    keep = np.zeros(n, np.uint8)
    cov = np.zeros((n, 3), np.float64)
    val = np.zeros(n, np.float64)
    v = to_oview(view)
    lib().or_keys(records.ctypes.data, n, C.byref(v), k.ctypes.data, keep.ctypes.data, cov.ctypes.data,
                  val.ctypes.data)
    return k, keep.astype(bool), cov, val


def sort(k, keep):
    k = np.ascontiguousarray(k, np.uint16)
    keep8 = np.ascontiguousarray(keep, np.uint8)
    order = np.zeros(max(len(k), 1), np.uint32)
    count = lib().or_sort(k.ctypes.data, keep8.ctypes.data, len(k), order.ctypes.data)
    return order[:count]


def run(records, view):
    """The whole path: (order, keys, keep, cov)."""
    k, keep, cov, _ = keys(records, view)
    return sort(k, keep), k, keep, cov


def bounds(records):
    records = np.ascontiguousarray(records, np.uint8).reshape(-1, 32)
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    bad = lib().or_bounds(records.ctypes.data, len(records), mn.ctypes.data, mx.ctypes.data)
    return int(bad), mn, mx


def prepare(view, projection, width, height, mn, mx, cull, world=None):
    out = OView()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    view, projection, mn, mx = f(view), f(projection), f(mn), f(mx)
    w = None if world is None else f(world)
    rc = lib().or_prepare(view.ctypes.data, None if w is None else w.ctypes.data, projection.ctypes.data,
                          width, height, mn.ctypes.data, mx.ctypes.data, cull, C.byref(out))
    return None if rc == 1 else out


def look_at(eye, center=(0, 0, 0), up=(0, 1, 0)):
    out = np.zeros(16, np.float32)
    d = lambda a: np.ascontiguousarray(a, np.float64)
    eye, center, up = d(eye), d(center), d(up)
    lib().or_look_at(out.ctypes.data, eye.ctypes.data, center.ctypes.data, up.ctypes.data)
    return out


def perspective(fovy, aspect, near):
    out = np.zeros(16, np.float32)
    lib().or_perspective(out.ctypes.data, fovy, aspect, near)
    return out


def mat4_multiply(a, b):
    out = np.zeros(16, np.float32)
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    lib().or_mat4_multiply(out.ctypes.data, a.ctypes.data, b.ctypes.data)
    return out


def eye(model_view):
    """The eye in the capture's space (SplatPass.prepare, src/render/splats.js:580), f32 x 3."""
    mv = np.ascontiguousarray(np.asarray(model_view, np.float32).reshape(16))
    out = np.zeros(3, np.float32)
    lib().or_eye(mv.ctypes.data, out.ctypes.data)
    return out


def sh_words(degree):
    return lib().or_sh_words(degree)


def shade(records, sh, degree, view):
    """The shade pass's colour (rgba bytes as u32) of every splat, FP64 (src/render/splats.js:335-381)."""
    records = np.ascontiguousarray(records, np.uint8)
    sh = np.ascontiguousarray(sh, np.uint32)
    n = records.nbytes // 32
    assert sh.size >= n * sh_words(degree)
    e = eye(to_oview(view).model_view)
    out = np.zeros(n, np.uint32)
    lib().or_shade_all(records.ctypes.data, n, sh.ctypes.data, degree, e.ctypes.data, out.ctypes.data)
    return out
