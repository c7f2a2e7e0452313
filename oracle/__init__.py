"""ctypes binding of the CPU oracle (oracle/splat_oracle.c).

TEST INFRASTRUCTURE ONLY. The parity tests, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py load it as the answer the GPU's results are
compared with; nothing in the product (winding_amd/) imports it. Build:
python -m winding_amd.build.
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
        L.or_ellipse.argtypes = [P, C.c_uint, P]
        L.or_ellipse.restype = C.c_int
        L.or_ellipses.argtypes = [P, U64, C.POINTER(OView), P]
        L.or_particle_edges.argtypes = [P]
        L.or_particle_keys.argtypes = [P, U64, P, P, P, P]
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
    """The eye in the capture's space (SplatPass.prepare, src/render/splats.js:585-586), f32 x 3."""
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


def ellipse(cov, abyte):
    """The vertex shader's axes for one footprint (src/render/splats.js:263-291), FP64: (along e, across e) or None."""
    f = np.ascontiguousarray(cov, np.float64)
    out = np.zeros(4, np.float64)
    return out if lib().or_ellipse(f.ctypes.data, int(abyte), out.ctypes.data) else None


def ellipses(records, view):
    """(n, 8) FP64: every splat's clip position and (along e, across e); zeros where nothing is drawn."""
    records = np.ascontiguousarray(records, np.uint8)
    n = records.nbytes // 32
    out = np.zeros((n, 8), np.float64)
    ov = to_oview(view)
    lib().or_ellipses(records.ctypes.data, n, C.byref(ov), out.ctypes.data)
    return out


def particle_edges():
    """The 4096 FP64 depth edges of the particle buckets (edges[0] = 0)."""
    out = np.zeros(4096, np.float64)
    lib().or_particle_edges(out.ctypes.data)
    return out


def particle_keys(pool, eye, forward):
    """Per particle (32-byte records): key (u16, 0 when dead), keep (bool)."""
    pool = np.ascontiguousarray(pool, np.uint8)
    n = pool.nbytes // 32
    e = np.ascontiguousarray(eye, np.float32)
    f = np.ascontiguousarray(forward, np.float32)
    keys = np.zeros(n, np.uint16)
    keep = np.zeros(n, np.uint8)
    lib().or_particle_keys(pool.ctypes.data, n, e.ctypes.data, f.ctypes.data, keys.ctypes.data, keep.ctypes.data)
    return keys, keep.astype(bool)


def particle_run(pool, eye, forward):
    """The draw order of a pool's live particles (indices, back to front, ties by index) and their keys."""
    keys, keep = particle_keys(pool, eye, forward)
    order = np.zeros(len(keys), np.uint32)
    keep8 = keep.astype(np.uint8)  # held: ctypes takes its address only
    v = lib().or_sort(keys.ctypes.data, keep8.ctypes.data, len(keys), order.ctypes.data)
    return order[:v], keys, keep


# ---- capture decoders (oracle/decode_oracle.c; the headers restated here) ----

_PLY_TYPES = {"char": (0, 1), "int8": (0, 1), "uchar": (1, 1), "uint8": (1, 1), "short": (2, 2), "int16": (2, 2),
              "ushort": (3, 2), "uint16": (3, 2), "int": (4, 4), "int32": (4, 4), "uint": (5, 4), "uint32": (5, 4),
              "float": (6, 4), "float32": (6, 4), "double": (7, 8), "float64": (7, 8)}
PLY_NEEDS = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
             "rot_0", "rot_1", "rot_2", "rot_3"]
_COEFS = [0, 3, 8, 15, 24]


class DecodeError(ValueError):
    pass


def _decode_lib():
    L = lib()
    if not getattr(L, "_decode_bound", False):
        P, U64 = C.c_void_p, C.c_uint64
        L.od_decode_ply.restype = C.c_int64
        L.od_decode_ply.argtypes = [P, U64, C.c_uint32, P, P, C.c_uint32, C.c_int, P, P, P]
        L.od_decode_spz.restype = C.c_int64
        L.od_decode_spz.argtypes = [P, U64, P, P, P, P, P]
        L.od_decode_sog.restype = C.c_int64
        L.od_decode_sog.argtypes = [C.c_uint32, C.c_int, C.c_uint32, P, P] + [P] * 7 + [C.c_uint32, U64, P, P, P,
                                                                                       P, P, P]
        L.od_half.restype = C.c_uint16
        L.od_half.argtypes = [C.c_double]
        L.od_byte.restype = C.c_uint
        L.od_byte.argtypes = [C.c_double]
        L._decode_bound = True
    return L


def half(v):
    """signedHalf(v): the reference's half bits of a number (src/render/hdr.js:152-163)."""
    return _decode_lib().od_half(float(v))


def colour_byte(v):
    """byte(v) (src/scene/splats.js:372)."""
    return _decode_lib().od_byte(float(v))


def _outputs(n, degree):
    w = (_COEFS[degree] * 3 + 1) // 2
    return (np.zeros((n, 4), np.float32), np.zeros((n, 6), np.float32), np.zeros((n, 2 * w), np.uint16))


def _result(centers, covs, halves, degree, bad):
    if bad >= 0:
        raise DecodeError("splats: splat %d has a position that is not a number" % bad)
    return {"count": len(centers), "degree": degree, "centers": centers, "covariances": covs,
            "sh": halves.view(np.uint32) if degree else None}


def ply_header(data):
    """parsePly's header (src/scene/splats.js:398-450): (end, count, stride, offsets, types, rest, degree)."""
    data = bytes(data)
    marker = b"end_header"
    end = -1
    for i in range(min(len(data), 65536) - len(marker)):
        if data[i:i + len(marker)] == marker:
            end = i + len(marker)
            break
    if end < 0:
        raise DecodeError("splats: a .ply with no end_header")
    if end < len(data) and data[end] == 0x0D:
        end += 1
    end += 1
    fmt, vertices = None, None
    for line in data[:end].decode("latin-1").replace("\r\n", "\n").split("\n"):
        words = line.split()
        if not words:
            continue
        if words[0] == "format":
            fmt = words[1] if len(words) > 1 else None
        elif words[0] == "element":
            w1 = words[1] if len(words) > 1 else "undefined"
            if vertices is None and w1 == "chunk":
                raise DecodeError("splats: a compressed .ply (from SuperSplat) is not read; export an uncompressed "
                                  ".ply or a .splat")
            if vertices is None and w1 != "vertex":
                raise DecodeError("splats: a .ply whose first element is '%s', not 'vertex'" % w1)
            if vertices is not None:
                break
            try:
                count = float(words[2])
            except (IndexError, ValueError):
                count = float("nan")
            vertices = {"count": count, "props": {}, "stride": 0}
        elif words[0] == "property" and vertices is not None:
            if words[1] == "list":
                raise DecodeError("splats: a .ply with list properties on its vertices is not a splat capture")
            if words[1] not in _PLY_TYPES:
                raise DecodeError("splats: unknown .ply property type '%s'" % words[1])
            t, size = _PLY_TYPES[words[1]]
            vertices["props"][words[2]] = (vertices["stride"], t)
            vertices["stride"] += size
    if vertices is None:
        raise DecodeError("splats: a .ply with no vertices")
    if fmt != "binary_little_endian":
        raise DecodeError("splats: a .ply in '%s' format; only binary_little_endian is read" % fmt)
    count = vertices["count"]
    if not (count > 0 and count == int(count)):
        raise DecodeError("splats: a .ply with %s vertices" % count)
    count = int(count)
    at = vertices["props"]
    missing = [name for name in PLY_NEEDS if name not in at]
    if missing:
        raise DecodeError("splats: this .ply is not a splat capture: no %s" % ", ".join(missing))
    stride = vertices["stride"]
    if end + count * stride > len(data):
        raise DecodeError("splats: the .ply is cut short: %d vertices need %d bytes, it has %d"
                          % (count, count * stride, len(data) - end))
    rest = 0
    while "f_rest_%d" % rest in at:
        rest += 1
    degree = _COEFS[:4].index(rest // 3) if rest % 3 == 0 and rest // 3 in _COEFS[:4] else -1
    if rest > 0 and degree < 1:
        raise DecodeError("splats: a .ply with %d f_rest_ properties; harmonics of degree 1, 2 or 3 have 9, 24 or 45"
                          % rest)
    names = PLY_NEEDS + ["f_rest_%d" % k for k in range(rest)]
    offs = np.array([at[nm][0] for nm in names] + [0] * (59 - len(names)), np.uint32)
    types = np.array([at[nm][1] for nm in names] + [0] * (59 - len(names)), np.uint8)
    return end, count, stride, offs, types, rest, max(degree, 0)


def decode_ply(data):
    """parsePly (src/scene/splats.js:398-468) in FP64: {count, degree, centers, covariances, sh}."""
    data = np.frombuffer(bytes(data), np.uint8)
    end, n, stride, offs, types, rest, degree = ply_header(data)
    body = np.ascontiguousarray(data[end:end + n * stride])
    centers, covs, halves = _outputs(n, degree)
    bad = _decode_lib().od_decode_ply(body.ctypes.data, n, stride, offs.ctypes.data, types.ctypes.data, rest, degree,
                                      centers.ctypes.data, covs.ctypes.data, halves.ctypes.data)
    return _result(centers, covs, halves, degree, bad)


_SPZ_ERRORS = {-2: "splats: gzipped, but not a .spz", -3: "splats: a .spz of version %(version)d; versions 1 to 3 are read",
               -4: "splats: a .spz with harmonics of degree %(degree)d", -5: "splats: a .spz with no splats",
               -6: "splats: the .spz is cut short"}


def decode_spz(data):
    """readSplats' gunzip and parseSpz (src/scene/splats.js:84-92,225-303) in FP64."""
    import gzip
    raw = bytes(data)
    if raw[:2] == b"\x1f\x8b":
        raw = gzip.decompress(raw)
    buf = np.frombuffer(raw, np.uint8).copy()
    count, degree = C.c_uint32(), C.c_int()
    code = _decode_lib().od_decode_spz(buf.ctypes.data, buf.size, C.byref(count), C.byref(degree), None, None, None)
    if code != -1:
        info = {"version": int.from_bytes(raw[4:8], "little") if len(raw) >= 8 else 0,
                "degree": raw[12] if len(raw) > 12 else 0}
        raise DecodeError(_SPZ_ERRORS.get(code, "splats: .spz %d") % info)
    centers, covs, halves = _outputs(count.value, degree.value)
    bad = _decode_lib().od_decode_spz(buf.ctypes.data, buf.size, C.byref(count), C.byref(degree), centers.ctypes.data,
                                      covs.ctypes.data, halves.ctypes.data)
    return _result(centers, covs, halves, degree.value, bad)


def decode_sog(meta, images):
    """readSog's loop (src/scene/splats.js:162-203) in FP64 on decoded images: meta as meta.json holds it,
    images {file name: (width, height, rgba uint8)}."""
    count = meta["count"]
    bands = meta["shN"]["bands"] if meta.get("shN") else 0
    degree = bands if bands in (1, 2, 3) else 0

    def img(name):
        return np.ascontiguousarray(images[name][2], np.uint8)

    def book(values):
        b = np.full(256, np.nan)
        v = np.asarray(values, np.float64)[:256]
        b[:len(v)] = v
        return b

    mins = np.asarray(meta["means"]["mins"], np.float64)
    maxs = np.asarray(meta["means"]["maxs"], np.float64)
    parts = [img(meta["means"]["files"][0]), img(meta["means"]["files"][1]), img(meta["quats"]["files"][0]),
             img(meta["scales"]["files"][0]), img(meta["sh0"]["files"][0])]
    if degree:
        cw, ch, crgba = images[meta["shN"]["files"][0]]
        parts += [img(meta["shN"]["files"][1]), np.ascontiguousarray(crgba, np.uint8)]
        sh_count, sh_book = meta["shN"].get("count", 0xFFFFFFFF), book(meta["shN"]["codebook"])
    else:
        cw, crgba = 0, np.zeros(4, np.uint8)
        parts += [np.zeros(4 * count, np.uint8), crgba]
        sh_count, sh_book = 0, book([])
    scale_book, colour_book = book(meta["scales"]["codebook"]), book(meta["sh0"]["codebook"])
    centers, covs, halves = _outputs(count, degree)
    bad = _decode_lib().od_decode_sog(count, degree, sh_count, mins.ctypes.data, maxs.ctypes.data,
                                      *[p.ctypes.data for p in parts], cw, parts[6].size,
                                      scale_book.ctypes.data, colour_book.ctypes.data, sh_book.ctypes.data,
                                      centers.ctypes.data, covs.ctypes.data, halves.ctypes.data)
    return _result(centers, covs, halves, degree, bad)


def run_sigma(centers, covs, view):
    """Order, keys, keep and FP64 cov (n x 3) of a cloud given as centres and f32 Sigma."""
    L = lib()
    if not getattr(L, "_sigma_bound", False):
        L.or_run_sigma.restype = C.c_uint32
        L.or_run_sigma.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(OView)] + [C.c_void_p] * 4
        L._sigma_bound = True
    centers = np.ascontiguousarray(centers, np.float32)
    covs = np.ascontiguousarray(covs, np.float32)
    n = centers.size // 4
    order = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint16)
    keep = np.zeros(n, np.uint8)
    cov = np.zeros((n, 3), np.float64)
    ov = to_oview(view)
    v = L.or_run_sigma(centers.ctypes.data, covs.ctypes.data, n, C.byref(ov), order.ctypes.data, keys.ctypes.data,
                       keep.ctypes.data, cov.ctypes.data)
    return order[:v], keys, keep.astype(bool), cov
