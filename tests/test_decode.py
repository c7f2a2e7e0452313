"""Capture decoders on the GPU (.ply, .spz, .sog -> centres, covariances and harmonics; src/scene/
splats.js:58-303,390-468) and sorts of such a cloud.

The oracle (oracle/decode_oracle.c, FP64 as the reference's JavaScript runs; the headers restated in
oracle/__init__.py) is pinned by the reference's own fixtures and known answers (test/splats.test.js:
73-329): the builders below restate its ply(), spz() and zip() helpers and its .sog, and each test
asserts what the reference's asserts. Its half-float rounding (halfBits, half up rather than to even,
saturating) is pinned against numpy's float16 away from ties and at constructed ties. The GPU is held
to the oracle: centres, colour words and harmonics bit for bit, covariances within 1e-6 of the largest
term (exp and the quaternion's length may differ in the last f64 bit), and the sort of a decoded cloud
bit for bit against the oracle's sort of the same centres and Sigma.
"""
import gzip
import json
import math
import struct

import numpy as np
import pytest

import oracle
import winding_amd as wa

SH_C0 = 0.28209479177387814
NEEDS = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
         "rot_0", "rot_1", "rot_2", "rot_3"]


def js_round(x):
    return math.floor(x + 0.5) if x - math.floor(x) != 0.5 else math.floor(x) + 1


def ply(properties, rows, types=None):
    """A binary .ply of the given properties (test/splats.test.js:32-42); every one a float unless typed."""
    types = types or {}
    fmt = {"float": "<f", "double": "<d", "uchar": "<B", "short": "<h", "int": "<i"}
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % len(rows)
    head += "".join("property %s %s\n" % (types.get(p, "float"), p) for p in properties) + "end_header\n"
    def value(t, v):
        return int(round(v)) if t in ("uchar", "short", "int") else v
    body = b"".join(struct.pack(fmt[types.get(p, "float")], value(types.get(p, "float"), row.get(p, 0)))
                    for row in rows for p in properties)
    return head.encode() + body


def colour_of(s, i):
    bits = int(s["centers"][i, 3:4].view(np.uint32)[0])
    return [bits & 255, (bits >> 8) & 255, (bits >> 16) & 255, bits >> 24]


def from_half(h):
    sign = -1 if h & 0x8000 else 1
    e, m = (h >> 10) & 31, h & 1023
    return sign * (m / 1024 * 2 ** -14 if e == 0 else (1 + m / 1024) * 2 ** (e - 15))


def u8(v):
    return max(0, min(255, js_round(v)))


def half_of(v):
    """The reference test's halfOf, for version 1 positions (test/splats.test.js:174-181)."""
    bits = struct.unpack("<I", struct.pack("<f", v))[0]
    return ((bits >> 16) & 0x8000) | ((((bits >> 23) & 255) - 127 + 15) << 10) | js_round((bits & 0x7FFFFF) / 8192)


FLIP = [-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1]


def spz(version, splats, degree=0, fraction=12, raw=False):
    """A .spz packed as the reference test packs one (test/splats.test.js:121-172): RUB, quantised, gzipped."""
    n, dim = len(splats), [0, 3, 8, 15, 24][degree]
    out = bytearray(struct.pack("<IIIBBBB", 0x5053474E, version, n, degree, fraction, 0, 0))
    for s in splats:
        x, y, z = s["position"]
        for v in (x, -y, -z):
            if version == 1:
                out += struct.pack("<H", half_of(v))
            else:
                f = js_round(v * (1 << fraction)) & 0xFFFFFF
                out += bytes([f & 255, (f >> 8) & 255, f >> 16])
    for s in splats:
        out.append(u8(255 / (1 + math.exp(-s["opacity"]))))
    for s in splats:
        out += bytes(u8(v * (0.15 * 255) + 0.5 * 255) for v in s["dc"])
    for s in splats:
        out += bytes(u8((v + 10) * 16) for v in s["logScale"])
    for s in splats:
        w, x, y, z = s["rotation"]
        ln = math.sqrt(w * w + x * x + y * y + z * z)
        q = [x / ln, -y / ln, -z / ln, w / ln]
        if version < 3:
            sign = -127.5 if q[3] < 0 else 127.5
            out += bytes(u8(q[k] * sign + 127.5) for k in range(3))
        else:
            largest = max(range(4), key=lambda k: (abs(q[k]), -k))
            negate = 1 if q[largest] < 0 else 0
            comp = largest
            for k in range(4):
                if k == largest:
                    continue
                negbit = (1 if q[k] < 0 else 0) ^ negate
                mag = math.floor(511 * (abs(q[k]) / math.sqrt(0.5)) + 0.5)
                comp = ((comp << 10) | (negbit << 9) | mag) & 0xFFFFFFFF
            out += struct.pack("<I", comp)
    for s in splats:
        for k in range(dim):
            out += bytes(u8(s["rest"][c][k] * (FLIP[k] if k < 15 else 1) * 128 + 128) for c in range(3))
    return bytes(out) if raw else gzip.compress(bytes(out))


def zip_stored(files):
    """A zip of the given files, stored, as the reference test's zip() (test/splats.test.js:227-262)."""
    import io
    import zipfile
    b = io.BytesIO()
    with zipfile.ZipFile(b, "w", zipfile.ZIP_STORED) as z:
        for name, data in files.items():
            z.writestr(name, data)
    return b.getvalue()


def decode_image(data):
    """The reference test's decodeImage: the "images" are their pixels as JSON."""
    d = json.loads(bytes(data).decode())
    return d["width"], d["height"], np.array(d["rgba"], np.uint8)


def as_file(rgba, width):
    return json.dumps({"width": width, "height": len(rgba) // 4 // width, "rgba": [int(v) for v in rgba]}).encode()


def sog_fixture():
    """The reference test's one-splat .sog and the .ply row it was made from (test/splats.test.js:264-320)."""
    scale_book = [-6 + k * 0.05 for k in range(256)]
    colour_book = [-2 + k / 64 for k in range(256)]
    sh_book = [-1 + k / 128 for k in range(256)]
    mins, maxs = [-1, -2, -1], [1, 2, 1]
    log = lambda x: math.copysign(math.log(abs(x) + 1), x)
    position = [1.5, -2.25, 0.75]
    stored = [js_round((log(x) - mins[a]) / (maxs[a] - mins[a]) * 65535) for a, x in enumerate(position)]
    r = [0.9, 0.3, -0.2, 0.25]
    ln = math.sqrt(sum(v * v for v in r))
    rotation = [v / ln for v in r]
    quat_bytes = [js_round((c * math.sqrt(2) + 1) / 2 * 255) for c in rotation[1:]]
    scale_index, colour_index = [100, 120, 130], [140, 100, 70]
    images = {
        "means_l.webp": [stored[0] & 255, stored[1] & 255, stored[2] & 255, 255],
        "means_u.webp": [stored[0] >> 8, stored[1] >> 8, stored[2] >> 8, 255],
        "quats.webp": quat_bytes + [252],
        "scales.webp": scale_index + [255],
        "sh0.webp": colour_index + [200],
        "shN_labels.webp": [1, 0, 0, 255],
    }
    centroids = np.zeros(192 * 4, np.uint8)
    rest_index = [[160, 64, 128], [128, 176, 128], [96, 128, 192]]
    for k in range(3):
        centroids[(3 + k) * 4:(4 + k) * 4] = [rest_index[0][k], rest_index[1][k], rest_index[2][k], 255]
    meta = {"version": 2, "count": 1,
            "means": {"mins": mins, "maxs": maxs, "files": ["means_l.webp", "means_u.webp"]},
            "scales": {"codebook": scale_book, "files": ["scales.webp"]}, "quats": {"files": ["quats.webp"]},
            "sh0": {"codebook": colour_book, "files": ["sh0.webp"]},
            "shN": {"count": 2, "bands": 1, "codebook": sh_book, "files": ["shN_centroids.webp", "shN_labels.webp"]}}
    files = {"meta.json": json.dumps(meta).encode()}
    files.update({k: as_file(v, 1) for k, v in images.items()})
    files["shN_centroids.webp"] = as_file(centroids, 192)
    row = {"x": position[0], "y": position[1], "z": position[2], "scale_0": scale_book[100],
           "scale_1": scale_book[120], "scale_2": scale_book[130], "rot_0": rotation[0], "rot_1": rotation[1],
           "rot_2": rotation[2], "rot_3": rotation[3], "opacity": math.log(200 / 55), "f_dc_0": colour_book[140],
           "f_dc_1": colour_book[100], "f_dc_2": colour_book[70]}
    for c in range(3):
        for k in range(3):
            row["f_rest_%d" % (c * 3 + k)] = sh_book[rest_index[c][k]]
    return meta, files, row


def sog_images(meta, files):
    return {name: decode_image(data) for name, data in files.items() if name != "meta.json"}


# ---- the oracle, pinned by the reference's fixtures ----

def test_ply_as_training_writes_it():
    # test/splats.test.js:73-88
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1",
             "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    s = oracle.decode_ply(ply(props, [
        {"x": -1, "y": 0, "z": 4, "f_dc_0": (1 - 0.5) / SH_C0, "f_dc_1": 0, "f_dc_2": (0 - 0.5) / SH_C0,
         "opacity": 0, "scale_0": math.log(2), "scale_1": math.log(3), "scale_2": 0, "rot_0": 2},
        {"x": 5, "y": -2, "z": 0, "rot_0": 1}]))
    assert s["count"] == 2 and s["degree"] == 0
    assert colour_of(s, 0) == [255, 128, 0, 128]
    c = s["covariances"][0]
    assert abs(c[0] - 4) < 1e-4 and abs(c[3] - 9) < 1e-4 and abs(c[5] - 1) < 1e-4
    assert list(s["centers"][:, :3].min(0)) == [-1, -2, 0] and list(s["centers"][:, :3].max(0)) == [5, 0, 4]


def test_ply_harmonics_degree_from_how_many():
    # test/splats.test.js:98-118
    rest = ["f_rest_%d" % k for k in range(9)]
    row = {"rot_0": 1}
    for k in range(3):
        row["f_rest_%d" % k], row["f_rest_%d" % (k + 3)], row["f_rest_%d" % (k + 6)] = k, 10 + k, -(k + 1)
    s = oracle.decode_ply(ply(NEEDS + rest, [row, row]))
    assert s["degree"] == 1 and s["sh"].shape == (2, 5)
    halves = s["sh"].view(np.uint16).reshape(2, 10)
    for i in (0, 1):
        for k in range(3):
            assert [from_half(int(h)) for h in halves[i, 3 * k:3 * k + 3]] == [k, 10 + k, -(k + 1)]
    with pytest.raises(oracle.DecodeError, match="2 f_rest_ properties; harmonics of degree 1, 2 or 3 have 9, 24 or 45"):
        oracle.decode_ply(ply(NEEDS + ["f_rest_0", "f_rest_1"], [{}]))


@pytest.mark.parametrize("data, message", [
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\n", "no end_header"),
    (ply(NEEDS, [{}]).replace(b"binary_little_endian", b"ascii"), "in 'ascii' format"),
    (b"ply\nformat binary_little_endian 1.0\nelement chunk 1\nend_header\n", "compressed .ply"),
    (b"ply\nformat binary_little_endian 1.0\nelement face 1\nend_header\n", "first element is 'face'"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty list uchar int v\nend_header\n", "list properties"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty half x\nend_header\n", "unknown .ply property type 'half'"),
    (ply(NEEDS[:5], [{}]), "not a splat capture: no f_dc_2, opacity, scale_0"),
    (ply(NEEDS, [{}, {}])[:-8], "cut short: 2 vertices need 112 bytes, it has 104"),
])
def test_ply_refusals(data, message):
    with pytest.raises(oracle.DecodeError, match=message.replace("(", r"\(").replace(".", r"\.")):
        oracle.decode_ply(data)


SPLAT = {"position": [1.5, -2.25, 0.75], "logScale": [-1, 0, 0.5], "rotation": [0.9, 0.3, -0.2, 0.25],
         "opacity": 1.4, "dc": [(0.7 - 0.5) / SH_C0, (0.4 - 0.5) / SH_C0, (0.2 - 0.5) / SH_C0],
         "rest": [[0.25, -0.5, 0.125], [0, 0.375, 0], [-0.25, 0, 0.5]]}


def splat_row():
    row = {"x": 1.5, "y": -2.25, "z": 0.75, "scale_0": -1, "scale_1": 0, "scale_2": 0.5, "rot_0": 0.9, "rot_1": 0.3,
           "rot_2": -0.2, "rot_3": 0.25, "opacity": 1.4, "f_dc_0": SPLAT["dc"][0], "f_dc_1": SPLAT["dc"][1],
           "f_dc_2": SPLAT["dc"][2]}
    for k in range(9):
        row["f_rest_%d" % k] = SPLAT["rest"][k // 3][k % 3]
    return row


@pytest.mark.parametrize("version", [1, 2, 3])
def test_spz_reads_as_the_ply_it_was_packed_from(version):
    # test/splats.test.js:183-215
    want = oracle.decode_ply(ply(NEEDS + ["f_rest_%d" % k for k in range(9)], [splat_row()]))
    got = oracle.decode_spz(spz(version, [SPLAT, SPLAT], degree=1))
    assert got["count"] == 2 and got["degree"] == 1
    eps = 2e-3 if version == 1 else 1e-3
    assert np.all(np.abs(got["centers"][1, :3] - want["centers"][0, :3]) <= eps)
    assert all(abs(a - b) <= 2 for a, b in zip(colour_of(got, 1), colour_of(want, 0)))
    assert np.all(np.abs(got["covariances"][1] - want["covariances"][0]) <= 0.03)
    gh = got["sh"].view(np.uint16).reshape(2, 10)[1, :9]
    wh = want["sh"].view(np.uint16).reshape(1, 10)[0, :9]
    assert all(abs(from_half(int(a)) - from_half(int(b))) <= 1 / 128 for a, b in zip(gh, wh))


def test_spz_refusals():
    with pytest.raises(oracle.DecodeError, match="gzipped, but not a .spz"):
        oracle.decode_spz(gzip.compress(bytes(32)))
    short = gzip.decompress(spz(3, [SPLAT]))[:30]
    with pytest.raises(oracle.DecodeError, match="the .spz is cut short"):
        oracle.decode_spz(gzip.compress(short))
    with pytest.raises(oracle.DecodeError, match="a .spz of version 5; versions 1 to 3 are read"):
        oracle.decode_spz(gzip.compress(struct.pack("<IIIBBBB", 0x5053474E, 5, 1, 0, 12, 0, 0) + bytes(64)))


def test_sog_reads_as_the_ply_it_was_made_from():
    # test/splats.test.js:264-320
    meta, files, row = sog_fixture()
    got = oracle.decode_sog(meta, sog_images(meta, files))
    want = oracle.decode_ply(ply(NEEDS + ["f_rest_%d" % k for k in range(9)], [row]))
    assert got["count"] == 1 and got["degree"] == 1
    assert np.all(np.abs(got["centers"][0, :3] - want["centers"][0, :3]) <= 1e-3)
    assert colour_of(got, 0) == colour_of(want, 0)
    assert np.all(np.abs(got["covariances"][0] - want["covariances"][0]) <= 0.03)
    assert list(got["sh"].view(np.uint16)[0, :9]) == list(want["sh"].view(np.uint16)[0, :9])


def test_half_bits_round_half_up_and_saturate():
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.normal(0, 3, 3000), rng.uniform(-1e-5, 1e-5, 1000)]).astype(np.float32).astype(np.float64)
    want = v.astype(np.float16).view(np.uint16)
    got = np.array([oracle.half(x) for x in v], np.uint16)
    # numpy rounds ties to even; a tie is a float32 whose 13 bits below the half's mantissa are 1000...0
    bits = v.astype(np.float32).view(np.uint32)
    tie = (bits & 0x1FFF) == 0x1000
    # below 2^-24 halfBits gives 0 where IEEE rounds [2^-25, 2^-24) up to the least subnormal
    tiny = np.abs(v) < 2.0 ** -24
    assert tiny.sum() >= 1 and np.all(got[tiny] & 0x7FFF == 0)
    assert np.array_equal(got[~tie & ~tiny], want[~tie & ~tiny])
    assert oracle.half(1 + 2 ** -11) == 0x3C01 and np.float16(1 + 2 ** -11).view(np.uint16) == 0x3C00
    assert oracle.half(70000) == 0x7BFF and oracle.half(-70000) == 0xFBFF
    assert oracle.half(float("nan")) == 0 and oracle.half(2 ** -25) == 0 and oracle.half(3 * 2 ** -25) == 2
    assert oracle.half(0.6) == 0x38CD  # the value test/gpu.test.js:1968-2035 names


# ---- the GPU, held to the oracle ----

def random_ply(n, degree, seed, types=None):
    rng = np.random.default_rng(seed)
    rest = ["f_rest_%d" % k for k in range(3 * [0, 3, 8, 15][degree])]
    props = NEEDS + rest
    rows = []
    for _ in range(n):
        q = rng.normal(size=4)
        row = dict(zip(["x", "y", "z"], rng.normal(0, 3, 3)))
        row.update({"f_dc_%d" % k: rng.normal(0, 1.5) for k in range(3)})
        row.update({"scale_%d" % k: rng.normal(-3, 1) for k in range(3)})
        row.update({"rot_%d" % k: q[k] for k in range(4)})
        row["opacity"] = rng.normal(0, 3)
        row.update({r: rng.normal(0, 0.3) for r in rest})
        rows.append(row)
    # harmonics at the edges of halfBits: ties (half up), both ends of the subnormals, saturation, NaN
    edges = [1 + 2 ** -11, -(1 + 2 ** -11), 0.5 + 2 ** -12, 2 ** -14 * (1 - 2 ** -11), 3 * 2 ** -25, 2 ** -25,
             65503.99, 65520.0, 1e9, -1e9, float("nan"), 0.0, -0.0, 2 ** -24]
    for j, r in enumerate(rest):
        rows[j % len(rows)][r] = edges[j % len(edges)]
    return ply(props, rows, types)


def check_cloud(cloud, want):
    centers, covs, sh, degree = cloud.read()
    assert cloud.count == want["count"] and degree == want["degree"]
    assert np.array_equal(centers.view(np.uint32), want["centers"].view(np.uint32))
    scale = np.maximum(np.abs(want["covariances"][:, [0, 3, 5]]).max(1, keepdims=True), 1e-30)
    err = np.abs(covs.astype(np.float64) - want["covariances"]) / scale
    exact = np.mean(covs.view(np.uint32) == want["covariances"].view(np.uint32))
    assert err.max() <= 1e-6, err.max()
    if degree:
        assert np.array_equal(sh, want["sh"])
    return float(exact)


@pytest.mark.gpu
@pytest.mark.parametrize("degree, types", [(0, None), (1, None), (2, None), (3, None),
                                           (1, {"x": "double", "opacity": "double", "f_dc_0": "short", "rot_0": "int"})])
def test_gpu_ply_matches_oracle(ctx, degree, types):
    data = random_ply(3000, degree, seed=degree, types=types)
    want = oracle.decode_ply(data)
    cloud = wa.decode_ply(ctx, data)
    try:
        exact = check_cloud(cloud, want)
        print("ply degree %d: covariance floats bit-equal %.4f" % (degree, exact), flush=True)
        assert np.array_equal(cloud.min, want["centers"][:, :3].min(0))
        assert np.array_equal(cloud.max, want["centers"][:, :3].max(0))
    finally:
        cloud.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("version, degree", [(1, 0), (2, 2), (3, 3), (3, 1), (3, 4), (2, 4)])
def test_gpu_spz_matches_oracle(ctx, version, degree):
    """2001 splats: the sections start at every offset modulo 4 (the decoder stages each workgroup's
    share of a section from the word holding its first byte), and degree 4 gives the widest rows (72
    bytes of harmonics a splat, of which the first 45 are read)."""
    rng = np.random.default_rng(version * 10 + degree)
    splats = []
    for _ in range(2001):
        splats.append({"position": list(rng.normal(0, 3, 3)), "logScale": list(rng.normal(-3, 1, 3)),
                       "rotation": list(rng.normal(size=4)), "opacity": rng.normal(0, 2),
                       "dc": list(rng.normal(0, 1.5, 3)), "rest": [list(rng.normal(0, 0.3, 24)) for _ in range(3)]})
    for raw in (False, True):
        data = spz(version, splats, degree=degree, raw=raw)
        want = oracle.decode_spz(data)
        cloud = wa.decode_spz(ctx, data)
        try:
            check_cloud(cloud, want)
        finally:
            cloud.destroy()


@pytest.mark.gpu
def test_gpu_sog_matches_oracle_and_the_reference_fixture(ctx):
    meta, files, row = sog_fixture()
    cloud = wa.read_sog(ctx, zip_stored(files), decode_image)
    try:
        check_cloud(cloud, oracle.decode_sog(meta, sog_images(meta, files)))
    finally:
        cloud.destroy()
    # a larger .sog: random images, the same codebooks and palette
    rng = np.random.default_rng(8)
    n = 5000
    meta = dict(meta, count=n)
    meta["shN"] = dict(meta["shN"], count=40)
    big = {"meta.json": json.dumps(meta).encode()}
    for name in ("means_l.webp", "means_u.webp", "scales.webp", "sh0.webp"):
        big[name] = as_file(rng.integers(0, 256, 4 * n), 50)
    quats = rng.integers(0, 256, (n, 4))
    quats[:, 3] = rng.integers(250, 256, n)
    big["quats.webp"] = as_file(quats.ravel(), 50)
    big["shN_labels.webp"] = as_file(np.c_[rng.integers(0, 48, n), np.zeros((n, 3), int)].ravel(), 50)
    big["shN_centroids.webp"] = as_file(rng.integers(0, 256, 192 * 4), 192)
    cloud = wa.read_sog(ctx, zip_stored(big), decode_image)
    try:
        check_cloud(cloud, oracle.decode_sog(meta, sog_images(meta, big)))
    finally:
        cloud.destroy()
    with pytest.raises(wa.WindingError, match="needs decodeImage"):
        wa.read_sog(ctx, zip_stored(files))
    with pytest.raises(wa.WindingError, match="version 1; version 2 is read"):
        wa.read_sog(ctx, zip_stored({"meta.json": b'{"means": {}}'}), decode_image)
    with pytest.raises(wa.WindingError, match="has no means_l.webp"):
        wa.read_sog(ctx, zip_stored({"meta.json": files["meta.json"]}), decode_image)


@pytest.mark.gpu
def test_gpu_box_and_first_bad_over_many_workgroups(ctx):
    """600,001 vertices: 4,688 workgroups' boxes, more than one pass of the box reduce (4,096 boxes),
    equal to the centres' min and max; then non-finite positions in two late vertices, in different
    workgroups, of which the first is the one reported."""
    n = 600001
    rng = np.random.default_rng(11)
    v = np.zeros((n, len(NEEDS)), np.float32)
    v[:, :3] = rng.normal(0, 3, (n, 3))
    v[:, 3:7] = rng.normal(0, 1, (n, 4))
    v[:, 7:10] = rng.normal(-3, 1, (n, 3))
    v[:, 10:14] = rng.normal(0, 1, (n, 4))
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    head = (head + "".join("property float %s\n" % p for p in NEEDS) + "end_header\n").encode()
    cloud = wa.decode_ply(ctx, head + v.tobytes())
    try:
        assert cloud.count == n
        assert np.array_equal(np.asarray(cloud.min, np.float32), v[:, :3].min(0))
        assert np.array_equal(np.asarray(cloud.max, np.float32), v[:, :3].max(0))
    finally:
        cloud.destroy()
    v[555555, 1] = np.nan
    v[590000, 0] = np.inf
    with pytest.raises(wa.WindingError, match="splat 555555 has a position that is not a number"):
        wa.decode_ply(ctx, head + v.tobytes())


@pytest.mark.gpu
def test_gpu_decoder_refusals(ctx):
    with pytest.raises(wa.WindingError, match="no end_header"):
        wa.decode_ply(ctx, b"ply\nformat binary_little_endian 1.0\n")
    with pytest.raises(wa.WindingError, match="2 f_rest_ properties"):
        wa.decode_ply(ctx, ply(NEEDS + ["f_rest_0", "f_rest_1"], [{}]))
    with pytest.raises(wa.WindingError, match="cut short: 2 vertices need 112 bytes, it has 104"):
        wa.decode_ply(ctx, ply(NEEDS, [{}, {}])[:-8])
    with pytest.raises(wa.WindingError, match="splat 1 has a position that is not a number"):
        wa.decode_ply(ctx, ply(NEEDS, [{}, {"y": float("nan")}]))
    with pytest.raises(wa.WindingError, match="version 4, compressed with zstd"):
        wa.decode_spz(ctx, struct.pack("<II", 0x5053474E, 4) + bytes(56))
    with pytest.raises(wa.WindingError, match="gzipped, but not a .spz"):
        wa.decode_spz(ctx, gzip.compress(bytes(32)))
    with pytest.raises(wa.WindingError, match="the .spz is cut short"):
        wa.decode_spz(ctx, gzip.compress(gzip.decompress(spz(3, [SPLAT]))[:30]))


@pytest.mark.gpu
@pytest.mark.parametrize("cull", [0.5, 0.0])
def test_gpu_sort_of_a_decoded_cloud_is_bit_exact(ctx, cull):
    # a decoded .ply of 20,000 splats, sorted from a camera at (0, 0, 12): the Sigma form of the projection
    data = random_ply(20000, 3, seed=5)
    cloud = wa.decode_ply(ctx, data)
    st = wa.State(ctx, cloud.count)
    try:
        centers, covs, sh, degree = cloud.read()
        v, p = wa.workloads.camera((0.0, 1.0, 12.0), 1280, 720)
        view = wa.prepare_view(v, p, 1280.0, 720.0, cloud.min, cloud.max, cull)
        st.sort(cloud, view)
        order, cov, stats = st.read(n_cov=cloud.count)
        want, _, keep, wcov = oracle.run_sigma(centers, covs, view)
        print("cull %.1f: %d of %d kept" % (cull, stats.count, cloud.count), flush=True)
        assert stats.count == len(want) and np.array_equal(order, want)
        s = np.maximum(np.abs(wcov[keep][:, [0, 2]]).max(1, keepdims=True), 1e-30)
        assert np.max(np.abs(cov[keep, :3] - wcov[keep]) / s) <= 1e-6
        # the same centres, Sigma and harmonics given from the host sort and shade the same
        host = wa.Cloud.from_splats(ctx, centers, covs)
        try:
            host.set_sh(sh, degree)
            st.sort(host, view)
            order2, cov2, _ = st.read(n_cov=cloud.count)
            assert np.array_equal(order2, order) and np.array_equal(cov2.view(np.uint32), cov.view(np.uint32))
        finally:
            host.destroy()
        # one batch of every form: the decoded cloud (centres, Sigma, harmonics), the same without
        # harmonics, and .splat records; each job's output equals its single sort's
        plain = wa.Cloud.from_splats(ctx, centers, covs)
        recs = wa.Cloud(ctx, wa.workloads.plain("plain-64k", n=30000)[0])
        jobs = [(cloud, view), (plain, view),
                (recs, wa.prepare_view(v, p, 1280.0, 720.0, recs.min, recs.max, cull))]
        states = [wa.State(ctx, c.count) for c, _ in jobs]
        b = wa.Batch(ctx, len(jobs))
        try:
            want = []
            for (c, vw), s in zip(jobs, states):
                s.sort(c, vw)
                o, cv, _ = s.read(n_cov=c.count)
                want.append((o, cv.view(np.uint32).copy()))
            b.set([(c, vw, s) for (c, vw), s in zip(jobs, states)])
            for run in range(2):
                b.run()
                for ((c, vw), s), (o, cv) in zip(zip(jobs, states), want):
                    got_o, got_c, _ = s.read(n_cov=c.count)
                    assert np.array_equal(got_o, o) and np.array_equal(got_c.view(np.uint32), cv), run
            # the harmonics changed since the jobs were set: refused until they are set again
            plain.set_sh(sh, degree)
            with pytest.raises(wa.WindingError, match="harmonics changed"):
                b.run()
        finally:
            b.destroy()
            for s in states:
                s.destroy()
            plain.destroy()
            recs.destroy()
    finally:
        st.destroy()
        cloud.destroy()
