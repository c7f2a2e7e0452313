"""Pins the CPU oracle: the reference's known answers, closed forms, invariants and brute force.

Known answers restated from the reference's own tests (Winding v1.9.0):
  test/splats.test.js:51-61, :63-71   a record's covariance, unturned and a quarter turn
  test/splats.test.js:338-339         a NaN position is refused
  test/core.test.js:214-223, :343-360 lookAt puts the target on -Z; reverse-Z near->1, far->0
  test/render.test.js:428-440         counting sort keeps input order within a bucket
  test/gpu.test.js:1804-1870          two splats on the view axis, drawn back to front from either side
  test/gpu.test.js:1914-1962          a faint speck culled at 0.5, kept at 0
"""
import math
import struct

import numpy as np
import pytest

import oracle
import winding_amd as wa


def record(pos, scale, rgba, quat):
    return np.frombuffer(struct.pack("<6f4B4B", *pos, *scale, *rgba, *quat), np.uint8).copy()


def cloud(*recs):
    return np.concatenate(recs).reshape(-1, 32)


def view_of(eye, width, height, fovy, recs, cull=0.5, target=(0, 0, 0)):
    v = oracle.look_at(eye, target)
    p = oracle.perspective(fovy, width / height, 0.1)
    bad, mn, mx = oracle.bounds(recs)
    assert bad == -1
    return oracle.prepare(v, p, width, height, mn, mx, cull)


def test_sigma_unturned_is_scale_squared():
    # test/splats.test.js:51-61: diag(0.25, 4, 1) for scales (0.5, 2, 1), no turn.
    s = oracle.sigma(record([1, 2, 3], [0.5, 2, 1], [10, 20, 30, 200], [255, 128, 128, 128]))
    assert list(s) == [0.25, 0.0, 0.0, 4.0, 0.0, 1.0]


def test_sigma_quarter_turn_swaps_axes():
    # test/splats.test.js:63-71: w = z = cos 45deg turns the long axis from y to x.
    b = 128 + round(math.sqrt(0.5) * 128)
    s = oracle.sigma(record([0, 0, 0], [0.5, 2, 1], [0, 0, 0, 255], [b, 128, 128, b]))
    xx, xy, _, yy, _, _ = s
    assert abs(xx - 4) < 1e-12 and abs(yy - 0.25) < 1e-12 and abs(xy) < 1e-12


def test_sigma_is_r_s_s_r_for_any_quaternion():
    # Closed form: Sigma = R diag(s^2) R^T with R the rotation of the unit quaternion (w,x,y,z),
    # built here from the axis-angle Rodrigues formula, a different route than put()'s.
    rng = np.random.default_rng(1)
    for _ in range(300):
        q = rng.integers(0, 256, 4)
        s = np.exp(rng.normal(-2, 1, 3)).astype(np.float32)
        rec = record([0, 0, 0], s, [0, 0, 0, 255], q)
        got = oracle.sigma(rec)
        qq = (q.astype(np.float64) - 128) / 128
        n = np.linalg.norm(qq)
        if n == 0:
            R = np.eye(3)
        else:
            w, v = qq[0] / n, qq[1:] / n
            angle = 2 * math.atan2(np.linalg.norm(v), w)
            axis = v / np.linalg.norm(v) if np.linalg.norm(v) > 0 else np.array([1.0, 0, 0])
            K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
            R = np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K
        S = R @ np.diag(s.astype(np.float64) ** 2) @ R.T
        want = [S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]]
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * float(np.max(s)) ** 2)


def test_zero_quaternion_is_identity():
    # put(): Math.hypot(...) || 1 -- a zero quaternion leaves R = I.
    s = oracle.sigma(record([0, 0, 0], [1, 2, 3], [0, 0, 0, 255], [128, 128, 128, 128]))
    assert list(s) == [1.0, 0.0, 0.0, 4.0, 0.0, 9.0]


def test_look_at_puts_target_on_minus_z():
    # test/core.test.js:214-223
    m = oracle.look_at((0, 0, 5), (0, 0, 0)).astype(np.float64).reshape(4, 4).T
    np.testing.assert_allclose(m @ [0, 0, 0, 1], [0, 0, -5, 1], atol=1e-6)
    np.testing.assert_allclose(m @ [0, 0, 5, 1], [0, 0, 0, 1], atol=1e-6)


def test_reverse_z_near_one_far_zero():
    # test/core.test.js:343-360
    p = oracle.perspective(math.pi / 3, 16 / 9, 0.1).astype(np.float64).reshape(4, 4).T
    ndc = lambda d: (p @ [0, 0, -d, 1])[2] / (p @ [0, 0, -d, 1])[3]
    assert abs(ndc(0.1) - 1) < 1e-6 and abs(ndc(1e9)) < 1e-6
    zs = [ndc(d) for d in [0.1, 0.5, 1, 10, 100, 1000, 100000]]
    assert all(a > b for a, b in zip(zs, zs[1:]))


def test_python_camera_helpers_match_oracle():
    for eye in [(0, 0, 10), (3, -2, 7.5), (-11, 4, 0.25)]:
        assert np.array_equal(wa.look_at(eye), oracle.look_at(eye))
    assert np.array_equal(wa.perspective(math.pi / 3, 16 / 9, 0.1), oracle.perspective(math.pi / 3, 16 / 9, 0.1))


def test_counting_sort_keeps_input_order_within_a_bucket():
    # test/render.test.js:428-440, verbatim keys and expected payload order.
    order = oracle.sort(np.array([7, 3, 7, 3, 7, 1], np.uint16), np.ones(6, bool))
    assert list(order) == [5, 1, 3, 0, 2, 4]


def test_sort_matches_stable_argsort_with_collisions_and_culls():
    rng = np.random.default_rng(2)
    for n, span in [(1, 5), (17, 3), (1000, 50), (20000, 65536), (50000, 7)]:
        k = rng.integers(0, span, n).astype(np.uint16)
        keep = rng.random(n) < 0.7
        got = oracle.sort(k, keep)
        idx = np.nonzero(keep)[0]
        want = idx[np.argsort(k[idx], kind="stable")]
        assert np.array_equal(got, want)


def red_blue():
    red = record([0, 0, 1], [0.4, 0.4, 0.4], [255, 0, 0, 255], [255, 128, 128, 128])
    blue = record([0, 0, -1], [0.4, 0.4, 0.4], [0, 0, 255, 255], [255, 128, 128, 128])
    return red, blue


@pytest.mark.parametrize("file_order", ["red first", "blue first"])
def test_back_to_front_follows_the_camera(file_order):
    # test/gpu.test.js:1804-1870, as orders: from (0,0,5) blue (farther) then red; from (0,0,-5) red then blue.
    red, blue = red_blue()
    recs = cloud(red, blue) if file_order == "red first" else cloud(blue, red)
    ri, bi = (0, 1) if file_order == "red first" else (1, 0)
    order, _, _, _ = oracle.run(recs, view_of((0, 0, 5), 160, 120, 0.8, recs))
    assert list(order) == [bi, ri]
    order, _, _, _ = oracle.run(recs, view_of((0, 0, -5), 160, 120, 0.8, recs))
    assert list(order) == [ri, bi]


def test_single_splat_spans_no_depth():
    # test/gpu.test.js:1860-1870: one splat alone still draws; its cloud spans no depth (depthScale floored).
    red, _ = red_blue()
    recs = cloud(red)
    v = view_of((0, 0, 5), 160, 120, 0.8, recs)
    assert v.depth_scale == np.float32(65535 / 1e-6)
    order, k, keep, _ = oracle.run(recs, v)
    assert list(order) == [0] and keep[0] and k[0] == 65535


def test_faint_speck_culled_at_default_kept_at_zero():
    # test/gpu.test.js:1914-1962: scale 5e-4, alpha 50 at 160x120, fovY 0.8, from (0,0,5):
    # 50/255 of the 1.9 pixels the draw gives it, 0.37 < 0.5. The large opaque splat stays.
    speck = record([0, 0, 0], [0.0005] * 3, [255, 255, 255, 50], [255, 128, 128, 128])
    big = record([1, 0, 0], [0.2] * 3, [255, 255, 255, 255], [255, 128, 128, 128])
    recs = cloud(speck, big)
    k, keep, cov, val = oracle.keys(recs, view_of((0, 0, 5), 160, 120, 0.8, recs, cull=0.5))
    assert list(keep) == [False, True]
    assert abs(val[0] - 2 * math.pi * 0.3 * 50 / 255) < 1e-3  # the speck adds 0.37 of a pixel
    k, keep, _, val = oracle.keys(recs, view_of((0, 0, 5), 160, 120, 0.8, recs, cull=0.0))
    assert list(keep) == [True, True] and math.isnan(val[0])


def test_key_buckets_and_edges_closed_form():
    # view = identity (eye at origin looking down -z), depthMin 0, depthScale 1: depth d lands in
    # bucket floor(d) clamped to 65535, key 65535 - bucket; an edge belongs to the bucket above it.
    v = oracle.OView()
    for i, x in enumerate([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]):
        v.model_view[i] = x
    for i, x in enumerate(oracle.perspective(math.pi / 2, 1.0, 0.1)):
        v.projection[i] = x
    v.viewport[0] = v.viewport[1] = 100
    v.depth_min, v.depth_scale, v.cull = 0.0, 1.0, 0.0
    below3 = float(np.nextafter(np.float32(3.0), np.float32(0)))
    cases = [(3.0, 3), (below3, 2), (0.25, 0), (70000.0, 65535), (65535.0, 65535), (65534.5, 65534)]
    for depth, bucket in cases:
        rec = record([0, 0, -depth], [0.1] * 3, [0, 0, 0, 255], [255, 128, 128, 128])
        k, keep, _, _ = oracle.keys(rec, v)
        assert keep[0] and k[0] == 65535 - bucket, (depth, k[0])
    # Behind the eye, nearer than near, outside 1.3 w, alpha <= 1: culled.
    for pos, a in [([0, 0, 1], 255), ([0, 0, -0.05], 255), ([2, 0, -1], 255), ([0, 0, -5], 1)]:
        k, keep, _, _ = oracle.keys(record(pos, [0.1] * 3, [0, 0, 0, a], [255, 128, 128, 128]), v)
        assert not keep[0], pos


def test_keys_monotone_in_depth():
    recs, view, _ = wa.workloads.plain("plain-64k", n=4096)
    k, keep, _, _ = oracle.keys(recs, view)
    pos = recs.view(np.float32).reshape(-1, 8)[:, :3].astype(np.float64)
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    depth = -(pos @ mv[2, :3] + mv[2, 3])
    idx = np.nonzero(keep)[0]
    by_depth = idx[np.argsort(depth[idx], kind="stable")]
    assert np.all(np.diff(k[by_depth].astype(np.int64)) <= 0)  # farther, smaller key


def test_cov_matches_closed_form_jacobian():
    # cov = J W Sigma W^T J^T with J the projection's derivative, here by central differences of the
    # pixel mapping u = fx * x / d (unclamped region), a route independent of footprint().
    recs, view, _ = wa.workloads.plain("plain-64k", n=512, cull=0.0)
    k, keep, cov, _ = oracle.keys(recs, view)
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    P = np.array(view.projection, np.float64)
    fx, fy = P[0] * view.viewport[0] * 0.5, P[5] * view.viewport[1] * 0.5
    checked = 0
    for i in np.nonzero(keep)[0][:200]:
        p = recs[i].view(np.float32)[:3].astype(np.float64)
        vp = mv[:3, :3] @ p + mv[:3, 3]
        if abs(vp[0] / -vp[2]) > 1.3 / P[0] or abs(vp[1] / -vp[2]) > 1.3 / P[5]:
            continue
        def pix(x):
            v = mv[:3, :3] @ x + mv[:3, 3]
            return np.array([fx * v[0] / -v[2], fy * v[1] / -v[2]])
        h = 1e-5
        J = np.stack([(pix(p + h * e) - pix(p - h * e)) / (2 * h) for e in np.eye(3)], axis=1)
        s = oracle.sigma(recs[i])
        S = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
        C2 = J @ S @ J.T
        want = [C2[0, 0], C2[0, 1], C2[1, 1]]
        np.testing.assert_allclose(cov[i], want, rtol=2e-5, atol=2e-5 * max(abs(C2[0, 0]), abs(C2[1, 1])))
        checked += 1
    assert checked > 100


def test_bounds_and_nan_refusal():
    # test/splats.test.js:338-339 (NaN position refused) and put()'s min/max.
    a = record([1, 2, 3], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128])
    b = record([-4, 5, 0.5], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128])
    bad, mn, mx = oracle.bounds(cloud(a, b))
    assert bad == -1 and list(mn) == [-4, 2, 0.5] and list(mx) == [1, 5, 3]
    nan = record([float("nan"), 0, 0], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128])
    assert oracle.bounds(cloud(a, nan))[0] == 1


def test_prepare_depth_range():
    # Camera at (0,0,10) on a [-1,1]^3 box: near 9, far 11, depthScale 65535/2; wholly behind: skipped.
    v = oracle.look_at((0, 0, 10))
    p = oracle.perspective(math.pi / 3, 16 / 9, 0.1)
    out = oracle.prepare(v, p, 1280, 720, [-1, -1, -1], [1, 1, 1], 0.5)
    assert out.depth_min == 9 and out.depth_scale == np.float32(65535 / 2)
    behind = oracle.look_at((0, 0, 10), (0, 0, 20))
    assert oracle.prepare(behind, p, 1280, 720, [-1, -1, -1], [1, 1, 1], 0.5) is None


def test_bounds_keep_the_first_of_equal_values():
    # put() updates min/max with strict compares in splat order (src/scene/splats.js:343-348), so of
    # equal values the first stays: a zero end takes the sign of the first zero met.
    def box(xs, ys):
        recs = cloud(*[record([x, y, 1.0], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128]) for x, y in zip(xs, ys)])
        bad, mn, mx = oracle.bounds(recs)
        assert bad == -1
        return mn.view(np.uint32), mx.view(np.uint32)
    mn, mx = box([0.0, -0.0, 3.0], [-3.0, -0.0, 0.0])
    assert mn[0] == 0x00000000 and mx[1] == 0x80000000
    mn, mx = box([-0.0, 0.0, 3.0], [-3.0, 0.0, -0.0])
    assert mn[0] == 0x80000000 and mx[1] == 0x00000000
    mn, mx = box([1.0, -0.0, 0.0, -2.0], [5.0, 0.0, -0.0, 1.0])
    assert np.float32(mn.view(np.float32)[0]) == -2.0 and mx.view(np.float32)[1] == 5.0


def test_cov_clamp_branch_is_the_margin_ray():
    # footprint() clamps x/d to +-1.3/p00 (src/render/splats.js:111-114): a splat past that margin, still
    # inside the frustum because the projection is off-axis (p[8] != 0, as a jittered projection has), has
    # the footprint of the same splat moved onto the margin ray at the same depth. An unclamped Jacobian
    # (central differences of the pixel map) gives a different one, so the check tells the clamp apart.
    v = oracle.look_at((0, 0, 10))
    p = oracle.perspective(math.pi / 3, 16 / 9, 0.1).copy()
    p[8] = 0.6
    p00 = float(p[0])
    lx = float(np.float64(np.float32(1.3)) / p00)
    d = 11.0
    recs, moved = [], []
    for sgn, ratio, yv in [(1, 1.6, 0.4), (1, 1.85, -0.7), (-1, 1.35, 0.0)]:
        if sgn < 0:
            ratio = 0.6  # on the left the frustum ends first at (0.6 - 1.3) / p00: stay inside the margin
        x = sgn * ratio / p00 * d
        q = [160, 100, 140, 120]
        recs.append(record([x, yv, 10.0 - d], [0.05, 0.02, 0.08], [200, 100, 50, 220], q))
        xm = np.float32(sgn * min(ratio / p00, lx) * d)
        moved.append(record([xm, yv, 10.0 - d], [0.05, 0.02, 0.08], [200, 100, 50, 220], q))
    recs = cloud(*recs, *moved)
    view = oracle.prepare(v, p, 1280, 720, [-30, -3, -3], [30, 3, 3], 0.0)
    k, keep, cov, _ = oracle.keys(recs, view)
    assert keep.all()
    n = len(recs) // 2
    for i in range(n):
        np.testing.assert_allclose(cov[i], cov[n + i], rtol=2e-6, atol=2e-6 * max(cov[i][0], cov[i][2]))
    # the unclamped Jacobian at the first splat is far from the clamped footprint
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    P = np.array(view.projection, np.float64)
    fx, fy = P[0] * 1280 * 0.5, P[5] * 720 * 0.5
    pos = recs[0].view(np.float32)[:3].astype(np.float64)

    def pix(x):
        vv = mv[:3, :3] @ x + mv[:3, 3]
        return np.array([fx * vv[0] / -vv[2], fy * vv[1] / -vv[2]])
    J = np.stack([(pix(pos + 1e-5 * e) - pix(pos - 1e-5 * e)) / 2e-5 for e in np.eye(3)], axis=1)
    s = np.array([float(np.float32(t)) for t in oracle.sigma(recs[0])])
    S = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
    free = J @ S @ J.T
    assert abs(free[0, 0] - cov[0][0]) > 0.05 * cov[0][0]


def test_record_form_equals_splats_data_form():
    # The same capture given as .splat records and as the reference's Splats data (centres with the colour
    # bits, Sigma as put() stores it in f32) is culled, keyed and sorted alike: the record form reads Sigma
    # at put()'s f32 rounding. Checked on every committed fixture, including edges_4096, whose survivors sit
    # on the cull threshold, at cull 0.5 and 2.0.
    import glob
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for path in sorted(glob.glob(os.path.join(here, "*.npz"))):
        with np.load(path, allow_pickle=False) as z:
            recs = z["records"]
            vf = z["view"].astype(np.float32)
        centers = np.zeros((len(recs), 4), np.float32)
        centers[:, :3] = recs.view(np.float32).reshape(-1, 8)[:, :3]
        centers.view(np.uint32)[:, 3] = recs[:, 24:28].copy().view(np.uint32).ravel()
        covs = np.array([oracle.sigma(r) for r in recs], np.float64).astype(np.float32)
        for cull in (float(vf[36]), 2.0):
            vf2 = vf.copy()
            vf2[36] = cull
            view = wa.View.from_buffer_copy(vf2.tobytes())
            o1, k1, keep1, c1 = oracle.run(recs, view)
            o2, k2, keep2, c2 = oracle.run_sigma(centers, covs, view)
            assert np.array_equal(keep1, keep2), os.path.basename(path)
            assert np.array_equal(k1[keep1], k2[keep2]) and np.array_equal(o1, o2)
            assert np.array_equal(c1[keep1], c2[keep2])
