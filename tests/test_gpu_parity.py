"""Parity of the HIP path with the CPU oracle, through the C ABI, on an MI355X.

The bar: keys, survivor flags and the draw order bit-exact; cov2d within 1e-5 of the FP64 oracle,
relative to max(|xx|, |yy|) per splat (the xy term can be near zero); the FP64 fallback bit-exact.
"""
import glob
import math
import os
import struct

import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu

COV_TOL = 1e-5


def record(pos, scale, rgba, quat):
    return np.frombuffer(struct.pack("<6f4B4B", *pos, *scale, *rgba, *quat), np.uint8).copy()


def run_gpu(ctx, recs, view, state=None):
    cloud = wa.Cloud(ctx, recs)
    own = state is None
    st = state or wa.State(ctx, len(recs))
    try:
        st.sort(cloud, view)
        order, cov, stats = st.read(n_cov=len(recs))
        keys, keep = st.read_keys(len(recs))
    finally:
        cloud.destroy()
        if own:
            st.destroy()
    return order, keys, keep, cov, stats


def cov_error(cov_gpu, cov64, keep):
    g = cov_gpu[keep, :3].astype(np.float64)
    w = cov64[keep]
    scale = np.maximum(np.abs(w[:, 0]), np.abs(w[:, 2]))
    return float(np.max(np.abs(g - w).max(axis=1) / scale)) if len(g) else 0.0


def assert_parity(ctx, recs, view, label=""):
    order, keys, keep, cov, stats = run_gpu(ctx, recs, view)
    o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
    print("[%s] n=%d V=%d recomputed=%d" % (label, len(recs), stats.count, stats.recomputed), flush=True)
    assert np.array_equal(keep, o_keep), "survivor flags differ at %s" % np.nonzero(keep != o_keep)[0][:10]
    assert np.array_equal(keys[keep], o_keys[o_keep]), "keys differ"
    assert stats.count == len(o_order)
    assert np.array_equal(order, o_order), "order differs first at %d" % int(np.argmax(order != o_order))
    err = cov_error(cov, o_cov, keep)
    assert err <= COV_TOL, "cov2d relative error %g" % err
    colours = cov[keep, 3].view(np.uint32)
    assert np.array_equal(colours, recs[keep, 24:28].copy().view(np.uint32).ravel())
    return stats, err


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures(ctx, path):
    """The committed fixtures (both oracles agreed on them): keys, flags, order bit-exact, cov2d to 1e-5."""
    with np.load(path, allow_pickle=False) as z:
        recs, fk, forder, fcov = z["records"], z["keys"], z["order"], z["cov"]
        view = wa.View.from_buffer_copy(z["view"].astype(np.float32).tobytes())
    order, keys, keep, cov, stats = run_gpu(ctx, recs, view)
    assert np.array_equal(keep, fk >= 0)
    assert np.array_equal(keys[keep].astype(np.int32), fk[keep])
    assert stats.count == len(forder) and np.array_equal(order, forder)
    err = cov_error(cov, np.nan_to_num(fcov), keep)
    print("%s: V=%d cov2d max rel err %.3g" % (os.path.basename(path), stats.count, err), flush=True)
    assert err <= COV_TOL


def test_selftest_dpp_scan_and_f64_rounding(ctx):
    rng = np.random.default_rng(5)
    # every sum of squares a byte quaternion can give (sqrt of S/16384), then random magnitudes
    a = np.concatenate([np.arange(0, 65537) / 16384.0, np.exp(rng.uniform(-40, 40, 100000))])
    b = np.concatenate([np.full(65537, 255.0), np.exp(rng.uniform(-40, 40, 100000))])
    args = np.stack([a, b], axis=1).ravel()
    expect = np.stack([np.sqrt(a), a / b], axis=1).ravel()
    mm = ctx.selftest(args, expect)
    assert list(mm) == [0, 0, 0], mm


@pytest.mark.parametrize("cull", [0.5, 0.0, 2.0])
def test_plain_64k_parity(ctx, cull):
    recs, view, _ = wa.workloads.plain("plain-64k", cull=cull)
    stats, err = assert_parity(ctx, recs, view, "plain-64k cull=%g" % cull)
    print("cov2d max rel err %.3g" % err)
    assert stats.recomputed < len(recs) // 1000


def test_fp64_definition_on_gpu_is_bit_exact(ctx):
    recs, view, _ = wa.workloads.plain("plain-64k", n=20000, cull=0.5)
    cloud = wa.Cloud(ctx, recs)
    try:
        got = cloud.canonical(view)
    finally:
        cloud.destroy()
    k, keep, cov, val = oracle.keys(recs, view)
    passed = got[:, 0] == 1
    tested = ~np.isnan(val)
    assert np.array_equal(passed, tested)
    assert np.array_equal(got[passed, 1].view(np.uint64), val[passed].view(np.uint64))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 255, 257, 4095, 4097, 12289, 70001])
def test_ragged_sizes(ctx, n):
    recs, view, _ = wa.workloads.plain("plain-64k", n=n)
    assert_parity(ctx, recs, view, "n=%d" % n)


def test_back_to_front_known_answer(ctx):
    # test/gpu.test.js:1804-1870 restated as orders.
    red = record([0, 0, 1], [0.4] * 3, [255, 0, 0, 255], [255, 128, 128, 128])
    blue = record([0, 0, -1], [0.4] * 3, [0, 0, 255, 255], [255, 128, 128, 128])
    for recs, ri, bi in [(np.stack([red, blue]), 0, 1), (np.stack([blue, red]), 1, 0)]:
        for eye, want in [((0, 0, 5), [bi, ri]), ((0, 0, -5), [ri, bi])]:
            view = wa.workloads.view_for(recs, eye, 160, 120)
            p = wa.perspective(0.8, 160 / 120, 0.1)
            view.projection[:] = list(p)
            order, _, _, _, _ = run_gpu(ctx, recs, view)
            assert list(order) == want


def test_faint_speck_cull(ctx):
    # test/gpu.test.js:1914-1962
    speck = record([0, 0, 0], [0.0005] * 3, [255, 255, 255, 50], [255, 128, 128, 128])
    big = record([1, 0, 0], [0.2] * 3, [255, 255, 255, 255], [255, 128, 128, 128])
    recs = np.stack([speck, big])
    for cull, want in [(0.5, [1]), (0.0, [0, 1])]:
        mn, mx = wa.workloads.bounds_of(recs)
        view = wa.prepare_view(wa.look_at((0, 0, 5)), wa.perspective(0.8, 160 / 120, 0.1), 160.0, 120.0, mn, mx, cull)
        order, _, _, _, _ = run_gpu(ctx, recs, view)
        assert sorted(order) == want


def test_all_culled_and_all_one_key(ctx):
    recs, view, _ = wa.workloads.plain("plain-64k", n=5000)
    behind = wa.workloads.view_for(recs, (0, 0, 10), 1280, 720)
    behind.model_view[14] += 100.0  # push everything behind the camera
    order, keys, keep, cov, stats = run_gpu(ctx, recs, behind)
    assert stats.count == 0 and not keep.any()
    # Every splat at one depth: one bucket, so the order is the survivors in index order.
    flat = recs.copy()
    pos = flat.view(np.float32).reshape(-1, 8)
    pos[:, 2] = 0.0
    view = wa.workloads.view_for(flat, (0, 0, 10), 1280, 720)
    assert_parity(ctx, flat, view, "one depth")


def test_near_edges_and_thresholds(ctx):
    # Splats placed on bucket edges and at the cull threshold: the f32 fast path must hand them to FP64.
    recs, view, _ = wa.workloads.plain("plain-64k", n=20000, cull=0.5)
    pos = recs.view(np.float32).reshape(-1, 8)
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    rng = np.random.default_rng(9)
    for i in range(0, len(recs), 3):  # move to the nearest bucket edge along the view axis
        p = pos[i, :3].astype(np.float64)
        d = -(mv[2, :3] @ p + mv[2, 3])
        t = (d - view.depth_min) * view.depth_scale
        edge = round(t) + rng.choice([-1e-4, 0, 1e-4])
        dd = edge / view.depth_scale + view.depth_min
        pos[i, :3] = (p + (d - dd) * mv[2, :3]).astype(np.float32)
    # rescale some splats so alpha * area sits on the cull threshold
    k, keep, cov, val = oracle.keys(recs, view)
    for i in np.nonzero(~np.isnan(val))[0][1::4]:
        pos[i, 3:6] *= np.float32(min(4.0, max(0.25, math.sqrt(0.5 / max(val[i], 1e-9)))))
    stats, _ = assert_parity(ctx, recs, view, "edges")


def test_orthographic_and_jittered_projections(ctx):
    recs, view, _ = wa.workloads.plain("plain-64k", n=30000, cull=0.5)
    jit = wa.View.from_buffer_copy(bytes(view))
    jit.projection[8] = 0.37 / 1280
    jit.projection[9] = -0.21 / 720
    assert_parity(ctx, recs, jit, "jittered")
    ortho = wa.View.from_buffer_copy(bytes(view))
    # mat4OrthographicReverseZ-like: x, y scaled, z mapped, w = 1 (the generic kernel path)
    P = np.zeros(16, np.float32)
    P[0], P[5], P[10], P[14], P[15] = 1 / 8.0, 1 / 4.5, 0.01, 0.5, 1.0
    ortho.projection[:] = list(P)
    assert_parity(ctx, recs, ortho, "orthographic")


def test_sort_is_deterministic_and_state_reusable(ctx):
    recs, view, _ = wa.workloads.plain("plain-64k", n=100000)
    st = wa.State(ctx, 100000)
    try:
        a = run_gpu(ctx, recs, view, st)[0]
        small, sview, _ = wa.workloads.plain("plain-64k", n=3000)
        b = run_gpu(ctx, small, sview, st)[0]
        c = run_gpu(ctx, recs, view, st)[0]
    finally:
        st.destroy()
    assert np.array_equal(a, c)
    assert np.array_equal(b, oracle.run(small, sview)[0])


def test_plain_1m_parity(ctx):
    recs, view, _ = wa.workloads.plain("plain-1m")
    assert_parity(ctx, recs, view, "plain-1m")


def test_slab_8m_parity_and_properties(ctx):
    recs, view, spec = wa.workloads.plain("slab-8m")
    order, keys, keep, cov, stats = run_gpu(ctx, recs, view)
    # size-independent properties first: a permutation of the survivors, keys ascending, ties by index
    assert stats.count == int(keep.sum()) and len(np.unique(order)) == len(order)
    assert keep[order].all()
    ko = keys[order].astype(np.int64)
    assert np.all(np.diff(ko) >= 0)
    same = np.diff(ko) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0)
    o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
    assert np.array_equal(keep, o_keep) and np.array_equal(order, o_order)
    assert cov_error(cov, o_cov, keep) <= COV_TOL
    print("slab-8m V=%d recomputed=%d" % (stats.count, stats.recomputed), flush=True)


def test_refusals(ctx):
    with pytest.raises(wa.WindingError, match="not a whole number of 32-byte"):
        wa.Cloud(ctx, np.zeros(33, np.uint8))
    nan = record([float("nan"), 0, 0], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128])
    ok = record([0, 0, 0], [1, 1, 1], [0, 0, 0, 0], [255, 128, 128, 128])
    with pytest.raises(wa.WindingError, match="splat 1 has a position that is not a number"):
        wa.Cloud(ctx, np.stack([ok, nan]))
    recs, view, _ = wa.workloads.plain("plain-64k", n=100)
    cloud = wa.Cloud(ctx, recs)
    st = wa.State(ctx, 50)
    with pytest.raises(wa.WindingError, match="more than this state holds"):
        st.sort(cloud, view)
    bad = wa.View.from_buffer_copy(bytes(view))
    bad.cull = -1
    st2 = wa.State(ctx, 100)
    with pytest.raises(wa.WindingError, match="splatCull must be 0 or more"):
        st2.sort(cloud, bad)
    cloud.destroy()
    st.destroy()
    st2.destroy()
    with pytest.raises(wa.WindingError, match="more than one sort holds"):
        wa.State(ctx, wa.WD_MAX_SORT + 1)


def _box_bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n", [1, 63, 4097, 70001])
def test_cloud_box_equals_put_box(ctx, n):
    """The box wd_cloud_create reports (k_bounds on the GPU; depthMin and depthScale come from it in the JS
    API) equals put()'s min/max (src/scene/splats.js:343-348) bit for bit: negative coordinates, ragged
    sizes, one splat, and zeros of both signs at the ends (put's strict compares keep the first zero)."""
    recs = wa.synth_records(0, 40 + n, n)
    pos = recs.view(np.float32).reshape(-1, 8)
    pos[:, :3] -= 7.5  # mostly negative
    rng = np.random.default_rng(n)
    if n > 1:
        # a zero of each sign at the minimum of x and the maximum of y, the first one positive or negative
        pos[:, 0] = np.abs(pos[:, 0]) + 1.0
        pos[:, 1] = -np.abs(pos[:, 1]) - 1.0
        i, j = sorted(rng.choice(n, 2, replace=False))
        first = rng.choice([0.0, -0.0])
        pos[i, 0], pos[j, 0] = first, -first
        pos[i, 1], pos[j, 1] = -first, first
    cases = [recs]
    if n == 4097:
        cases.append(wa.workloads.plain("slab-8m")[0])
    for r in cases:
        cloud = wa.Cloud(ctx, r)
        try:
            bad, mn, mx = oracle.bounds(r)
            assert bad < 0 or bad >= len(r)
            print("n=%d box %s %s" % (len(r), mn, mx), flush=True)
            assert np.array_equal(_box_bits(cloud.min), _box_bits(mn)), (cloud.min, mn)
            assert np.array_equal(_box_bits(cloud.max), _box_bits(mx)), (cloud.max, mx)
        finally:
            cloud.destroy()


def test_off_axis_projection_takes_the_clamp_branch(ctx):
    """An off-axis (jittered) projection, p[8] and p[9] nonzero: many survivors lie past footprint()'s
    1.3 margin (src/render/splats.js:111-114) and take its clamp; keys, order and cov2d still match."""
    recs = wa.synth_records(0, 77, 65536)
    pos = recs.view(np.float32).reshape(-1, 8)
    pos[:, 0] *= 2.5  # wide in x: many splats beyond the margin
    v = oracle.look_at((0.0, 0.0, 10.0))
    p = oracle.perspective(math.pi / 3, 16 / 9, 0.1).copy()
    p[8], p[9] = 0.6, -0.4
    _, mn, mx = oracle.bounds(recs)
    view = wa.View.from_buffer_copy(bytes(oracle.prepare(v, p, 1280.0, 720.0, mn, mx, 0.5)))
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    vp = pos[:, :3].astype(np.float64) @ mv[:3, :3].T + mv[:3, 3]
    _, _, keep, _ = oracle.run(recs, view)
    clamped = keep & (np.abs(vp[:, 0] / -vp[:, 2]) > (1.3 / p[0]))
    print("survivors %d, clamped %d" % (keep.sum(), clamped.sum()), flush=True)
    assert clamped.sum() > 1000
    assert_parity(ctx, recs, view, "off-axis")


@pytest.mark.parametrize("which", ["edges_4096", "plain-64k"])
def test_record_form_equals_splats_data_form_on_gpu(ctx, which):
    """One capture as .splat records and as the reference's Splats data (centres, Sigma rounded to f32 as
    put() stores it): the same survivors, keys and order on the GPU, and both equal the oracle's."""
    if which == "plain-64k":
        recs, view, _ = wa.workloads.plain("plain-64k")
    else:
        with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", which + ".npz"),
                     allow_pickle=False) as z:
            recs = z["records"]
            view = wa.View.from_buffer_copy(z["view"].astype(np.float32).tobytes())
    centers = np.zeros((len(recs), 4), np.float32)
    centers[:, :3] = recs.view(np.float32).reshape(-1, 8)[:, :3]
    centers.view(np.uint32)[:, 3] = recs[:, 24:28].copy().view(np.uint32).ravel()
    covs = np.array([oracle.sigma(r) for r in recs], np.float64).astype(np.float32)
    o_order, o_keys, o_keep, _ = oracle.run(recs, view)
    a = wa.Cloud(ctx, recs)
    b = wa.Cloud.from_splats(ctx, centers, covs)
    st = wa.State(ctx, len(recs))
    try:
        for cloud in (a, b):
            st.sort(cloud, view)
            order, _, stats = st.read()
            keys, keep = st.read_keys(len(recs))
            assert np.array_equal(keep, o_keep)
            assert np.array_equal(keys[keep], o_keys[o_keep])
            assert np.array_equal(order, o_order)
    finally:
        st.destroy()
        a.destroy()
        b.destroy()


@pytest.mark.parametrize("kpt", [4, 8, 16])
@pytest.mark.parametrize("n", [1, 1023, 1025, 2049, 4097, 70001, 300000])
def test_every_tile_size_matches_oracle(ctx, kpt, n):
    """The single sort's three tile builds (1024, 2048 and 4096 splats a tile) forced on the same inputs:
    each bit-exact against the oracle, at sizes on and beside every tile edge."""
    recs = wa.synth_records(n % 2, 90 + n, n)
    view = wl.view_for(recs, (0.5, 1.0, 10.0), 1280, 720)
    st = wa.State(ctx, n)
    st.set_keys_per_thread(kpt)
    try:
        order, keys, keep, cov, stats = run_gpu(ctx, recs, view, st)
    finally:
        st.destroy()
    o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
    assert np.array_equal(keep, o_keep) and np.array_equal(keys[keep], o_keys[o_keep])
    assert np.array_equal(order, o_order), "kpt %d n %d: order differs" % (kpt, n)
    assert cov_error(cov, o_cov, keep) <= COV_TOL


def test_forced_small_tiles_past_the_buffers_are_refused(ctx):
    """A state of 600,000 splats holds per-tile buffers for tiles of 2048 (293) or, up to 512Ki splats,
    of 1024 (512): tiles of 1024 over 600,000 splats (586) are refused by name, and the size's own
    choice then runs."""
    n = 600_000
    recs = wa.synth_records(0, 5, n)
    view = wl.view_for(recs, (0.0, 0.0, 10.0), 1280, 720)
    cloud = wa.Cloud(ctx, recs)
    st = wa.State(ctx, n)
    try:
        st.set_keys_per_thread(4)
        with pytest.raises(wa.WindingError, match="586 tiles; this state's buffers hold 512"):
            st.sort(cloud, view)
        with pytest.raises(wa.WindingError, match="keys a thread; 0"):
            st.set_keys_per_thread(5)
        st.set_keys_per_thread(0)
        st.sort(cloud, view)
        order, _, _ = st.read()
        assert np.array_equal(order, oracle.run(recs, view)[0])
    finally:
        cloud.destroy()
        st.destroy()


def test_two_contexts_sort_side_by_side(ctx):
    """Two contexts on one device, each its own stream, sorts enqueued in turn with no wait between:
    the streams run side by side and share nothing but the device, so each result is the oracle's."""
    other = wa.Context(0)
    recs_a, view_a, _ = wl.plain("plain-1m")
    recs_b = wa.synth_records(1, 77, 300_000)
    view_b = wl.view_for(recs_b, (1.0, 0.5, 9.0), 1280, 720)
    ca, cb = wa.Cloud(ctx, recs_a), wa.Cloud(other, recs_b)
    sa, sb = wa.State(ctx, len(recs_a)), wa.State(other, len(recs_b))
    try:
        for _ in range(6):
            sa.sort(ca, view_a)
            sb.sort(cb, view_b)
        ctx.sync()
        other.sync()
        assert np.array_equal(sa.read()[0], oracle.run(recs_a, view_a)[0])
        assert np.array_equal(sb.read()[0], oracle.run(recs_b, view_b)[0])
    finally:
        for x in (ca, sa, cb, sb):
            x.destroy()
        other.close()
