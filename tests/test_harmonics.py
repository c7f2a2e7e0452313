"""Spherical-harmonic shading fused into the projection (the reference's shade pass,
src/render/splats.js:315-381, on the `sh` layout of src/scene/splats.js:306-332).

The oracle is pinned by the reference's own known answer (test/gpu.test.js:1968-2035, restated on the
shaded colour: a grey splat whose only harmonic is red's -x term reads redder from +x and less red from -x,
and the same from -x once the capture is turned half round about y; blue's last degree-3 term likewise), by
the eye as the camera position for a look-at view, and by an independent evaluation of the real SH basis
written as a table of polynomials. The GPU is held to the oracle: each colour byte within one unit (the GPU
sums in f32, the oracle in f64, and the byte rounds), alpha and everything the sort writes unchanged.
"""
import math
import struct

import numpy as np
import pytest

import oracle
import winding_amd as wa

NCOEF = [0, 3, 8, 15]


def sh_words(values):
    """(n, k, 3) float coefficients -> the reference's `sh` words (halves, coefficient by coefficient, RGB)."""
    n, k, _ = values.shape
    degree = NCOEF.index(k)
    w = oracle.sh_words(degree)
    halves = np.zeros(n * w * 2, np.uint16)
    idx = (np.arange(n)[:, None, None] * w * 2 + np.arange(k)[None, :, None] * 3 + np.arange(3)[None, None, :])
    halves[idx.ravel()] = values.astype(np.float16).view(np.uint16).ravel()
    return halves.view(np.uint32), degree


def record(pos, scale, rgba, quat=(255, 128, 128, 128)):
    return np.frombuffer(struct.pack("<6f4B4B", *pos, *scale, *rgba, *quat), np.uint8).copy()


def grey_view(eye, recs, world=None):
    v = oracle.look_at(eye, (0, 0, 0))
    p = oracle.perspective(0.8, 160 / 120, 0.1)
    bad, mn, mx = oracle.bounds(recs)
    assert bad == -1
    return oracle.prepare(v, p, 160, 120, mn, mx, 0.0, world)


def rgb(u):
    return [(int(u) >> (8 * c)) & 255 for c in range(3)]


def test_eye_is_the_camera_for_a_look_at_view():
    rng = np.random.default_rng(3)
    for _ in range(20):
        eye = rng.uniform(-20, 20, 3)
        mv = oracle.look_at(eye, rng.uniform(-1, 1, 3))
        assert np.allclose(oracle.eye(mv), eye, rtol=1e-5, atol=1e-5)


def test_eye_is_the_inverse_translation_with_a_world():
    rng = np.random.default_rng(4)
    for _ in range(20):
        mv = oracle.look_at(rng.uniform(-20, 20, 3), (0, 0, 0))
        a = rng.normal(size=(3, 3))
        world = np.eye(4, dtype=np.float32)
        world[:3, :3] = a.T  # column-major storage of a
        world[3, :3] = rng.uniform(-5, 5, 3)
        m = oracle.mat4_multiply(mv, world.ravel())
        M = np.asarray(m, np.float64).reshape(4, 4).T  # row-major maths matrix
        want = np.linalg.inv(M)[:3, 3]
        assert np.allclose(oracle.eye(m), want, rtol=1e-4, atol=1e-4)


def test_known_answer_degree_one_red_minus_x_term():
    # test/gpu.test.js:1968-2035: a grey splat, f_rest_2 = 0.6 (red's third coefficient, the -x term).
    rec = record((0, 0, 0), (0.4, 0.4, 0.4), (128, 128, 128, 255))
    k = np.zeros((1, 3, 3))
    k[0, 2, 0] = 0.6
    words, degree = sh_words(k)
    plus = rgb(oracle.shade(rec, words, degree, grey_view((5, 0, 0), rec.reshape(1, 32)))[0])
    minus = rgb(oracle.shade(rec, words, degree, grey_view((-5, 0, 0), rec.reshape(1, 32)))[0])
    assert plus[0] > plus[1] + 40 and minus[0] < minus[1] - 40 and abs(plus[1] - minus[1]) <= 2
    # turned half round about y: the capture's +x faces the world's -x, so from -x it reads as from +x
    turn = np.diag([-1.0, 1.0, -1.0, 1.0]).astype(np.float32).ravel()
    turned = rgb(oracle.shade(rec, words, degree, grey_view((-5, 0, 0), rec.reshape(1, 32), turn))[0])
    assert abs(turned[0] - plus[0]) <= 2
    # the closed form: 128/255 + 0.4886025 * 0.6 from +x (direction -x), bytes rounded
    assert plus[0] == math.floor(0.5 + 255 * (128 / 255 + 0.4886025119029199 * float(np.float16(0.6))))


def test_known_answer_degree_three_blue_last_term():
    # the second splat of the reference's degree-3 file: only f_rest_44, blue's -x (x^2 - 3y^2) term
    rec = record((0, 0, 0), (0.4, 0.4, 0.4), (128, 128, 128, 255))
    k = np.zeros((1, 15, 3))
    k[0, 14, 2] = 0.6
    words, degree = sh_words(k)
    near = rgb(oracle.shade(rec, words, degree, grey_view((5, 0, 0), rec.reshape(1, 32)))[0])
    away = rgb(oracle.shade(rec, words, degree, grey_view((-5, 0, 0), rec.reshape(1, 32)))[0])
    assert near[2] > near[1] + 40 and away[2] < away[1] - 40


def basis(d):
    """The real SH basis past the constant term as a table of polynomials (3DGS sign convention)."""
    x, y, z = d
    return np.array([
        -0.4886025119029199 * y, 0.4886025119029199 * z, -0.4886025119029199 * x,
        1.0925484305920792 * x * y, -1.0925484305920792 * y * z, 0.31539156525252005 * (2 * z * z - x * x - y * y),
        -1.0925484305920792 * x * z, 0.5462742152960396 * (x * x - y * y),
        -0.5900435899266435 * y * (3 * x * x - y * y), 2.890611442640554 * x * y * z,
        -0.4570457994644658 * y * (4 * z * z - x * x - y * y), 0.3731763325901154 * z * (2 * z * z - 3 * x * x - 3 * y * y),
        -0.4570457994644658 * x * (4 * z * z - x * x - y * y), 1.445305721320277 * z * (x * x - y * y),
        -0.5900435899266435 * x * (x * x - 3 * y * y)])


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_shade_matches_the_basis_table(degree):
    rng = np.random.default_rng(10 + degree)
    recs, view, _ = wa.workloads.plain("plain-64k", n=300)
    k = rng.normal(0, 0.3, (300, NCOEF[degree], 3))
    words, _ = sh_words(k)
    got = oracle.shade(recs, words, degree, view)
    eye = oracle.eye(oracle.to_oview(view).model_view).astype(np.float64)
    pos = recs.view(np.float32).reshape(-1, 8)[:, :3].astype(np.float64)
    kq = k.astype(np.float16).astype(np.float64)
    for i in range(300):
        d = (pos[i] - eye) / np.linalg.norm(pos[i] - eye)
        c = basis(d)[:NCOEF[degree]] @ kq[i]
        base = recs[i, 24:27].astype(np.float64) / 255
        want = np.floor(0.5 + 255 * np.clip(base + c, 0, 1)).astype(int)
        assert np.abs(np.array(rgb(got[i])) - want).max() <= 1, i
        assert (int(got[i]) >> 24) == recs[i, 27]


# ---- the GPU, held to the oracle ----

@pytest.mark.gpu
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_gpu_shading_matches_oracle_and_leaves_the_sort_alone(ctx, degree):
    rng = np.random.default_rng(20 + degree)
    recs, view, _ = wa.workloads.plain("plain-64k", n=20000, cull=0.5)
    n = len(recs)
    k = rng.normal(0, 0.3, (n, NCOEF[degree], 3))
    words, _ = sh_words(k)
    cloud = wa.Cloud(ctx, recs)
    st = wa.State(ctx, n)
    try:
        st.sort(cloud, view)
        order0, cov0, stats0 = st.read(n_cov=n)
        cloud.set_sh(words, degree)
        st.sort(cloud, view)
        order, cov, stats = st.read(n_cov=n)
        assert stats.count == stats0.count and np.array_equal(order, order0)
        assert np.array_equal(cov[:, :3].view(np.uint32), cov0[:, :3].view(np.uint32))
        keep = order
        want = oracle.shade(recs, words, degree, view)[keep]
        got = cov[keep, 3].view(np.uint32)
        diff = np.abs(np.array([[(int(g) >> (8 * c)) & 255 for c in range(4)] for g in got], np.int64) -
                      np.array([[(int(w) >> (8 * c)) & 255 for c in range(4)] for w in want], np.int64))
        print("degree %d: %d survivors, bytes off by one: %d" % (degree, len(keep), int((diff == 1).sum())), flush=True)
        assert diff[:, :3].max() <= 1 and diff[:, 3].max() == 0
        # degree 0 takes them away: the record's colour again
        cloud.set_sh(None, 0)
        st.sort(cloud, view)
        _, cov2, _ = st.read(n_cov=n)
        assert np.array_equal(cov2.view(np.uint32), cov0.view(np.uint32))
    finally:
        st.destroy()
        cloud.destroy()


@pytest.mark.gpu
def test_gpu_known_answer_and_refusals(ctx):
    rec = record((0, 0, 0), (0.4, 0.4, 0.4), (128, 128, 128, 255))
    k = np.zeros((1, 3, 3))
    k[0, 2, 0] = 0.6
    words, degree = sh_words(k)
    cloud = wa.Cloud(ctx, rec)
    st = wa.State(ctx, 1)
    try:
        cloud.set_sh(words, degree)
        seen = {}
        for name, eye in (("plus", (5, 0, 0)), ("minus", (-5, 0, 0))):
            v, p = wa.workloads.camera(eye, 160, 120)
            view = wa.prepare_view(v, p, 160.0, 120.0, cloud.min, cloud.max, 0.0)
            st.sort(cloud, view)
            _, cov, stats = st.read(n_cov=1)
            assert stats.count == 1
            seen[name] = rgb(cov[0, 3].view(np.uint32))
            assert seen[name] == rgb(oracle.shade(rec, words, degree, view)[0])
        assert seen["plus"][0] > seen["plus"][1] + 40 and seen["minus"][0] < seen["minus"][1] - 40
        with pytest.raises(wa.WindingError, match="degree 4"):
            cloud.set_sh(words, 4)
        with pytest.raises(wa.WindingError, match="take 23 words"):
            cloud.set_sh(words, 3)
        # a batch shades as the single sort does; under a projection other than the reverse-Z
        # perspective it refuses such a cloud
        batch = wa.Batch(ctx, 1)
        try:
            v, p = wa.workloads.camera((5, 0, 0), 160, 120)
            view = wa.prepare_view(v, p, 160.0, 120.0, cloud.min, cloud.max, 0.0)
            batch.set([(cloud, view, st)])
            batch.run()
            _, cov, stats = st.read(n_cov=1)
            assert stats.count == 1 and rgb(cov[0, 3].view(np.uint32)) == seen["plus"]
            ortho = np.zeros(16, np.float32)
            ortho[0], ortho[5], ortho[10], ortho[14], ortho[15] = 0.1, 0.1, -0.01, 0.5, 1.0
            view2 = wa.prepare_view(v, ortho, 160.0, 120.0, cloud.min, cloud.max, 0.0)
            with pytest.raises(wa.WindingError, match="reverse-Z perspective only"):
                batch.set([(cloud, view2, st)])
        finally:
            batch.destroy()
    finally:
        st.destroy()
        cloud.destroy()
