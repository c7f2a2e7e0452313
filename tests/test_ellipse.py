"""The draw's per-splat ellipse set-up, made by the projection beside cov2d (the splat vertex
shader's per-instance work, src/render/splats.js:263-291).

The reference has no test of these numbers, so the oracle (oracle/splat_oracle.c or_ellipse,
or_ellipses, FP64) is pinned by the mathematics: a circle's closed form, numpy's symmetric
eigen-solver on random footprints (lengths sqrt(2 lambda) and the eigenvector's direction), the
1024-pixel cap, the reach's closed form per alpha byte, and the clip position as numpy's product of
the two matrices. The GPU is held to the oracle: clip positions bit for bit (both round the same
FP64 value), zeros exactly where the oracle draws nothing, and the axes within 1e-5 of the major
axis's length.
"""
import math

import numpy as np
import pytest

import oracle
import winding_amd as wa

C03 = float(np.float32(0.3))


def reach_scale(abyte):
    return math.sqrt(min(math.log(255.0 * (abyte / 255.0)), 4.0))


def test_circle_closed_form():
    for s, abyte in ((1.0, 255), (7.5, 50), (0.01, 2)):
        got = oracle.ellipse((s, 0.0, s), abyte)
        e = reach_scale(abyte)
        r = math.sqrt(2 * (s + C03)) * e
        # a circle: along x, across a quarter turn clockwise (down the screen's y)
        assert np.allclose(got, [r, 0.0, 0.0, -r], rtol=1e-15, atol=0)


def test_taller_than_wide_takes_y_first():
    got = oracle.ellipse((1.0, 0.0, 4.0), 255)
    e = reach_scale(255)
    assert np.allclose(got, [0.0, math.sqrt(2 * (4 + C03)) * e, math.sqrt(2 * (1 + C03)) * e, 0.0], rtol=1e-15)


def test_axes_are_the_eigenvectors_of_the_footprint():
    rng = np.random.default_rng(5)
    for _ in range(500):
        g = rng.normal(size=(2, 2)) * rng.uniform(0.1, 30)
        f = g @ g.T
        cov = (f[0, 0], f[0, 1], f[1, 1])
        abyte = int(rng.integers(2, 256))
        got = oracle.ellipse(cov, abyte)
        e = reach_scale(abyte)
        C = np.array([[cov[0] + C03, cov[1]], [cov[1], cov[2] + C03]])
        lam, vec = np.linalg.eigh(C)
        A, B = got[:2], got[2:]
        la, lb = min(math.sqrt(2 * lam[1]), 1024) * e, min(math.sqrt(2 * lam[0]), 1024) * e
        assert math.isclose(np.linalg.norm(A), la, rel_tol=1e-9)
        assert math.isclose(np.linalg.norm(B), lb, rel_tol=1e-9)
        if lam[1] - lam[0] > 1e-6 * lam[1]:
            assert abs(abs(A @ vec[:, 1]) / la - 1) < 1e-8  # along the larger eigenvector
        # across is along turned a quarter clockwise: (x, y) -> (y, -x)
        assert np.allclose(B * la, lb * np.array([A[1], -A[0]]), rtol=0, atol=1e-9 * la * lb)
        # the quad's axes rebuild 2 e^2 C
        M = np.outer(A, A) + np.outer(B, B)
        assert np.allclose(M, 2 * e * e * C, rtol=1e-9, atol=1e-9 * M.max())


def test_the_1024_pixel_cap():
    got = oracle.ellipse((4e6, 0.0, 100.0), 255)
    e = reach_scale(255)
    assert math.isclose(got[0], 1024 * e, rel_tol=1e-15)
    assert math.isclose(-got[3], math.sqrt(2 * (100 + C03)) * e, rel_tol=1e-12)


def test_reach_per_alpha_byte():
    assert oracle.ellipse((1.0, 0.0, 1.0), 0) is None  # log 0: nothing drawn
    assert oracle.ellipse((1.0, 0.0, 1.0), 1) is None  # log 1 = 0, not above zero
    r = math.sqrt(2 * (1 + C03))
    for abyte, e in ((2, math.sqrt(math.log(2))), (50, math.sqrt(math.log(50))), (60, 2.0), (255, 2.0)):
        assert math.isclose(oracle.ellipse((1.0, 0.0, 1.0), abyte)[0], r * e, rel_tol=1e-12), abyte


def test_cloud_ellipses_clip_and_zeros():
    recs, view, _ = wa.workloads.plain("plain-64k", n=3000, cull=0.5)
    _, keep, cov, _ = oracle.keys(recs, view)
    E = oracle.ellipses(recs, view)
    drawn = np.any(E != 0, axis=1)
    assert np.array_equal(drawn, keep)  # every survivor has alpha >= 2 and a positive minor axis
    ov = oracle.to_oview(view)
    MV = np.array(ov.model_view, np.float64).reshape(4, 4).T
    P = np.array(ov.projection, np.float64).reshape(4, 4).T
    pos = recs.view(np.float32).reshape(-1, 8)[:, :3].astype(np.float64)
    clip = (P @ (MV @ np.c_[pos, np.ones(len(pos))].T)).T
    assert np.allclose(E[keep, :4], clip[keep], rtol=1e-12, atol=1e-9)
    for i in np.nonzero(keep)[0][:200]:
        assert np.array_equal(E[i, 4:], oracle.ellipse(cov[i], recs[i, 27]))


# ---- the GPU, held to the oracle ----

def check_against_oracle(recs, view, got):
    want = oracle.ellipses(recs, view)
    drawn = np.any(want != 0, axis=1)
    assert np.array_equal(np.any(got != 0, axis=1), drawn)
    assert not np.any(got[~drawn])
    # clip: the same FP64 value rounded to f32 on both sides
    assert np.array_equal(got[drawn, :4], want[drawn, :4].astype(np.float32))
    A, B = got[drawn, 4:6].astype(np.float64), got[drawn, 6:8].astype(np.float64)
    Aw, Bw = want[drawn, 4:6], want[drawn, 6:8]
    la = np.linalg.norm(Aw, axis=1)
    tol = 1e-5 * la
    assert np.all(np.abs(np.linalg.norm(A, axis=1) - la) <= tol)
    assert np.all(np.abs(np.linalg.norm(B, axis=1) - np.linalg.norm(Bw, axis=1)) <= tol)
    # where the axes are well apart, each component; everywhere, the quad's matrix
    gap = 1 - (np.linalg.norm(Bw, axis=1) / la) ** 2
    sharp = gap > 0.05
    comp = np.max(np.abs(np.c_[A, B] - np.c_[Aw, Bw]), axis=1) / la
    assert np.all(comp[sharp] <= 1e-5), comp[sharp].max()
    M = lambda a, b: np.stack([a[:, 0] ** 2 + b[:, 0] ** 2, a[:, 0] * a[:, 1] + b[:, 0] * b[:, 1],
                               a[:, 1] ** 2 + b[:, 1] ** 2], 1)
    merr = np.max(np.abs(M(A, B) - M(Aw, Bw)), axis=1) / la ** 2
    assert np.all(merr <= 1e-5), merr.max()
    return int(drawn.sum()), float(comp[sharp].max(initial=0)), float(merr.max(initial=0))


@pytest.mark.gpu
@pytest.mark.parametrize("cull", [0.5, 0.0])
def test_gpu_ellipses_match_oracle_and_leave_the_sort_alone(ctx, cull):
    recs, view, _ = wa.workloads.plain("plain-64k", n=30000, cull=cull)
    n = len(recs)
    cloud = wa.Cloud(ctx, recs)
    st = wa.State(ctx, n)
    try:
        st.sort(cloud, view)
        order0, cov0, stats0 = st.read(n_cov=n)
        st.set_ellipses(True)
        st.sort(cloud, view)
        order, cov, stats = st.read(n_cov=n)
        assert stats.count == stats0.count and np.array_equal(order, order0)
        assert np.array_equal(cov.view(np.uint32), cov0.view(np.uint32))
        got = st.read_ellipses(n)
        drawn, comp, merr = check_against_oracle(recs, view, got)
        print("cull %.1f: %d drawn, worst axis component %.2e, worst quad matrix %.2e" % (cull, drawn, comp, merr),
              flush=True)
        st.set_ellipses(False)
        with pytest.raises(wa.WindingError, match="does not take them"):
            st.read_ellipses(n)
    finally:
        st.destroy()
        cloud.destroy()


@pytest.mark.gpu
def test_gpu_ellipses_general_projection_and_harmonics(ctx):
    # an orthographic projection takes the general projection form; harmonics the shaded one
    recs, view, _ = wa.workloads.plain("plain-64k", n=5000, cull=0.0)
    proj = np.zeros(16, np.float32)
    proj[0], proj[5], proj[10], proj[14], proj[15] = 0.1, 0.1, -0.01, 0.5, 1.0
    mv = np.array(view.model_view[:], np.float32)
    cloud = wa.Cloud(ctx, recs)
    v2 = wa.prepare_view(mv, proj, 640.0, 480.0, cloud.min, cloud.max, 0.0)
    rng = np.random.default_rng(3)
    words = rng.integers(0, 1 << 14, size=len(recs) * oracle.sh_words(1)).astype(np.uint32)
    st = wa.State(ctx, len(recs))
    try:
        st.set_ellipses(True)
        st.sort(cloud, v2)
        check_against_oracle(recs, v2, st.read_ellipses(len(recs)))
        cloud.set_sh(words, 1)
        st.sort(cloud, view)
        check_against_oracle(recs, view, st.read_ellipses(len(recs)))
        batch = wa.Batch(ctx, 1)
        try:
            cloud.set_sh(None, 0)
            with pytest.raises(wa.WindingError, match="reverse-Z perspective only"):
                batch.set([(cloud, v2, st)])
        finally:
            batch.destroy()
    finally:
        st.destroy()
        cloud.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [wa.WD_SCHEDULE_PHASED, wa.WD_SCHEDULE_PERSISTENT])
def test_gpu_batch_ellipses_equal_single_sorts(ctx, schedule):
    # one batch: records with ellipses, centres and Sigma with harmonics and ellipses, records without;
    # each job's order, cov2d and ellipses equal its own single sort's, bit for bit
    recs, view, _ = wa.workloads.plain("plain-64k", n=20000, cull=0.0)
    rng = np.random.default_rng(11)
    centers = np.frombuffer(recs.tobytes(), np.float32).reshape(-1, 8)[:, :4].copy()
    centers[:, 3] = np.frombuffer(recs[:, 24:28].tobytes(), np.float32)
    covs = np.stack([oracle.sigma(r) for r in recs[:6000]]).astype(np.float32)
    words = rng.integers(0, 1 << 14, size=6000 * oracle.sh_words(2)).astype(np.uint32)
    a = wa.Cloud(ctx, recs)
    b = wa.Cloud.from_splats(ctx, centers[:6000], covs)
    b.set_sh(words, 2)
    c = wa.Cloud(ctx, recs[:9000])
    jobs = [(a, view, True), (b, view, True), (c, view, False)]
    states = [wa.State(ctx, cl.count) for cl, _, _ in jobs]
    singles = [wa.State(ctx, cl.count) for cl, _, _ in jobs]
    batch = wa.Batch(ctx, 4)
    try:
        for (cl, v, ell), st, one in zip(jobs, states, singles):
            st.set_ellipses(ell)
            one.set_ellipses(ell)
            one.sort(cl, v)
        batch.schedule(schedule)
        batch.set([(cl, v, st) for (cl, v, _), st in zip(jobs, states)])
        for run in range(2):
            batch.run()
            for (cl, _, ell), st, one in zip(jobs, states, singles):
                o1, c1, s1 = st.read(n_cov=cl.count)
                o2, c2, s2 = one.read(n_cov=cl.count)
                print("run %d: %d splats, %d drawn, ellipses %s" % (run, cl.count, s1.count, ell), flush=True)
                assert s1.count == s2.count and np.array_equal(o1, o2)
                assert np.array_equal(c1.view(np.uint32), c2.view(np.uint32))
                if ell:
                    assert np.array_equal(st.read_ellipses(cl.count).view(np.uint32),
                                          one.read_ellipses(cl.count).view(np.uint32))
        states[2].set_ellipses(True)
        with pytest.raises(wa.WindingError, match="turned on or off since the jobs were set"):
            batch.run()
    finally:
        batch.destroy()
        for st in states + singles:
            st.destroy()
        for cl in (a, b, c):
            cl.destroy()
