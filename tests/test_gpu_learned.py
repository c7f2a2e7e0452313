"""The learned sort: the predictor trained on the exact sort's own quantiles, then up to 8 views of one
cloud binned by its depth CDF. Every view's keys, survivors and order must equal the oracle's (and the
exact sort's), bit for bit, whether the fast path held or the check sent the view back to the exact path;
and the trained predictor must keep views on the fast path where even key-width bins (the method without
it) do not."""
import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu

N = 262144
W, H = 1280, 720


RECS = {}


@pytest.fixture(scope="module")
def scene(ctx):
    recs = wa.synth_records(0, 7, N)
    cloud = wa.Cloud(ctx, recs)
    RECS[id(cloud)] = recs
    exact = wa.State(ctx, N)
    states = [wa.State(ctx, N) for _ in range(8)]
    yield cloud, exact, states
    for s in states:
        s.destroy()
    exact.destroy()
    cloud.destroy()


@pytest.fixture(scope="module")
def trained(ctx, scene):
    cloud, exact, _ = scene
    X, Y = wl.training_set(ctx, cloud, exact, wl.training_poses(512, 99), W, H)
    mlp = wa.Mlp(ctx, seed=1)
    before, _ = mlp.grad(X, Y)
    mlp.train(X, Y, epochs=150, batch=64, lr=2e-3, seed=2)
    after, _ = mlp.grad(X, Y)
    entropy = float(-(Y * np.log(np.maximum(Y, 1e-30))).sum(axis=1).mean())
    print("predictor: %d poses, loss %.4f -> %.4f (target entropy %.4f)" % (len(X), before, after, entropy),
          flush=True)
    assert after - entropy < 0.5 * (before - entropy)
    yield mlp
    mlp.destroy()


def orbit_views(cloud, first, count, step=37):
    eyes = wl.orbit_sequence(1024)
    return [wl.view_at(cloud, eyes[(first + k * step) % 1024], W, H) for k in range(count)]


def check_oracle(st, recs, view, order, cov, label):
    """One view's learned result against the oracle itself: survivors, keys and order bit for bit,
    cov2d within 1e-5 of the FP64 footprint."""
    o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
    keys, keep = st.read_keys(len(recs))
    assert np.array_equal(keep, o_keep), "%s: survivors differ from the oracle's" % label
    assert np.array_equal(keys[keep], o_keys[o_keep]), "%s: keys differ from the oracle's" % label
    assert np.array_equal(order, o_order), "%s: order differs from the oracle's" % label
    if keep.any():
        g, w = cov[keep, :3].astype(np.float64), o_cov[keep]
        err = float(np.max(np.abs(g - w).max(axis=1) / np.maximum(np.abs(w[:, 0]), np.abs(w[:, 2]))))
        assert err <= 1e-5, "%s: cov2d error %g against the oracle" % (label, err)


def check_views(ctx, cloud, exact, states, views, label, n=N, recs=None):
    """Each view's learned result against the oracle (when the records are known) and against wd_sort's;
    returns how many views fell back."""
    fell = 0
    recs = RECS.get(id(cloud)) if recs is None else recs
    for k, (view, st) in enumerate(zip(views, states)):
        order, cov, stats = st.read(n_cov=n)
        if recs is not None:
            check_oracle(st, recs, view, order, cov, "%s view %d" % (label, k))
        exact.sort(cloud, view)
        e_order, e_cov, e_stats = exact.read(n_cov=n)
        assert stats.count == e_stats.count, "%s view %d: V %d vs %d" % (label, k, stats.count, e_stats.count)
        assert np.array_equal(order, e_order), "%s view %d: order differs" % (label, k)
        # cov2d: the same survivors and colours; the footprint to the FP32 bar (both kernels compile the f32
        # footprint with contraction allowed, so the last bits may differ between them)
        assert np.array_equal(cov[:, 3].view(np.uint32), e_cov[:, 3].view(np.uint32))
        a, b = cov[:, :3].astype(np.float64), e_cov[:, :3].astype(np.float64)
        scale = np.maximum(np.maximum(np.abs(b[:, 0]), np.abs(b[:, 2])), 1e-30)
        err = float(np.max(np.abs(a - b).max(axis=1) / scale))
        assert err <= 1e-5, "%s view %d: cov2d differs by %g" % (label, k, err)
        fell += stats.fallback
    return fell


def test_learned_is_exact_and_stays_on_the_fast_path(ctx, scene, trained):
    cloud, exact, states = scene
    learned = wa.Learned(ctx, 8)
    try:
        for first in (0, 300):
            views = orbit_views(cloud, first, 8)
            learned.run(trained, cloud, views, states)
            fell = check_views(ctx, cloud, exact, states, views, "trained")
            print("orbit from %d: %d of 8 views fell back" % (first, fell), flush=True)
            assert fell == 0
        # a learned run keeps the keys as well: its state's training targets are the exact sort's
        exact.sort(cloud, views[0])
        assert np.array_equal(states[0].quantiles(), exact.quantiles())
        # fewer views than the run holds, and a single view
        for k in (3, 1):
            views = orbit_views(cloud, 500, k)
            learned.run(trained, cloud, views, states[:k])
            assert check_views(ctx, cloud, exact, states[:k], views, "%d views" % k) == 0
    finally:
        learned.destroy()


def test_a_failed_check_runs_the_exact_path(ctx, scene, trained):
    cloud, exact, states = scene
    learned = wa.Learned(ctx, 8)
    try:
        views = orbit_views(cloud, 100, 8)
        learned.run(trained, cloud, views, states, cap=1)  # every bin is over: every view falls back
        assert check_views(ctx, cloud, exact, states, views, "cap 1") == 8
        for st in states:
            assert st.stats().overflow_bins > 0
        learned.run(trained, cloud, views, states)  # and the next run is back on the fast path
        assert check_views(ctx, cloud, exact, states, views, "after fallback") == 0
    finally:
        learned.destroy()


def test_trained_predictor_beats_key_width_bins(ctx, scene, trained):
    """The method without the network: bins of equal key width (zero weights give uniform widths, i.e.
    the key's high byte). With a cap of 3x an even share, those bins overflow on the orbit; the trained
    network's bins do not. Both stay exact."""
    cloud, exact, states = scene
    flat = wa.Mlp(ctx, seed=0)
    flat.set_params(np.zeros(wa.WD_MLP_PARAMS, np.float32))
    learned = wa.Learned(ctx, 8)
    cap = 3 * N // 256
    try:
        views = orbit_views(cloud, 700, 8)
        learned.run(flat, cloud, views, states, cap=cap)
        fell_flat = check_views(ctx, cloud, exact, states, views, "key-width bins")
        learned.run(trained, cloud, views, states, cap=cap)
        fell_trained = check_views(ctx, cloud, exact, states, views, "trained bins")
        print("cap %d: key-width bins fell back on %d of 8 views, trained bins on %d" % (cap, fell_flat, fell_trained),
              flush=True)
        assert fell_flat >= 6 and fell_trained == 0
    finally:
        learned.destroy()
        flat.destroy()


@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_ragged_sizes_and_a_view_with_no_survivors(ctx, n):
    """Clouds that end inside a wave, a round and a tile, and a view that culls every splat (no
    survivors, so every bin is empty): still the exact order, with no fallback."""
    recs = wa.synth_records(0, 11, n)
    cloud = wa.Cloud(ctx, recs)
    exact = wa.State(ctx, n)
    states = [wa.State(ctx, n) for _ in range(3)]
    flat = wa.Mlp(ctx, seed=0)
    flat.set_params(np.zeros(wa.WD_MLP_PARAMS, np.float32))
    learned = wa.Learned(ctx, 3)
    try:
        eyes = wl.orbit_sequence(1024)
        views = [wl.view_at(cloud, eyes[k * 300], W, H) for k in range(2)]
        away = wl.view_at(cloud, eyes[0], W, H)
        away.model_view[14] += 1000.0  # everything behind the camera
        views.append(away)
        learned.run(flat, cloud, views, states)
        assert check_views(ctx, cloud, exact, states, views, "n=%d" % n, n=n, recs=recs) == 0
        assert states[2].stats().count == 0
    finally:
        learned.destroy()
        flat.destroy()
        for st in states:
            st.destroy()
        exact.destroy()
        cloud.destroy()
