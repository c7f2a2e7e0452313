"""The learned workload at its own size: 4,194,304 splats of the mixture generator, the predictor trained
on the exact sort's quantiles at 4,096 random poses of the orbit family (as the benchmark trains it), then
runs of 8 poses of the 1,024-pose orbit. Every view is checked against the oracle itself: survivors, keys
and draw order bit for bit, cov2d within 1e-5 of the FP64 footprint; the fast path must hold on every view
with the default cap, and a cap of 1 must send every view down the exact path with the same result."""
import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu

SPEC = wl.LEARNED


def test_learned_4m_against_the_oracle(ctx):
    n, W, H = SPEC["n"], SPEC["width"], SPEC["height"]
    recs = wa.synth_records(SPEC["kind"], SPEC["seed"], n)
    cloud = wa.Cloud(ctx, recs)
    states = [wa.State(ctx, n) for _ in range(8)]
    mlp = wa.Mlp(ctx, seed=1)
    learned = wa.Learned(ctx, 8)
    try:
        X, Y = wl.training_set(ctx, cloud, states[0], wl.training_poses(SPEC["train_poses"], SPEC["train_seed"]),
                               W, H)
        assert len(X) == SPEC["train_poses"]
        mlp.train(X, Y, epochs=40, batch=256, lr=2e-3, seed=2)
        eyes = wl.orbit_sequence(1024)
        cases = [(0, 0), (8, 0), (512, 1)]  # (first pose of the group, cap); cap 1 forces the exact path
        for first, cap in cases:
            views = [wl.view_at(cloud, eyes[first + k], W, H) for k in range(8)]
            learned.run(mlp, cloud, views, states, cap=cap)
            fell = 0
            for k, (view, st) in enumerate(zip(views, states)):
                label = "poses %d+%d cap %d" % (first, k, cap)
                order, cov, stats = st.read(n_cov=n)
                keys, keep = st.read_keys(n)
                o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
                assert np.array_equal(keep, o_keep), label
                assert np.array_equal(keys[keep], o_keys[o_keep]), label
                assert stats.count == len(o_order) and np.array_equal(order, o_order), label
                g, w = cov[keep, :3].astype(np.float64), o_cov[keep]
                err = float(np.max(np.abs(g - w).max(axis=1) / np.maximum(np.abs(w[:, 0]), np.abs(w[:, 2]))))
                assert err <= 1e-5, "%s: cov2d error %g" % (label, err)
                fell += stats.fallback
            print("poses %d..%d, cap %d: V %d..., %d of 8 fell back" % (first, first + 7, cap, stats.count, fell),
                  flush=True)
            assert fell == (8 if cap == 1 else 0)
    finally:
        learned.destroy()
        mlp.destroy()
        for s in states:
            s.destroy()
        cloud.destroy()
