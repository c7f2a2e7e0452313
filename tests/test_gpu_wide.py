"""Clouds past 2^24 splats: the two scatter passes carry 8-byte elements (high byte << 32 | index)
instead of 4-byte ones (high byte << 24 | index). The same parity bar as every other sort: keys,
survivor flags and order bit-exact against the oracle, cov2d within 1e-5."""
import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_sort_past_2_24_splats_matches_oracle(ctx):
    n = wa.WD_MAX_NARROW + 4097
    recs = wa.synth_records(0, 2024, n)
    view = wl.view_for(recs, (1.0, 2.0, 10.0), 1920, 1080)
    cloud = wa.Cloud(ctx, recs)
    st = wa.State(ctx, n)
    try:
        st.sort(cloud, view)
        order, cov, stats = st.read(n_cov=n)
        keys, keep = st.read_keys(n)
        print("n=%d V=%d" % (n, stats.count), flush=True)
        o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
        assert np.array_equal(keep, o_keep)
        assert np.array_equal(keys[keep], o_keys[o_keep])
        assert stats.count == len(o_order) and stats.count > wa.WD_MAX_NARROW // 2
        assert np.array_equal(order, o_order), "order differs first at %d" % int(np.argmax(order != o_order))
        assert int(order.max()) >= wa.WD_MAX_NARROW  # indices past 24 bits made it through
        g, w = cov[keep, :3].astype(np.float64), o_cov[keep]
        err = float(np.max(np.abs(g - w).max(axis=1) / np.maximum(np.abs(w[:, 0]), np.abs(w[:, 2]))))
        assert err <= 1e-5, "cov2d error %g" % err
        # the same state sorting a small cloud takes the 4-byte form again
        small = wa.synth_records(1, 7, 70001)
        sview = wl.view_for(small, (0.0, 0.0, 10.0), 1280, 720)
        scloud = wa.Cloud(ctx, small)
        st.sort(scloud, sview)
        s_order, _, s_stats = st.read()
        assert np.array_equal(s_order, oracle.run(small, sview)[0])
        scloud.destroy()
        # a batch job takes up to 2^24 splats and says so
        batch = wa.Batch(ctx, 1)
        with pytest.raises(wa.WindingError, match="a batch job takes up to"):
            batch.set([(cloud, view, st)])
        batch.destroy()
    finally:
        cloud.destroy()
        st.destroy()


def test_sort_at_the_size_limit(ctx):
    """WD_MAX_SORT splats (2^28): a 2^24-splat cloud tiled 16 times, so that every copy has the same
    survivors, keys and cov2d, and the stable order of the whole follows from the copy's: within a key,
    copy 0's indices, then copy 1's, and so on. The copy's own sort is held to the oracle above."""
    n_small, copies = 1 << 24, 16
    n = n_small * copies
    assert n == wa.WD_MAX_SORT
    small = wa.synth_records(0, 99, n_small)
    view = wl.view_for(small, (0.5, 1.5, 10.0), 1920, 1080)
    st = wa.State(ctx, n)
    try:
        cloud = wa.Cloud(ctx, small)
        try:
            st.sort(cloud, view)
            s_order, s_cov, s_stats = st.read(n_cov=n_small)
            s_keys, s_keep = st.read_keys(n_small)
            s_box = np.concatenate([cloud.min, cloud.max]).view(np.uint32).copy()
        finally:
            cloud.destroy()
        big = np.tile(small, (copies, 1))
        cloud = wa.Cloud(ctx, big)
        del big
        try:
            assert np.array_equal(np.concatenate([cloud.min, cloud.max]).view(np.uint32), s_box)  # copies keep the box
            st.sort(cloud, view)
            order, _, stats = st.read()
            print("n=%d V=%d (copy V=%d)" % (n, stats.count, s_stats.count), flush=True)
            assert stats.count == copies * s_stats.count
            # the order the copies' stable sort implies: a key's group of the copy's order, once a copy
            k_sorted = s_keys[s_order].astype(np.int64)
            counts = np.bincount(k_sorted, minlength=65536)
            starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
            p = np.arange(len(s_order), dtype=np.int64)
            s, L = starts[k_sorted], counts[k_sorted]
            want = np.empty(copies * len(s_order), dtype=np.uint32)
            for c in range(copies):
                want[copies * s + c * L + (p - s)] = s_order + np.uint32(c * n_small)
            assert np.array_equal(order, want), "order differs first at %d" % int(np.argmax(order != want))
            assert int(order.max()) >= n - n_small  # indices of the last copy made it through
            # cov2d of the last copy equals the copy's own, bit for bit
            st_cov = st.read(n_cov=n)[1]
            assert np.array_equal(st_cov[n - n_small:][s_keep].view(np.uint32), s_cov[s_keep].view(np.uint32))
        finally:
            cloud.destroy()
    finally:
        st.destroy()
