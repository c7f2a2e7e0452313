"""The particle sort (src/render/particles.js:169-240): the same stable LSD passes as the splat sort,
after a key pass of its own (4096 log-depth buckets, far first).

The oracle (oracle/splat_oracle.c or_particle_keys, or_particle_edges) is pinned by:
- its 4095 bucket edges against 2^((30 b - 40950) / 4095) to 40 digits (Python's decimal): within one
  ulp of exp2 at the FP64 exponent, and eight of the true edge;
- the closed form of the bucket at exact powers of two;
- floor((log2 d + 10) 4095 / 30) on random depths away from an edge;
- the live test as the shader writes it, and both clamps;
- the reference's known answer (test/gpu.test.js:2666-2760: an alpha emitter's particles come out far
  to near whatever slots they were born in);
- a brute-force stable sort.
The GPU is held to it bit for bit.
"""
import math
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle
import winding_amd as wa

EYE = wa.workloads.PARTICLE_EYE
FWD = wa.workloads.PARTICLE_FORWARD


def particle(pos, age=0.5, lifetime=2.0, vel=(0.0, 0.0, 0.0)):
    return np.frombuffer(struct.pack("<8f", *pos, age, *vel, lifetime), np.uint8).copy()


def pool_of(parts):
    return np.stack(parts).reshape(-1, 32)


def test_edges_are_the_powers_of_two_to_forty_digits():
    getcontext().prec = 40
    edges = oracle.particle_edges()
    assert edges[0] == 0 and np.all(np.diff(edges[1:]) > 0)
    ln2 = Decimal(2).ln()
    for b in list(range(1, 4096, 7)) + [1, 1365, 2730, 4095]:
        got = Decimal(float(edges[b]))
        # exp2 of the FP64 exponent (30 b - 40950) / 4095, to one ulp ...
        at = (Decimal(float((30.0 * b - 40950.0) / 4095.0)) * ln2).exp()
        assert abs(got - at) <= at * Decimal(2) ** -52, b
        # ... and so within 8 ulps of the true edge (the exponent's own rounding, half an ulp of up to 20,
        # times ln 2, adds up to about 6)
        exact = ((Decimal(30 * b) - 40950) / 4095 * ln2).exp()
        assert abs(got - exact) <= exact * Decimal(2) ** -49, b


def test_bucket_closed_form_at_powers_of_two():
    parts, want = [], []
    for k in range(-12, 23):
        parts.append(particle((0.0, 0.0, EYE[2] - 2.0 ** k)))
        want.append(min(max(math.floor((k + 10) * 4095 / 30), 0), 4095))
    keys, keep = oracle.particle_keys(pool_of(parts), EYE, FWD)
    assert keep.all()
    assert list(4095 - keys.astype(int)) == want


def test_bucket_is_floor_of_the_log_depth_away_from_edges():
    rng = np.random.default_rng(2)
    d = np.exp2(rng.uniform(-11, 21, 3000))
    parts = [particle((0.0, 0.0, float(np.float32(EYE[2] - x)))) for x in d]
    pool = pool_of(parts)
    keys, keep = oracle.particle_keys(pool, EYE, FWD)
    z = pool.view(np.float32).reshape(-1, 8)[:, 2].astype(np.float64)
    depth = np.maximum((z - EYE[2]) * FWD[2], float(np.float32(1e-6)))
    t = (np.log2(depth) + 10) * 4095 / 30
    clear = np.abs(t - np.round(t)) > 1e-9
    want = np.clip(np.floor(t), 0, 4095).astype(int)
    assert clear.sum() > 2900
    assert np.array_equal((4095 - keys.astype(int))[clear], want[clear])


def test_live_test_and_clamps():
    nan = float("nan")
    parts = [
        particle((0, 0, 90.0), age=2.0, lifetime=2.0),     # dead: age reached lifetime
        particle((0, 0, 90.0), age=0.0, lifetime=0.0),     # never used (a zeroed slot)
        particle((0, 0, 90.0), age=nan, lifetime=1.0),     # NaN >= x is false: live, as the shader tests it
        particle((0, 0, 150.0)),                           # behind the eye: depth clamps to 1e-6, nearest
        particle((0, 0, 100.0 - 2.0 ** 25)),               # past the far end: bucket 4095, farthest
    ]
    keys, keep = oracle.particle_keys(pool_of(parts), EYE, FWD)
    assert list(keep) == [False, False, True, True, True]
    assert keys[3] == 4095 and keys[4] == 0
    assert keys[2] == 4095 - math.floor((math.log2(10.0) + 10) * 4095 / 30)


def test_known_answer_far_to_near_whatever_the_slots():
    # test/gpu.test.js:2666-2760: particles born in any order draw far to near
    depths = [3.0, 40.0, 0.5, 12.0, 200.0, 7.0]
    pool = pool_of([particle((0.1 * k, 0.0, EYE[2] - d)) for k, d in enumerate(depths)])
    order, keys, keep = oracle.particle_run(pool, EYE, FWD)
    assert list(order) == list(np.argsort(depths)[::-1])


def test_order_is_a_stable_sort_by_key():
    pool = wa.workloads.particle_pool(20000, seed=5)
    order, keys, keep = oracle.particle_run(pool, EYE, FWD)
    live = np.nonzero(keep)[0]
    want = live[np.lexsort((live, keys[live]))]
    assert np.array_equal(order, want)
    assert 0.6 < len(order) / len(pool) < 0.8
    k = 4095 - keys[live].astype(int)
    assert k.min() == 0 and k.max() == 4095  # both clamps are reached


# ---- the GPU, held to the oracle ----

@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 4095, 70001, 1 << 20])
def test_gpu_particle_order_is_bit_exact(ctx, n):
    pool = wa.workloads.particle_pool(n, seed=n)
    want, _, keep = oracle.particle_run(pool, EYE, FWD)
    p = wa.Pool(ctx, pool)
    st = wa.State(ctx, n)
    try:
        st.sort_particles(p, wa.particle_view(EYE, FWD))
        order, _, stats = st.read()
        print("%d particles: %d live" % (n, stats.count), flush=True)
        assert stats.count == len(want)
        assert np.array_equal(order, want)
    finally:
        st.destroy()
        p.destroy()


@pytest.mark.gpu
def test_gpu_pool_with_no_live_particles(ctx):
    """Every particle's age has reached its lifetime: nothing is drawn, the count is 0, and the same state
    then sorts a live pool exactly."""
    pool = wa.workloads.particle_pool(5000, seed=3)
    words = pool.view(np.float32).reshape(len(pool), -1)
    dead = pool.copy()
    dead.view(np.float32).reshape(len(pool), -1)[:, 3] = words[:, 7]  # age = lifetime
    want_dead, _, keep = oracle.particle_run(dead, EYE, FWD)
    assert len(want_dead) == 0 and not keep.any()
    st = wa.State(ctx, 5000)
    try:
        for data, want in ((dead, want_dead), (pool, oracle.particle_run(pool, EYE, FWD)[0])):
            p = wa.Pool(ctx, data)
            try:
                st.sort_particles(p, wa.particle_view(EYE, FWD))
                order, _, stats = st.read()
                assert stats.count == len(want)
                assert np.array_equal(order, want)
            finally:
                p.destroy()
    finally:
        st.destroy()


@pytest.mark.gpu
def test_gpu_emitter_rings_of_one_pool(ctx):
    # two emitters' rings back to back in one pool: each sort gives its ring's slots, far to near
    pool = wa.workloads.particle_pool(50000, seed=9)
    p = wa.Pool(ctx, pool)
    st = wa.State(ctx, 30000)
    view = wa.particle_view(EYE, FWD)
    try:
        for first, count in ((0, 20000), (20000, 30000)):
            st.sort_particles(p, view, first, count)
            order, _, stats = st.read()
            want, _, _ = oracle.particle_run(pool[first:first + count], EYE, FWD)
            assert np.array_equal(order, want + first)
        with pytest.raises(wa.WindingError, match="slots 40000 to 60000"):
            st.sort_particles(p, view, 40000, 20000)
        big = wa.Pool(ctx, wa.workloads.particle_pool(40000, seed=1))
        try:
            with pytest.raises(wa.WindingError, match="more than this state holds"):
                st.sort_particles(big, view)
        finally:
            big.destroy()
    finally:
        st.destroy()
        p.destroy()
