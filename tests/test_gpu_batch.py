"""The batch path (many (cloud, view) jobs in one hipGraph launch, on the phased and the persistent
schedule) against the CPU oracle, job by job: order bit-exact, survivor count, cov2d within 1e-5."""
import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu


def check_job(state, recs, view, label):
    order, cov, stats = state.read(n_cov=len(recs))
    o_order, _, keep, o_cov = oracle.run(recs, view)
    assert stats.count == len(o_order), label
    assert np.array_equal(order, o_order), "%s: order differs" % label
    if keep.any():
        g, w = cov[keep, :3].astype(np.float64), o_cov[keep]
        err = np.max(np.abs(g - w).max(axis=1) / np.maximum(np.abs(w[:, 0]), np.abs(w[:, 2])))
        assert err <= 1e-5, "%s: cov2d error %g" % (label, err)


@pytest.mark.parametrize("schedule,super_tiles,kpt", [(wa.WD_SCHEDULE_PHASED, 2, 16), (wa.WD_SCHEDULE_PHASED, 1, 16),
                                                      (wa.WD_SCHEDULE_PERSISTENT, 2, 16),
                                                      (wa.WD_SCHEDULE_PHASED, 2, 8),
                                                      (wa.WD_SCHEDULE_PERSISTENT, 5, 8)])
def test_batch_uneven_jobs_and_replays(ctx, schedule, super_tiles, kpt):
    sizes = [1, 100, 4096, 5000, 70001, 300000, 12289, 65536]
    st = wl.Stream(11)
    jobs, data = [], []
    for j, n in enumerate(sizes):
        recs = wa.synth_records(j % 2, 100 + j, n)
        cloud = wa.Cloud(ctx, recs)
        state = wa.State(ctx, n)
        view = wl.view_for(recs, wl.orbit_pose(st), 1280, 720)
        if view is None:
            view = wl.view_for(recs, (0.0, 0.0, 10.0), 1280, 720)
        jobs.append((cloud, view, state))
        data.append(recs)
    batch = wa.Batch(ctx, len(jobs))
    try:
        batch.configure(schedule, super_tiles, kpt)
        batch.set(jobs)
        for replay in range(3):
            if replay == 2:  # new views, same clouds and states: the captured graph is replayed
                jobs = [(c, wl.view_for(r, wl.orbit_pose(st), 1920, 1080), s) for (c, _, s), r in zip(jobs, data)]
                batch.set(jobs)
            batch.run()
            ctx.sync()
            assert batch.stuck() == 0
            print("replay %d ok" % replay, flush=True)
            for (cloud, view, state), recs in zip(jobs, data):
                check_job(state, recs, view, "replay %d n=%d" % (replay, len(recs)))
        # a single-view sort on a state a batch used: epochs do not collide
        cloud, view, state = jobs[4]
        state.sort(cloud, view)
        check_job(state, data[4], view, "single after batch")
    finally:
        batch.destroy()
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()


@pytest.mark.parametrize("schedule,super_tiles,kpt", [(wa.WD_SCHEDULE_PHASED, 2, 8), (wa.WD_SCHEDULE_PERSISTENT, 2, 16)])
def test_batch_job_with_no_survivors(ctx, schedule, super_tiles, kpt):
    """A job whose view culls every splat (all behind the camera) between two that keep most: its count
    is 0 and its passes after the projection have no tiles; the jobs beside it still match the oracle."""
    jobs, data = [], []
    for j, n in enumerate([70001, 300000, 12289]):
        recs = wa.synth_records(j % 2, 300 + j, n)
        view = wl.view_for(recs, (0.0, 1.0, 10.0), 1280, 720)
        if j == 1:
            view.model_view[14] += 1000.0  # everything behind the camera
        jobs.append((wa.Cloud(ctx, recs), view, wa.State(ctx, n)))
        data.append(recs)
    batch = wa.Batch(ctx, len(jobs))
    try:
        batch.configure(schedule, super_tiles, kpt)
        batch.set(jobs)
        for replay in range(2):
            batch.run()
            ctx.sync()
            assert batch.stuck() == 0
            for j, ((cloud, view, state), recs) in enumerate(zip(jobs, data)):
                check_job(state, recs, view, "replay %d job %d" % (replay, j))
            assert jobs[1][2].stats().count == 0
            print("replay %d ok" % replay, flush=True)
    finally:
        batch.destroy()
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()


def test_batch_64_scenes_32m(ctx):
    seed, sizes = wl.batch_sizes()
    st = wl.Stream(seed)
    jobs, data = [], []
    try:
        for j, n in enumerate(sizes):
            recs = wa.synth_records(0, 1000 + j, n)
            view = None
            while view is None:
                view = wl.view_for(recs, wl.orbit_pose(st), 1280, 720)
            jobs.append((wa.Cloud(ctx, recs), view, wa.State(ctx, n)))
            data.append(recs)
        batch = wa.Batch(ctx, 64)
        batch.set(jobs)
        t = batch.tune(reps=3)
        rep = t.report()
        print("64 scenes, %d splats:" % sum(sizes), rep, flush=True)
        assert len(rep["timed_ms"]) == wa.WD_TUNE_CONFIGS and min(rep["timed_ms"].values()) > 0
        assert t.tile == 256 * t.keys_per_thread and (t.radix_bits, t.passes) == (8, 2)
        kept = [x for x in t.timed if (x.schedule, x.super_tiles, x.keys_per_thread)
                == (t.schedule, t.super_tiles, t.keys_per_thread)]
        assert kept and kept[0].ms == min(x.ms for x in t.timed)
        for which, sup, kpt in ((wa.WD_SCHEDULE_PERSISTENT, 8, 16), (wa.WD_SCHEDULE_PHASED, 1, 16),
                                (wa.WD_SCHEDULE_PHASED, 2, 16), (wa.WD_SCHEDULE_PHASED, 1, 8),
                                (wa.WD_SCHEDULE_PERSISTENT, 3, 8)):
            batch.configure(which, sup, kpt)
            for cloud, view, state in jobs:  # stale results must not pass for new ones
                state.sort(cloud, wl.view_for(data[0], (0.0, 0.0, 50.0), 64, 64) or view)
            batch.run()
            ctx.sync()
            assert batch.stuck() == 0
            for (cloud, view, state), recs in zip(jobs, data):
                check_job(state, recs, view, "schedule %d/%d/kpt%d scene n=%d" % (which, sup, kpt, len(recs)))
        batch.destroy()
    finally:
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()


def _small_jobs(ctx, sizes, seed):
    st = wl.Stream(seed)
    jobs, data = [], []
    for j, n in enumerate(sizes):
        recs = wa.synth_records(0, 500 + j, n)
        view = wl.view_for(recs, wl.orbit_pose(st), 1280, 720) or wl.view_for(recs, (0.0, 0.0, 10.0), 1280, 720)
        jobs.append((wa.Cloud(ctx, recs), view, wa.State(ctx, n)))
        data.append(recs)
    return jobs, data


def test_batch_refuses_a_destroyed_state_even_at_a_reused_address(ctx):
    """A job whose state was destroyed is refused at run time, also when a new state was given the
    same address (the batch compares each object's serial, not its address)."""
    jobs, data = _small_jobs(ctx, [5000, 70001], 21)
    batch = wa.Batch(ctx, 2)
    made = []
    try:
        batch.set(jobs)
        batch.run()
        ctx.sync()
        cloud, view, state = jobs[1]
        old = state.ptr.value
        state.destroy()
        for _ in range(64):  # the allocator most often hands the address out again at once
            made.append(wa.State(ctx, len(data[1])))
            if made[-1].ptr.value == old:
                break
        print("address reused: %s" % (made[-1].ptr.value == old), flush=True)
        with pytest.raises(wa.WindingError, match="destroyed; set the jobs again") as e:
            batch.run()
        assert e.value.code == -5
        jobs[1] = (cloud, view, made[-1])
        batch.set(jobs)  # set again: runs, and the results are the new state's
        batch.run()
        ctx.sync()
        for (c, v, s), recs in zip(jobs, data):
            check_job(s, recs, v, "after set again n=%d" % len(recs))
    finally:
        batch.destroy()
        for s in made[:-1]:
            s.destroy()
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()


def test_a_stuck_persistent_run_fails_its_states_once(ctx):
    """The persistent kernel's "a wait ran out" flag, raised from the host (no hang): the run's states
    then fail to read, the batch reports it once, and the next run is whole again."""
    jobs, data = _small_jobs(ctx, [4096, 100000, 300000], 33)
    batch = wa.Batch(ctx, 3)
    try:
        batch.schedule(wa.WD_SCHEDULE_PERSISTENT, 1)
        batch.set(jobs)
        batch.run()
        ctx.sync()
        assert batch.stuck() == 0
        batch.force_stuck()
        batch.run()
        ctx.sync()
        for _, _, state in jobs:
            with pytest.raises(wa.WindingError, match="bounded wait ran out"):
                state.stats()
        assert batch.stuck() == 1
        assert batch.stuck() == 0  # reported once
        batch.run()
        ctx.sync()
        assert batch.stuck() == 0
        for (c, v, s), recs in zip(jobs, data):
            check_job(s, recs, v, "after a stuck run n=%d" % len(recs))
    finally:
        batch.destroy()
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()


def test_batch_job_past_the_held_scan_rows(ctx):
    """A job of 9,000,000 splats beside a small one: at tiles of 4096 its 2,198 tiles are more than
    the 8 rows a scan thread holds (256 x 8 = 2,048 tiles), and at tiles of 2048 more again, so the
    batch scans take the path that reads the rows twice. Every schedule and tile, against the oracle."""
    sizes = [9_000_000, 3000]
    st = wl.Stream(31)
    jobs, data = [], []
    try:
        for j, n in enumerate(sizes):
            recs = wa.synth_records(j % 2, 700 + j, n)
            view = wl.view_for(recs, wl.orbit_pose(st), 1920, 1080) or wl.view_for(recs, (0.0, 0.0, 10.0), 1920, 1080)
            jobs.append((wa.Cloud(ctx, recs), view, wa.State(ctx, n)))
            data.append(recs)
        batch = wa.Batch(ctx, len(jobs))
        batch.set(jobs)
        for which, sup, kpt in ((wa.WD_SCHEDULE_PHASED, 1, 16), (wa.WD_SCHEDULE_PHASED, 2, 8),
                                (wa.WD_SCHEDULE_PERSISTENT, 8, 16), (wa.WD_SCHEDULE_PERSISTENT, 4, 8)):
            batch.configure(which, sup, kpt)
            batch.run()
            ctx.sync()
            assert batch.stuck() == 0
            print("schedule %d/%d/kpt%d ran" % (which, sup, kpt), flush=True)
            for (cloud, view, state), recs in zip(jobs, data):
                check_job(state, recs, view, "schedule %d/%d/kpt%d n=%d" % (which, sup, kpt, len(recs)))
        batch.destroy()
    finally:
        for cloud, _, state in jobs:
            cloud.destroy()
            state.destroy()
