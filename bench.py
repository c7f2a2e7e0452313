"""Benchmark of the splat hot path on one MI355X: project + key + stable back-to-front sort.

One step = one sort of one view of the workload's cloud (records resident in HBM), the
reference's makeKeys -> scanCounts -> place (src/render/splats.js:155-219). Prints one JSON line.

  value     M splats projected+sorted per second (input splats N per step / step time)
  roofline  the dominant kernel (k_project) against HBM: its algorithmic bytes per launch
            (32 N record bytes read + 16 V cov2d/colour bytes written) / its mean duration,
            from HIP events recorded on the library's stream around each launch
  path      the whole step against HBM: B = 32 N + 20 V bytes / step time
  cpu_baseline  the CPU oracle on this host, one thread, on a bounded sample of the same cloud
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter runs (the `traffic:` lines of
    profiles/*.md: FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a copy of known size), or None."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "*.md")):
        for line in open(path):
            f = line.split()
            if len(f) == 5 and f[0] == "traffic:" and f[1] == workload and f[2] == kernel:
                return float(f[3]) + float(f[4])
    return None


def cpu_baseline(recs, view, budget_s=15.0):
    """The JavaScript CPU path (oracle/splat_oracle.js, a one-thread restatement of the reference's
    put + makeKeys + stable counting sort) in Node on this host, on a prefix of the cloud; the C
    restatement when Node is absent."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    n = min(len(recs), 1 << 20)
    sample = np.ascontiguousarray(recs[:n])
    node = shutil.which("node")
    if node:
        with tempfile.TemporaryDirectory() as tmp:
            rp, vp = os.path.join(tmp, "records.bin"), os.path.join(tmp, "view.bin")
            sample.tofile(rp)
            view.as_array().tofile(vp)
            out = subprocess.run([node, os.path.join(ROOT, "oracle", "run_js.js"), rp, vp, "--time", str(budget_s)],
                                 capture_output=True, text=True, timeout=budget_s * 4 + 60)
        if out.returncode == 0 and out.stdout.strip():
            r = json.loads(out.stdout.strip().splitlines()[-1])
            return {"value": r["splats_per_s"] / 1e6, "unit": "M splats/s", "cores": r["threads"], "kind": "port",
                    "sample": "first %d splats of the same cloud and view; median of %d runs of the JS CPU path "
                              "(oracle/splat_oracle.js) in Node %s, 1 thread, on a %d-CPU host (%s)"
                              % (n, r["runs"], r["node"], r["host_cpus"], r["cpu_model"].strip())}
    import oracle
    t0 = time.perf_counter()
    oracle.run(sample, view)
    t = time.perf_counter() - t0
    reps = max(1, min(5, int(budget_s / max(t, 1e-3))))
    times = [t]
    for _ in range(reps - 1):
        t0 = time.perf_counter()
        oracle.run(sample, view)
        times.append(time.perf_counter() - t0)
    best = sorted(times)[len(times) // 2]
    return {"value": n / best / 1e6, "unit": "M splats/s", "cores": 1, "kind": "port",
            "sample": "first %d splats of the same cloud and view, median of %d runs, oracle/splat_oracle.c" % (n, len(times))}


def bench_particles(args):
    """particles-8m: one emitter ring of 8,388,608 particles (a quarter dead, 5% never used), sorted far to near
    (wd_particles_sort): the key pass, then the splat sort's five passes."""
    import numpy as np
    import winding_amd as wa
    n = 8 << 20
    pool = wa.workloads.particle_pool(n, seed=11)
    ctx = wa.Context(0)
    p = wa.Pool(ctx, pool)
    st = wa.State(ctx, n)
    view = wa.particle_view(wa.workloads.PARTICLE_EYE, wa.workloads.PARTICLE_FORWARD)
    for _ in range(args.warmup):
        st.sort_particles(p, view)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.sort_particles(p, view)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    v = st.stats().count
    ctx.profile(True)
    for _ in range(args.steps):
        st.sort_particles(p, view)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    kernels = {k: {"ms_mean": msum / cnt, "launches": cnt} for k, (msum, cnt) in prof.items()}
    key_ms = kernels["k_particle_keys"]["ms_mean"]
    alg_key = 32.0 * n + 2.0 * n  # the 32-byte particles read, the 16-bit keys written
    line = {
        "metric": "M particles keyed+sorted/s", "value": n / (ms * 1e-3) / 1e6, "unit": "M particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded generator, winding_amd/workloads.py particle_pool)",
        "config": {"workload": "particles-8m", "particles": n, "live": v, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": alg_key / (key_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg_key / (key_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("particles-8m", "k_particle_keys"),
                     "kernel": "k_particle_keys", "kernel_ms": key_ms, "bytes_per_launch": alg_key},
        "path": {"bytes_per_step": 32.0 * n + 4.0 * v,
                 "frac_of_peak": (32.0 * n + 4.0 * v) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)
    st.destroy()
    p.destroy()
    ctx.close()
    return 0


def synth_ply(n, seed=3):
    """A binary .ply of n splats with degree-3 harmonics, as training writes it: 62 floats a vertex."""
    import numpy as np
    rng = np.random.default_rng(seed)
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
             "rot_0", "rot_1", "rot_2", "rot_3"] + ["f_rest_%d" % k for k in range(45)] + ["nx", "ny", "nz"]
    v = np.empty((n, len(names)), np.float32)
    v[:, :3] = rng.normal(0, 3, (n, 3))
    v[:, 3:6] = rng.normal(0, 1.5, (n, 3))
    v[:, 6] = rng.normal(0, 3, n)
    v[:, 7:10] = rng.normal(-3, 1, (n, 3))
    v[:, 10:14] = rng.normal(0, 1, (n, 4))
    v[:, 14:59] = rng.normal(0, 0.3, (n, 45))
    v[:, 59:] = 0
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    head += "".join("property float %s\n" % nm for nm in names) + "end_header\n"
    return np.frombuffer(head.encode() + v.tobytes(), np.uint8)


def synth_spz(n, seed=4):
    """A gunzipped .spz of version 3 with degree-3 harmonics: 65 bytes a splat in its sections."""
    import numpy as np
    rng = np.random.default_rng(seed)
    head = np.frombuffer(np.array([0x5053474E, 3, n], "<u4").tobytes() + bytes([3, 12, 0, 0]), np.uint8)
    body = rng.integers(0, 256, n * (9 + 1 + 3 + 3 + 4 + 45), dtype=np.uint8)
    rot = body[n * 16:n * 20].view("<u4")
    rot &= np.uint32(0x3FFFFFFF)
    rot |= (rng.integers(0, 4, n).astype(np.uint32) << np.uint32(30))
    return np.concatenate([head, body])


def bench_decode(args):
    """decode-1m: 1,048,576 splats with degree-3 harmonics as a .ply (248 bytes a splat) and as a
    gunzipped .spz of version 3 (65 bytes a splat), each decoded into a cloud of centres, Sigma and
    harmonics. The kernel is timed with HIP events; the whole call (header, upload, decode, box and
    check read back, the cloud's allocations) with the host clock."""
    import winding_amd as wa
    n = 1 << 20
    ctx = wa.Context(0)
    out = {}
    for kind, data, fn, in_bytes in (("ply", synth_ply(n), wa.decode_ply, 248.0),
                                     ("spz", synth_spz(n), wa.decode_spz, 65.0)):
        for _ in range(max(1, args.warmup // 4)):
            fn(ctx, data).destroy()
        ctx.profile(True)
        reps = max(5, args.steps // 10)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(ctx, data).destroy()
        call_ms = (time.perf_counter() - t0) * 1e3 / reps
        prof = ctx.profile_read()
        ctx.profile(False)
        k_ms = prof["k_decode"][0] / prof["k_decode"][1]
        alg = (in_bytes + 16 + 24 + 92) * n  # the file's splat bytes read; centre, Sigma and harmonics written
        out[kind] = {"kernel_ms": k_ms, "call_ms": call_ms, "splats_per_s_kernel": n / (k_ms * 1e-3),
                     "splats_per_s_call": n / (call_ms * 1e-3), "bytes_per_launch": alg,
                     "frac_of_peak": alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ply = out["ply"]
    line = {
        "metric": "M splats decoded/s (.ply, degree 3)", "value": ply["splats_per_s_kernel"] / 1e6,
        "unit": "M splats/s", "n_gpus": 1, "steps": reps, "warmup": max(1, args.warmup // 4),
        "ms_per_step": ply["kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded numpy .ply and .spz)",
        "config": {"workload": "decode-1m", "splats": n, "sh_degree": 3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ply["bytes_per_launch"] / (ply["kernel_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ply["frac_of_peak"],
                     "traffic": measured_traffic("decode-1m", "k_decode_ply"),
                     "kernel": "k_decode_ply", "kernel_ms": ply["kernel_ms"], "bytes_per_launch": ply["bytes_per_launch"]},
        "decode": out,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def bench_batch(args, emit=True):
    """64 scenes of log-uniform size (about 32M splats), one random orbit view each: one hipGraph launch
    of the batch (schedule picked by the autotuner) against the same 64 sorts as single-view launches."""
    import winding_amd as wa
    wl = wa.workloads
    seed, sizes = wl.batch_sizes()
    st = wl.Stream(seed)
    ctx = wa.Context(0)
    jobs = []
    for j, n in enumerate(sizes):
        recs = wa.synth_records(0, 1000 + j, n)
        view = None
        while view is None:
            view = wl.view_for(recs, wl.orbit_pose(st), 1280, 720, cull=args.cull)
        jobs.append((wa.Cloud(ctx, recs), view, wa.State(ctx, n)))
        del recs
    total = sum(sizes)
    batch = wa.Batch(ctx, len(jobs))
    batch.set(jobs)
    tuning = batch.tune(reps=5)
    for _ in range(args.warmup):
        batch.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    v = sum(state.stats().count for _, _, state in jobs)
    # the same 64 sorts, one launch sequence each
    for _ in range(max(1, args.warmup // 4)):
        for cloud, view, state in jobs:
            state.sort(cloud, view)
    ctx.sync()
    reps = max(3, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        for cloud, view, state in jobs:
            state.sort(cloud, view)
    ctx.sync()
    ms_plain = (time.perf_counter() - t0) * 1e3 / reps
    path_bytes = 32.0 * total + 20.0 * v
    line = {
        "metric": "M splats projected+sorted/s and % HBM peak; batched vs plain",
        "value": total / (ms * 1e-3) / 1e6,
        "unit": "M splats/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded generator, winding_amd/workloads.py)",
        "config": {"workload": "batch-64", "scenes": len(sizes), "splats": total, "survivors": v,
                   "width": 1280, "height": 720, "cull": args.cull, "size_seed": seed,
                   "parallelism": "single GPU"},
        "batch": {**tuning.report(), "plain_ms": ms_plain,
                  "plain_value": total / (ms_plain * 1e-3) / 1e6, "speedup_vs_plain": ms_plain / ms},
        "path": {"bytes_per_step": path_bytes, "achieved_gbs": path_bytes / (ms * 1e-3) / 1e9,
                 "frac_of_peak": path_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    batch.destroy()
    for cloud, _, state in jobs:
        cloud.destroy()
        state.destroy()
    ctx.close()
    if not emit:
        return line
    print(json.dumps(line), flush=True)
    return 0


def bench_learned(args, emit=True):
    """4,194,304 splats, a 1024-pose orbit. The predictor is trained first (untimed) on the exact sort's
    quantiles at 4096 random poses of the same family (a disjoint seed). Timed: the whole orbit sorted
    (a) plain, one wd_sort a pose; (b) batched, 8 poses a phased batch launch; (c) batched + learned, 8 poses
    a wd_learned_run (one read of the records for the 8, bins from the predictor, exact fallback gated)."""
    import numpy as np
    import winding_amd as wa
    wl = wa.workloads
    spec = wl.LEARNED
    n, W, H, K = spec["n"], spec["width"], spec["height"], 8
    ctx = wa.Context(0)
    recs = wa.synth_records(spec["kind"], spec["seed"], n)
    cloud = wa.Cloud(ctx, recs)
    del recs
    states = [wa.State(ctx, n) for _ in range(K)]
    t0 = time.perf_counter()
    X, Y = wl.training_set(ctx, cloud, states[0], wl.training_poses(spec["train_poses"], spec["train_seed"]), W, H,
                           args.cull)
    t_data = time.perf_counter() - t0
    mlp = wa.Mlp(ctx, seed=1)
    before, _ = mlp.grad(X, Y)
    t0 = time.perf_counter()
    mlp.train(X, Y, epochs=40, batch=256, lr=2e-3, seed=2)
    t_train = time.perf_counter() - t0
    after, _ = mlp.grad(X, Y)
    entropy = float(-(Y * np.log(np.maximum(Y, 1e-30))).sum(axis=1).mean())
    views = [wl.view_at(cloud, e, W, H, args.cull) for e in wl.orbit_sequence(1024)]
    groups = [(views[i:i + K], states[:len(views[i:i + K])]) for i in range(0, len(views), K)]
    learned = wa.Learned(ctx, K)
    batch = wa.Batch(ctx, K)
    batch.set([(cloud, v, s) for v, s in zip(views[:K], states)])
    batch.tune(reps=5)

    def orbit_plain():
        for v in views:
            states[0].sort(cloud, v)

    def orbit_batched():
        for vs, sts in groups:
            batch.set([(cloud, v, s) for v, s in zip(vs, sts)])
            batch.run()

    def orbit_learned():
        for vs, sts in groups:
            learned.run(mlp, cloud, vs, sts)

    def timed(fn, reps):
        fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t) / reps

    reps = max(1, args.steps // 50)
    t_plain = timed(orbit_plain, reps)
    t_batched = timed(orbit_batched, reps)
    t_learned = timed(orbit_learned, reps)
    # fallbacks over one pass of the orbit, and the survivors of every pose
    fell, v_total = 0, 0
    for vs, sts in groups:
        learned.run(mlp, cloud, vs, sts)
        for s in sts:
            st = s.stats()
            fell += st.fallback
            v_total += st.count
    poses = len(views)
    path_bytes = 32.0 * n * poses + 20.0 * v_total
    line = {
        "metric": "M splats projected+sorted/s and % HBM peak; batched+learned vs plain",
        "value": n * poses / t_learned / 1e6,
        "unit": "M splats/s",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": t_learned * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded generator, winding_amd/workloads.py)",
        "config": {"workload": "learned-4m", "splats": n, "poses": poses, "survivors_total": v_total,
                   "width": W, "height": H, "cull": args.cull, "views_per_run": K, "parallelism": "single GPU"},
        "learned": {"plain_value": n * poses / t_plain / 1e6, "batched_value": n * poses / t_batched / 1e6,
                    "speedup_vs_plain": t_plain / t_learned, "speedup_batched_exact_vs_plain": t_plain / t_batched,
                    "fallback_views": fell, "train_poses": len(X), "train_loss": [before, after],
                    "target_entropy": entropy, "train_data_s": t_data, "train_s": t_train,
                    "ms_per_orbit": {"plain": t_plain * 1e3, "batched": t_batched * 1e3, "learned": t_learned * 1e3}},
        "path": {"bytes_per_step": path_bytes, "achieved_gbs": path_bytes / t_learned / 1e9,
                 "frac_of_peak": path_bytes / t_learned / 1e9 / HBM_PEAK_GBS},
    }
    learned.destroy()
    batch.destroy()
    mlp.destroy()
    for s in states:
        s.destroy()
    cloud.destroy()
    ctx.close()
    if not emit:
        return line
    print(json.dumps(line), flush=True)
    return 0


def second_half(args):
    """The metric's second half, measured in the same run as the slab-8m line: batch-64 (the batched form
    against per-scene sorts) and learned-4m (plain, batched exact and batched + learned over the orbit)."""
    b = bench_batch(args, emit=False)
    lr = bench_learned(args, emit=False)
    bb, ll = b["batch"], lr["learned"]
    batch = {"workload": "batch-64", "scenes": b["config"]["scenes"], "splats": b["config"]["splats"],
             "survivors": b["config"]["survivors"], "replay_ms": b["ms_per_step"],
             "per_scene_ms": b["ms_per_step"] / b["config"]["scenes"], "value": b["value"],
             "plain_ms": bb["plain_ms"], "speedup_vs_plain": bb["speedup_vs_plain"],
             "schedule": bb["schedule"],
             "tuning": {k: v for k, v in bb.items() if k not in ("plain_ms", "plain_value", "speedup_vs_plain")},
             "path_frac_of_peak": b["path"]["frac_of_peak"]}
    learned = {"workload": "learned-4m", "splats": lr["config"]["splats"], "poses": lr["config"]["poses"],
               "ms_per_orbit": ll["ms_per_orbit"], "value": lr["value"], "plain_value": ll["plain_value"],
               "batched_value": ll["batched_value"], "speedup_vs_plain": ll["speedup_vs_plain"],
               "speedup_batched_exact_vs_plain": ll["speedup_batched_exact_vs_plain"],
               "fallback_views": ll["fallback_views"], "train_poses": ll["train_poses"],
               "train_loss": ll["train_loss"], "target_entropy": ll["target_entropy"],
               "path_frac_of_peak": lr["path"]["frac_of_peak"]}
    return batch, learned


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="slab-8m",
                    choices=["plain-64k", "plain-1m", "slab-8m", "batch-64", "learned-4m", "sh3-1m", "particles-8m", "decode-1m"])
    ap.add_argument("--cull", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-second-half", action="store_true",
                    help="slab-8m only: leave out the batch-64 and learned-4m objects of the default line")
    ap.add_argument("--calibrate", action="store_true",
                    help="time a plain 512 MiB device copy instead (HBM yardstick for the traffic counters)")
    args = ap.parse_args()
    if args.calibrate:
        import winding_amd as wa
        ctx = wa.Context(0)
        nbytes = 512 << 20
        ms = ctx.copy_ms(nbytes, reps=args.steps)
        ctx.close()
        print(json.dumps({"calibration": "device copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "ms_mean": ms, "gbs": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not run (each box has one MI355X, "
              "and the multi-GPU split is a design note, not code)" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import winding_amd as wa
    if args.workload == "batch-64":
        return bench_batch(args)
    if args.workload == "learned-4m":
        return bench_learned(args)
    if args.workload == "particles-8m":
        return bench_particles(args)
    if args.workload == "decode-1m":
        return bench_decode(args)
    # sh3-1m: the 1,048,576-splat cloud with degree-3 harmonics (seeded), shaded in the projection
    sh_degree = 3 if args.workload == "sh3-1m" else 0
    recs, view, spec = wa.workloads.plain("plain-1m" if sh_degree else args.workload, cull=args.cull)
    n = len(recs)
    ctx = wa.Context(0)
    cloud = wa.Cloud(ctx, recs)
    sh_bytes = 0
    if sh_degree:
        words = wa.workloads.sh_words(n, sh_degree, seed=7)
        cloud.set_sh(words, sh_degree)
        sh_bytes = 4.0 * (words.size // n)
    st = wa.State(ctx, n)
    for _ in range(args.warmup):
        st.sort(cloud, view)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.sort(cloud, view)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1e3 / args.steps
    stats = st.stats()
    v = stats.count

    # Kernel durations from HIP events on the library's stream, over a second timed region.
    ctx.profile(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        st.sort(cloud, view)
    ctx.sync()
    prof_elapsed = time.perf_counter() - t1
    prof = ctx.profile_read()
    ctx.profile(False)
    kernels = {k: {"ms_mean": msum / cnt, "launches": cnt} for k, (msum, cnt) in prof.items()}
    proj_ms = kernels["k_project"]["ms_mean"]
    alg_proj = 32.0 * n + (16.0 + sh_bytes) * v  # the survivors' harmonics are read once, by the projection
    achieved = alg_proj / (proj_ms * 1e-3) / 1e9
    path_bytes = 32.0 * n + (20.0 + sh_bytes) * v
    path_gbs = path_bytes / (ms * 1e-3) / 1e9

    # The spread of single sorts: HIP events around each, on the library's stream.
    each = np.sort(st.time_sorts(cloud, view, max(100, args.steps)))
    latency = {"median_ms": float(np.median(each)), "p95_ms": float(each[int(0.95 * (len(each) - 1))]),
               "sorts": int(len(each))}

    # The same cloud and view with the area cull off (cull = 0): more survivors, the same path.
    view0 = wa.workloads.plain("plain-1m" if sh_degree else args.workload, cull=0.0)[1]
    for _ in range(args.warmup):
        st.sort(cloud, view0)
    ctx.sync()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        st.sort(cloud, view0)
    ctx.sync()
    ms0 = (time.perf_counter() - t2) * 1e3 / args.steps
    v0 = st.stats().count
    cull0 = {"value": n / (ms0 * 1e-3) / 1e6, "ms_per_step": ms0, "survivors": v0,
             "path_frac_of_peak": (32.0 * n + (20.0 + sh_bytes) * v0) / (ms0 * 1e-3) / 1e9 / HBM_PEAK_GBS}

    # The same sort with the draw's ellipse set-up written beside cov2d (32 more bytes a splat).
    st.set_ellipses(True)
    for _ in range(args.warmup):
        st.sort(cloud, view)
    ctx.sync()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        st.sort(cloud, view)
    ctx.sync()
    ms_e = (time.perf_counter() - t3) * 1e3 / args.steps
    ctx.profile(True)
    for _ in range(args.steps):
        st.sort(cloud, view)
    ctx.sync()
    proj_e = ctx.profile_read()["k_project"]
    ctx.profile(False)
    st.set_ellipses(False)
    proj_e_ms = proj_e[0] / proj_e[1]
    alg_e = alg_proj + 32.0 * v  # the survivors' set-up (culled slots are written too, as zeros)
    ellipses = {"value": n / (ms_e * 1e-3) / 1e6, "ms_per_step": ms_e, "k_project_ms": proj_e_ms,
                "k_project_frac_of_peak": alg_e / (proj_e_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    line = {
        "metric": "M splats projected+sorted/s and % HBM peak",
        "value": n / (ms * 1e-3) / 1e6,
        "unit": "M splats/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, winding_amd/workloads.py)",
        "config": {"workload": args.workload, "splats": n, "survivors": v, "width": spec["width"],
                   "height": spec["height"], "cull": args.cull, "sh_degree": sh_degree, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, "k_project"),
                     "kernel": "k_project",
                     "kernel_ms": proj_ms, "bytes_per_launch": alg_proj},
        "path": {"bytes_per_step": path_bytes, "achieved_gbs": path_gbs, "frac_of_peak": path_gbs / HBM_PEAK_GBS,
                 "recomputed_fp64": stats.recomputed, "profiled_ms_per_step": prof_elapsed * 1e3 / args.steps},
        "kernels": kernels,
        "latency": latency,
        "cull0": cull0, "ellipses": ellipses,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(recs, view)
    st.destroy()
    cloud.destroy()
    ctx.close()
    if args.workload == "slab-8m" and not args.no_second_half:
        line["batch"], line["learned"] = second_half(args)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
