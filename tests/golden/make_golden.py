"""Writes the golden fixtures of the splat path: small seeded clouds, their per-view uniform, and the
expected keys, survivor flags, draw order and FP64 cov2d.

Every expected value is computed twice, by the C restatement (oracle/splat_oracle.c) and by the
JavaScript one (oracle/splat_oracle.js, run in Node), and written only when the two agree exactly.

  python tests/golden/make_golden.py        (from the repository root; needs Node for the second oracle)

Fixtures (numpy .npz, no pickles):
  mixture_64     the first 64 splats of the 65,536-splat mixture cloud (seed 0), eye (0, 0, 10), 1280x720
  mixture_4096   the first 4,096 of the same cloud and view
  edges_4096     those 4,096 with every third splat moved onto a depth-bucket edge (+-1e-4 of a bucket)
                 and every fourth survivor rescaled so its opacity x area sits on the 0.5 cull threshold
  slab_4096      4,096 splats of the slab generator (heavy key collisions), 1920x1080
  cactus_7070    every 64th splat of the reference's own capture, demo/cactus.splat (452,451 splats, CC0),
                 placed and seen as its demo shows it: turned upright (world diag(1, -1, -1)), eye
                 (0, 1.6, 3.2) looking at (0, 0.7, 0), fovY pi/3, near 0.05, 1280x720; written only when
                 the reference tree is at /root/reference (the committed fixture carries the records)
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import winding_amd as wa  # noqa: E402


def js_oracle(recs, view):
    with tempfile.TemporaryDirectory() as tmp:
        rp, vp, out = os.path.join(tmp, "r.bin"), os.path.join(tmp, "v.bin"), os.path.join(tmp, "o")
        np.ascontiguousarray(recs).tofile(rp)
        view.as_array().tofile(vp)
        subprocess.run(["node", os.path.join(ROOT, "oracle", "run_js.js"), rp, vp, out], check=True)
        keys = np.fromfile(out + "_keys.bin", np.int32)
        order = np.fromfile(out + "_order.bin", np.uint32)
        cov = np.fromfile(out + "_cov.bin", np.float64).reshape(-1, 3)
    return keys, order, cov


def edges_case(recs, view):
    recs = recs.copy()
    pos = recs.view(np.float32).reshape(-1, 8)
    mv = np.array(view.model_view, np.float64).reshape(4, 4).T
    rng = np.random.default_rng(9)
    for i in range(0, len(recs), 3):
        p = pos[i, :3].astype(np.float64)
        d = -(mv[2, :3] @ p + mv[2, 3])
        t = (d - view.depth_min) * view.depth_scale
        edge = round(t) + rng.choice([-1e-4, 0, 1e-4])
        dd = edge / view.depth_scale + view.depth_min
        pos[i, :3] = (p + (d - dd) * mv[2, :3]).astype(np.float32)
    _, _, _, val = oracle.keys(recs, view)
    for i in np.nonzero(~np.isnan(val))[0][1::4]:
        pos[i, 3:6] *= np.float32(min(4.0, max(0.25, math.sqrt(0.5 / max(val[i], 1e-9)))))
    return recs


def cases():
    recs, view, _ = wa.workloads.plain("plain-64k")
    yield "mixture_64", recs[:64].copy(), wa.workloads.view_for(recs[:64], (0.0, 0.0, 10.0), 1280, 720)
    r4 = recs[:4096].copy()
    v4 = wa.workloads.view_for(r4, (0.0, 0.0, 10.0), 1280, 720)
    yield "mixture_4096", r4, v4
    yield "edges_4096", edges_case(r4, v4), v4
    slab = wa.synth_records(1, 0, 4096)
    yield "slab_4096", slab, wa.workloads.view_for(slab, (0.0, 0.0, 10.0), 1920, 1080)
    capture = "/root/reference/demo/cactus.splat"
    if os.path.exists(capture):
        cactus = np.fromfile(capture, np.uint8).reshape(-1, 32)[::64].copy()
        v = wa.workloads.look_at((0.0, 1.6, 3.2), (0.0, 0.7, 0.0), (0.0, 1.0, 0.0))
        p = wa.workloads.perspective(math.pi / 3, 1280 / 720, 0.05)
        world = np.diag([1.0, -1.0, -1.0, 1.0]).astype(np.float32).ravel()
        bad, mn, mx = oracle.bounds(cactus)
        assert bad == -1
        yield "cactus_7070", cactus, wa.prepare_view(v, p, 1280.0, 720.0, mn, mx, 0.5, world)


def main():
    for name, recs, view in cases():
        order, keys, keep, cov = oracle.run(recs, view)
        j_keys, j_order, j_cov = js_oracle(recs, view)
        c_keys = np.where(keep, keys.astype(np.int32), -1)
        assert np.array_equal(j_keys, c_keys), name + ": JS and C keys differ"
        assert np.array_equal(j_order, order), name + ": JS and C order differ"
        c_cov = np.where(keep[:, None], cov, np.nan)
        assert np.array_equal(np.isnan(j_cov), np.isnan(c_cov)) and np.array_equal(
            j_cov[keep], c_cov[keep]), name + ": JS and C cov2d differ"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), records=np.ascontiguousarray(recs, np.uint8),
                            view=view.as_array(), keys=c_keys, order=order.astype(np.uint32), cov=c_cov)
        print("%s: n=%d V=%d" % (name, len(recs), len(order)))


if __name__ == "__main__":
    main()
