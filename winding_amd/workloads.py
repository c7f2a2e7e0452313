"""The measured workloads: seeded inputs and cameras, by name.

  plain-64k   65,536 splats, Gaussian mixture, seed 0, eye (0,0,10), fovY 60deg, 1280x720
  plain-1m    1,048,576 splats, the same generator and camera
  slab-8m     8,388,608 splats, 80% in a slab |z|<=0.25 at distance 10, 20% in a 40^3 box, 1920x1080
  batch-64    64 scenes, sizes log-uniform in [1e4, 2e6] (about 32M splats), one random orbit view each
  learned-4m  4,194,304 splats, mixture generator, a 1024-pose orbit; the predictor trains on 4096 random poses

The cull is the reference's default, 0.5 (src/render/splats.js:78), unless a caller passes another.
"""
import math

import numpy as np

from . import WD_SPLAT_CULL, look_at, perspective, prepare_view, synth_records

NEAR = 0.1
FOVY = math.pi / 3


class Stream:
    """splitmix64, the generator's own stream, for sizes and poses."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


PLAIN = {
    "plain-64k": dict(n=65536, kind=0, seed=0, eye=(0.0, 0.0, 10.0), width=1280, height=720),
    "plain-1m": dict(n=1048576, kind=0, seed=0, eye=(0.0, 0.0, 10.0), width=1280, height=720),
    "slab-8m": dict(n=8388608, kind=1, seed=0, eye=(0.0, 0.0, 10.0), width=1920, height=1080),
}


def camera(eye, width, height, target=(0.0, 0.0, 0.0)):
    """(view, projection) as Camera.update makes them (src/scene/camera.js:258-278)."""
    return look_at(eye, target, (0.0, 1.0, 0.0)), perspective(FOVY, width / height, NEAR)


def plain(name, cull=WD_SPLAT_CULL, n=None):
    """(records, wd_view, spec) of a single-view workload; n overrides the size (same generator)."""
    spec = dict(PLAIN[name])
    if n is not None:
        spec["n"] = n
    rec = synth_records(spec["kind"], spec["seed"], spec["n"])
    return rec, view_for(rec, spec["eye"], spec["width"], spec["height"], cull), spec


def bounds_of(records):
    pos = np.ascontiguousarray(records).view(np.float32).reshape(-1, 8)[:, :3]
    return pos.min(axis=0), pos.max(axis=0)


def view_for(records, eye, width, height, cull=WD_SPLAT_CULL, target=(0.0, 0.0, 0.0)):
    mn, mx = bounds_of(records)
    v, p = camera(eye, width, height, target)
    return prepare_view(v, p, float(width), float(height), mn, mx, cull)


def orbit_pose(stream, rmin=5.0, rmax=20.0):
    """A random eye on a sphere of radius U[rmin, rmax] about the origin."""
    r = rmin + (rmax - rmin) * stream.uniform()
    z = -1.0 + 2.0 * stream.uniform()
    phi = 2.0 * math.pi * stream.uniform()
    s = math.sqrt(max(0.0, 1.0 - z * z))
    return (r * s * math.cos(phi), r * z, r * s * math.sin(phi))


def batch_sizes(seed=None, count=64, lo=1e4, hi=2e6, target=32_000_000):
    """Scene sizes log-uniform in [lo, hi]. With no seed, the seed in 0..999 whose total is nearest target."""
    def sizes(s):
        st = Stream(s)
        return [int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * st.uniform()))) for _ in range(count)]
    if seed is None:
        seed = min(range(1000), key=lambda s: abs(sum(sizes(s)) - target))
    return seed, sizes(seed)


def orbit_sequence(count=1024, radius=10.0):
    """The learned workload's camera path: a full circle about y at radius 10, height swinging +-3."""
    eyes = []
    for k in range(count):
        a = 2.0 * math.pi * k / count
        eyes.append((radius * math.sin(a), 3.0 * math.sin(3.0 * a), radius * math.cos(a)))
    return eyes


LEARNED = dict(n=4194304, kind=0, seed=0, width=1280, height=720, train_poses=4096, train_seed=4242)


def view_at(cloud, eye, width, height, cull=WD_SPLAT_CULL, target=(0.0, 0.0, 0.0)):
    """The per-view uniform for a cloud already on the device (its box from the upload check)."""
    v, p = camera(eye, width, height, target)
    return prepare_view(v, p, float(width), float(height), cloud.min, cloud.max, cull)


def training_poses(count, seed, radius=(8.0, 12.0), height=3.0):
    """Random eyes of the orbit family: any angle about y, radius and height drawn uniformly."""
    st = Stream(seed)
    eyes = []
    for _ in range(count):
        a = 2.0 * math.pi * st.uniform()
        r = radius[0] + (radius[1] - radius[0]) * st.uniform()
        y = height * (2.0 * st.uniform() - 1.0)
        eyes.append((r * math.sin(a), y, r * math.cos(a)))
    return eyes


def training_set(ctx, cloud, state, eyes, width, height, cull=WD_SPLAT_CULL):
    """(poses, targets) for the predictor: each eye's 12 pose floats, and the 256 equal-mass bin widths of
    the keys its exact sort produces (the sort generates its own training data)."""
    from . import pose_features, widths_from_edges
    X, Y = [], []
    for eye in eyes:
        view = view_at(cloud, eye, width, height, cull)
        if view is None:
            continue
        state.sort(cloud, view)
        X.append(pose_features(view))
        Y.append(widths_from_edges(state.quantiles()))
    return np.array(X, np.float32), np.array(Y, np.float32)


SH_COEFFICIENTS = (0, 3, 8, 15)


def sh_words(n, degree, seed=7, sigma=0.2):
    """Harmonics of `degree` for n splats in the reference's `sh` layout (src/scene/splats.js:306-332):
    coefficients N(0, sigma) as IEEE halves, coefficient by coefficient, red, green and blue, ceil(3k / 2)
    32-bit words a splat (k = 3, 8, 15). numpy's seeded generator."""
    import numpy as np
    k = SH_COEFFICIENTS[degree]
    w = (3 * k + 1) // 2
    halves = np.zeros((n, 2 * w), np.uint16)
    vals = np.random.default_rng(seed).normal(0.0, sigma, (n, 3 * k)).astype(np.float16)
    halves[:, :3 * k] = vals.view(np.uint16)
    return halves.reshape(-1).view(np.uint32)


PARTICLE_EYE = (0.0, 0.0, 100.0)
PARTICLE_FORWARD = (0.0, 0.0, -1.0)


def particle_pool(n, seed=11, dead=0.25, unused=0.05):
    """A pool of n 32-byte particles (position, age, velocity, lifetime; src/render/particles.js:51-56) seen
    from PARTICLE_EYE along PARTICLE_FORWARD: depths log-uniform over the buckets' whole range and past both
    ends (2^-12 to 2^21, and some behind the eye), a fraction dead (age past lifetime) and a fraction never
    used (zeroed, lifetime 0). numpy's seeded generator."""
    import numpy as np
    rng = np.random.default_rng(seed)
    depth = np.exp2(rng.uniform(-12.0, 21.0, n))
    depth[rng.random(n) < 0.02] *= -1.0
    pos = np.zeros((n, 3), np.float64)
    pos[:, :2] = rng.normal(0.0, 1.0, (n, 2)) * (0.2 * depth[:, None])
    pos[:, 2] = PARTICLE_EYE[2] - depth
    rec = np.zeros((n, 8), np.float32)
    rec[:, :3] = pos
    rec[:, 4:7] = rng.normal(0.0, 1.0, (n, 3))
    life = rng.uniform(1.0, 5.0, n)
    age = rng.uniform(0.0, 1.0, n) * life
    gone = rng.random(n) < dead
    age[gone] = life[gone] + rng.uniform(0.0, 2.0, gone.sum())
    rec[:, 3] = age
    rec[:, 7] = life
    rec[rng.random(n) < unused] = 0.0
    return rec.view(np.uint8).reshape(n, 32)
