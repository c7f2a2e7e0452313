"""The C ABI without a GPU: the library loads, exports what include/*.h declares, and its host-only
entry points (seeded records, the per-view uniform) agree with the oracle bit for bit."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import winding_amd as wa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    declared = header_functions()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", wa.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r"\bT (wd_\w+)", out))
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    assert sorted(wa.EXPORTS) == declared
    L = wa.lib()
    for n in declared:
        assert getattr(L, n) is not None


def test_addon_is_built_and_links_the_library():
    addon = os.path.join(ROOT, "winding_amd", "winding_amd.node")
    assert os.path.exists(addon)
    out = subprocess.run(["readelf", "-d", addon], stdout=subprocess.PIPE, text=True).stdout
    assert "libwinding_amd.so" in out and "$ORIGIN" in out


def test_synth_is_deterministic_and_seeded():
    a = wa.synth_records(0, 7, 5000)
    b = wa.synth_records(0, 7, 5000)
    c = wa.synth_records(0, 8, 5000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # A prefix of a larger cloud is the smaller cloud: each splat has its own stream.
    assert np.array_equal(wa.synth_records(0, 7, 70000)[:5000], a)
    pos = a.view(np.float32).reshape(-1, 8)
    assert np.all(np.isfinite(pos[:, :6])) and np.all(pos[:, 3:6] > 0)
    s = wa.synth_records(1, 0, 100000).view(np.float32).reshape(-1, 8)
    slab = np.abs(s[:, 2]) <= 0.25
    assert 0.79 < slab.mean() < 0.82  # 80% in the slab, plus the box's share of |z| <= 0.25


def test_prepare_matches_oracle_bit_for_bit():
    rng = np.random.default_rng(3)
    for _ in range(200):
        eye = rng.uniform(-20, 20, 3)
        world = np.eye(4)
        world[:3, 3] = rng.uniform(-2, 2, 3)
        world[:3, :3] *= rng.uniform(0.5, 2)
        world = world.T.reshape(-1).astype(np.float32)
        v = oracle.look_at(eye)
        p = oracle.perspective(rng.uniform(0.3, 2.0), rng.uniform(0.5, 2.5), 0.1)
        mn = rng.uniform(-5, 0, 3).astype(np.float32)
        mx = (mn + rng.uniform(0, 5, 3)).astype(np.float32)
        cull = float(rng.choice([0.0, 0.5, 1.0]))
        got = wa.prepare_view(v, p, 1280.0, 720.0, mn, mx, cull, world=world)
        want = oracle.prepare(v, p, 1280.0, 720.0, mn, mx, cull, world=world)
        assert (got is None) == (want is None)
        if got is not None:
            assert bytes(got) == bytes(want)


def test_prepare_refuses_negative_cull():
    # renderer.js:1528-1529: "splatCull must be 0 or more, got -1"
    with pytest.raises(wa.WindingError, match="splatCull must be 0 or more"):
        wa.prepare_view(wa.look_at((0, 0, 5)), wa.perspective(0.8, 4 / 3, 0.1), 160.0, 120.0, [0, 0, 0], [1, 1, 1], -1.0)


def test_workload_sizes():
    seed, sizes = wa.workloads.batch_sizes()
    assert len(sizes) == 64 and min(sizes) >= 10000 and max(sizes) <= 2000000
    assert abs(sum(sizes) - 32_000_000) < 500_000
    eyes = wa.workloads.orbit_sequence()
    assert len(eyes) == 1024 and all(abs(math.hypot(e[0], e[2]) - 10) < 1e-9 for e in eyes)


def test_open_without_a_gpu_fails_cleanly():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(wa.WindingError):
        wa.Context(0)
