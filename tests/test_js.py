"""The JavaScript side: the JS CPU path against the C oracle bit for bit, the JS API's own checks
(node tests/js/cpu_test.js, no GPU), and the JS API end to end on the GPU (node tests/js/gpu_test.js)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import winding_amd as wa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def node(*args, timeout=300):
    r = subprocess.run(["node"] + list(args), cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    return r.stdout


def js_oracle(recs, view):
    with tempfile.TemporaryDirectory() as d:
        rp, vp, out = os.path.join(d, "r.bin"), os.path.join(d, "v.bin"), os.path.join(d, "o")
        np.ascontiguousarray(recs).tofile(rp)
        open(vp, "wb").write(bytes(view))
        node("oracle/run_js.js", rp, vp, out)
        keys = np.fromfile(out + "_keys.bin", np.int32)
        order = np.fromfile(out + "_order.bin", np.uint32)
        cov = np.fromfile(out + "_cov.bin", np.float64).reshape(-1, 3)
    return keys, order, cov


@pytest.mark.parametrize("name,n,cull", [("plain-64k", 65536, 0.5), ("plain-64k", 20000, 0.0), ("slab-8m", 50000, 0.5)])
def test_js_cpu_path_equals_c_oracle(name, n, cull):
    recs, view, _ = wa.workloads.plain(name, cull=cull, n=n)
    keys, order, cov = js_oracle(recs, view)
    o_order, o_keys, o_keep, o_cov = oracle.run(recs, view)
    assert np.array_equal(keys >= 0, o_keep)
    assert np.array_equal(keys[o_keep].astype(np.uint16), o_keys[o_keep])
    assert np.array_equal(order, o_order)
    assert np.array_equal(cov[o_keep].view(np.uint64), o_cov[o_keep].view(np.uint64))


def test_js_api_cpu():
    out = node("tests/js/cpu_test.js")
    assert "checks passed" in out


def test_js_esm_entry_loads():
    out = node("--input-type=module", "-e",
               "import wd from './winding_amd/js/index.mjs'; import { SplatPass, SPLAT_CULL } from './winding_amd/js/index.mjs';"
               "console.log(typeof SplatPass, SPLAT_CULL, typeof wd.parseSplats)")
    assert "function 0.5 function" in out


@pytest.mark.gpu
def test_js_api_gpu():
    out = node("tests/js/gpu_test.js")
    assert "checks passed" in out
