"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the C and
the JavaScript oracles, which had to agree): the oracle still reproduces them, and they satisfy what the
reference's sort fixes whatever the arithmetic -- the order is the survivors sorted by key, ties by index
(a brute-force stable sort of the fixture's own keys), and the survivors are exactly the keyed splats."""
import glob
import os

import numpy as np
import pytest

import oracle
import winding_amd as wa

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def view_of(f):
    return wa.View.from_buffer_copy(f["view"].astype(np.float32).tobytes())


def test_fixtures_exist():
    names = {os.path.basename(p) for p in FIXTURES}
    assert {"mixture_64.npz", "mixture_4096.npz", "edges_4096.npz", "slab_4096.npz"} <= names


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    f = load(path)
    order, keys, keep, cov = oracle.run(f["records"], view_of(f))
    assert np.array_equal(np.where(keep, keys.astype(np.int32), -1), f["keys"])
    assert np.array_equal(order, f["order"])
    assert np.array_equal(cov[keep], f["cov"][keep])


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_fixture_order_is_the_stable_sort_of_its_keys(path):
    f = load(path)
    k = f["keys"]
    survivors = [i for i in range(len(k)) if k[i] >= 0]
    brute = sorted(survivors, key=lambda i: (int(k[i]), i))  # Python's sort is stable; the pair makes it total
    assert f["order"].tolist() == brute
    assert np.all(np.isfinite(f["cov"][k >= 0])) and np.all(np.isnan(f["cov"][k < 0]))
    # xx and yy of a footprint are squared lengths: never negative
    assert np.all(f["cov"][k >= 0][:, [0, 2]] >= 0)
