"""The depth-CDF predictor: the CPU oracle pinned by central differences, the GPU's forward and backward
kernels against the oracle, training, and save/load."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def random_case(rng, n=8, scale=1.0):
    p = np.zeros(mo.N_PARAMS)
    p[mo.O_W1:mo.O_B1] = rng.normal(0, np.sqrt(2 / 12), 128 * 12) * scale
    p[mo.O_B1:mo.O_W2] = rng.normal(0, 0.1, 128)
    p[mo.O_W2:mo.O_B2] = rng.normal(0, np.sqrt(2 / 128), 128 * 128) * scale
    p[mo.O_B2:mo.O_W3] = rng.normal(0, 0.1, 128)
    p[mo.O_W3:mo.O_B3] = rng.normal(0, np.sqrt(2 / 128), 256 * 128) * scale
    p[mo.O_B3:] = rng.normal(0, 0.1, 256)
    X = rng.normal(0, 1, (n, 12))
    T = rng.dirichlet(np.full(256, 0.3), n)
    T[0, :128] = 0  # zero-mass bins, as exact histograms have
    T[0] /= T[0].sum()
    return p, X, T


def test_bf16_rounding_nearest_even():
    # exact bf16 values stay; halfway cases go to even; others to nearest
    assert mo.bf16(np.float32(1.0 + 2 ** -9))[()] == 1.0
    assert mo.bf16(np.float32(1.0 + 2 ** -8 + 2 ** -12))[()] == 1.0 + 2 ** -7
    assert mo.bf16(np.float32(1.0 + 2 ** -7))[()] == 1.0 + 2 ** -7
    assert mo.bf16(np.float32(1.0 + 2 ** -8))[()] == 1.0          # halfway, to even (1.0)
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -8))[()] == 1.0 + 2 ** -6  # halfway, to even
    assert mo.bf16(np.float32(-2.5))[()] == -2.5


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    p, X, T = random_case(rng, n=4)
    loss, g, _ = mo.forward_backward(p, X, T, bf16_points=False)
    idx = np.concatenate([rng.integers(lo, hi, 12) for lo, hi in
                          [(mo.O_W1, mo.O_B1), (mo.O_B1, mo.O_W2), (mo.O_W2, mo.O_B2), (mo.O_B2, mo.O_W3),
                           (mo.O_W3, mo.O_B3), (mo.O_B3, mo.N_PARAMS)]])
    for i in idx:
        h = 1e-6
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (mo.loss_only(a, X, T) - mo.loss_only(b, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_oracle_softmax_head_gradient_is_p_minus_t():
    rng = np.random.default_rng(1)
    p, X, T = random_case(rng, n=3)
    _, g, P = mo.forward_backward(p, X, T, bf16_points=False)
    np.testing.assert_allclose(g[mo.O_B3:], (P - T).sum(axis=0) / 3, atol=1e-12)


def tensor_errors(g, w):
    out = {}
    for name, lo, hi in [("W1", mo.O_W1, mo.O_B1), ("b1", mo.O_B1, mo.O_W2), ("W2", mo.O_W2, mo.O_B2),
                         ("b2", mo.O_B2, mo.O_W3), ("W3", mo.O_W3, mo.O_B3), ("b3", mo.O_B3, mo.N_PARAMS)]:
        a, b = g[lo:hi].astype(np.float64), w[lo:hi]
        out[name] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))
    return out


@pytest.mark.gpu
def test_gpu_gradients_match_oracle(ctx):
    import winding_amd as wa
    rng = np.random.default_rng(2)
    p, X, T = random_case(rng, n=300)
    mlp = wa.Mlp(ctx, seed=3)
    try:
        mlp.set_params(p.astype(np.float32))
        p32 = mlp.params().astype(np.float64)
        loss, g = mlp.grad(X.astype(np.float32), T.astype(np.float32))
        w_loss, w_g, _ = mo.forward_backward(p32, X.astype(np.float32).astype(np.float64),
                                             T.astype(np.float32).astype(np.float64))
        errs = tensor_errors(g, w_g)
        print("loss gpu %.6f oracle %.6f; per-tensor max rel err %s" % (loss, w_loss, errs), flush=True)
        assert abs(loss - w_loss) <= 1e-3 * abs(w_loss)
        # per tensor, the largest difference relative to the tensor's largest entry: fp32 sums in another
        # order than the oracle's float64 ones, with the same bf16 rounding points
        assert max(errs.values()) <= 1e-4, errs
        cos = float(g.astype(np.float64) @ w_g / (np.linalg.norm(g) * np.linalg.norm(w_g)))
        assert cos > 0.9999
        # the same inputs give the same gradients, bit for bit (the loss is summed by atomics: to rounding)
        loss2, g2 = mlp.grad(X.astype(np.float32), T.astype(np.float32))
        assert abs(loss2 - loss) <= 1e-5 * abs(loss) and np.array_equal(g2, g)
    finally:
        mlp.destroy()


@pytest.mark.gpu
def test_gpu_inference_matches_oracle_forward(ctx):
    """The one-launch inference kernel (the learned sort's and predict's) against the oracle's forward with the
    same bf16 rounding points: equal to fp32 summation order (an activation whose fp32 sum lands the other side of
    a bf16 rounding edge moves by one bf16 step; 1% of a probability covers that)."""
    import winding_amd as wa
    rng = np.random.default_rng(8)
    p, X, T = random_case(rng, n=40)
    mlp = wa.Mlp(ctx, seed=3)
    try:
        mlp.set_params(p.astype(np.float32))
        p32 = mlp.params().astype(np.float64)
        P = mlp.predict(X.astype(np.float32)).astype(np.float64)
        _, _, Po = mo.forward_backward(p32, X.astype(np.float32).astype(np.float64), T)
        rel = np.abs(P - Po) / np.maximum(Po, 1e-12)
        print("inference vs oracle: max rel %.3g, median rel %.3g" % (rel.max(), np.median(rel)), flush=True)
        assert np.allclose(P.sum(axis=1), 1.0, atol=1e-5)
        assert np.median(rel) < 1e-5 and np.allclose(P, Po, rtol=1e-2, atol=1e-7)
    finally:
        mlp.destroy()


@pytest.mark.gpu
def test_gpu_training_reduces_loss_and_saves(ctx):
    import winding_amd as wa
    rng = np.random.default_rng(4)
    n = 1024
    X = rng.normal(0, 1, (n, 12)).astype(np.float32)
    # targets: a bump whose centre and width follow the pose, a learnable map
    centre = 128 + 60 * np.tanh(X[:, 0] + 0.5 * X[:, 3])
    width = 10 + 5 * (1 + np.tanh(X[:, 7]))
    k = np.arange(256)
    T = np.exp(-0.5 * ((k[None, :] - centre[:, None]) / width[:, None]) ** 2)
    T = (T / T.sum(axis=1, keepdims=True)).astype(np.float32)
    # cross-entropy against soft targets bottoms out at their entropy: compare what is above it (the KL)
    entropy = float(-(T * np.log(np.maximum(T, 1e-30))).sum(axis=1).mean())
    mlp = wa.Mlp(ctx, seed=5)
    try:
        before, _ = mlp.grad(X, T)
        mlp.train(X, T, epochs=100, batch=128, lr=2e-3, seed=6)
        final, _ = mlp.grad(X, T)
        print("loss before %.4f, after %.4f, target entropy %.4f" % (before, final, entropy), flush=True)
        assert final - entropy < 0.25 * (before - entropy)
        pred = mlp.predict(X[:64])
        assert np.allclose(pred.sum(axis=1), 1, atol=1e-4)
        blob = mlp.save()
        other = wa.Mlp(ctx, seed=99)
        other.load(blob)
        assert np.array_equal(other.predict(X[:64]), pred)
        other.destroy()
    finally:
        mlp.destroy()
