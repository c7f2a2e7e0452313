"""CPU oracle of the depth-CDF predictor (numpy). TEST INFRASTRUCTURE ONLY.

The same network and loss as winding_amd/csrc/mlp_host.cpp: 12 -> relu 128 -> relu 128 -> 256, softmax,
cross-entropy against target distributions, mean over rows. With bf16=True it rounds to bf16 (nearest
even) exactly where the GPU's matrix-core GEMMs read their operands -- weights, the pose inputs, the
hidden activations and the backward deltas -- and accumulates in float64, so the GPU's gradients must
agree to fp32 accumulation error. With bf16=False everything is float64: the analytic gradient is then
pinned by central differences (tests/test_mlp.py).
"""
import numpy as np

IN, H, OUT = 12, 128, 256
O_W1, O_B1 = 0, 128 * 12
O_W2, O_B2 = O_B1 + 128, O_B1 + 128 + 128 * 128
O_W3, O_B3 = O_B2 + 128, O_B2 + 128 + 256 * 128
N_PARAMS = O_B3 + 256


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(128, 12), p[O_B1:O_W2], p[O_W2:O_B2].reshape(128, 128), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(256, 128), p[O_B3:])


def forward_backward(params, X, T, bf16_points=True):
    """(loss, grads[N_PARAMS], probs) for poses X (n x 12) and targets T (n x 256)."""
    r = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16_points else (lambda a: a)
    W1, b1, W2, b2, W3, b3 = unpack(params)
    W1b, W2b, W3b = (r(W.astype(np.float32)) if bf16_points else W for W in (W1, W2, W3))
    n = len(X)
    X = np.asarray(X, np.float64)
    Xb = r(X.astype(np.float32)) if bf16_points else X
    H1 = f32(np.maximum(Xb @ W1b.T + b1, 0))
    H1b = r(H1.astype(np.float32)) if bf16_points else H1
    H2 = f32(np.maximum(H1b @ W2b.T + b2, 0))
    H2b = r(H2.astype(np.float32)) if bf16_points else H2
    Z = f32(H2b @ W3b.T + b3)
    Z = Z - Z.max(axis=1, keepdims=True)
    E = np.exp(Z)
    P = E / E.sum(axis=1, keepdims=True)
    T = np.asarray(T, np.float64)
    with np.errstate(divide="ignore"):
        logp = np.where(T > 0, np.log(np.maximum(P, 1e-30)), 0.0)
    loss = -(T * logp).sum() / n
    dZ = f32((P - T) / n)
    dZb = r(dZ.astype(np.float32)) if bf16_points else dZ
    gW3 = dZb.T @ H2b
    gb3 = dZ.sum(axis=0)
    dH2 = f32((dZb @ W3b) * (H2 > 0))
    dH2b = r(dH2.astype(np.float32)) if bf16_points else dH2
    gW2 = dH2b.T @ H1b
    gb2 = dH2.sum(axis=0)
    dH1 = f32((dH2b @ W2b) * (H1 > 0))
    dH1b = r(dH1.astype(np.float32)) if bf16_points else dH1
    gW1 = dH1b.T @ Xb
    gb1 = dH1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g, P


def loss_only(params, X, T):
    return forward_backward(params, X, T, bf16_points=False)[0]
