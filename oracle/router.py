"""FP64 NumPy oracle of the learned probe router (test infrastructure only).

MLP d -> h (ReLU) -> c, mean softmax cross-entropy against targets uniform
over each row's labels.  `bf16=True` rounds every GEMM operand to bfloat16
(round to nearest even) exactly where the GPU kernels do, then computes in FP64.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to bfloat16 (RNE), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params, d, h, c):
    p = np.asarray(params, np.float64)
    W1 = p[:h * d].reshape(h, d)
    b1 = p[h * d:h * d + h]
    o = h * d + h
    W2 = p[o:o + c * h].reshape(c, h)
    b2 = p[o + c * h:]
    return W1, b1, W2, b2


def targets(labels, c):
    """labels: (n, L) 1-based list ids, 0 = none."""
    n = labels.shape[0]
    T = np.zeros((n, c))
    for i in range(n):
        ls = [int(v) - 1 for v in labels[i] if v > 0]
        for l in ls:
            T[i, l] += 1.0 / len(ls)
    return T


def prior_logits(X, C, gamma):
    """Centroid prior gamma * (2 x.c - |c|^2) (FP64); C is (c, d)."""
    X = np.asarray(X, np.float64)
    C = np.asarray(C, np.float64)
    return gamma * (2.0 * X @ C.T - (C * C).sum(1))


def loss_and_grad(params, X, labels, d, h, c, bf16=False, prior=None):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32)) if bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2 = split(params, d, h, c)
    Xf = np.asarray(X, np.float32)
    B = Xf.shape[0]
    H = np.maximum(r(Xf) @ r(f32(W1)).T + b1, 0.0)
    H32 = f32(H)
    Z = r(H32) @ r(f32(W2)).T + b2
    if prior is not None:
        Z = Z + prior
    Z32 = np.asarray(f32(Z), np.float64)
    T = targets(labels, c)
    zmax = Z32.max(1, keepdims=True)
    lse = zmax + np.log(np.exp(Z32 - zmax).sum(1, keepdims=True))
    loss = float(-(T * (Z32 - lse)).sum() / B)
    G = (np.exp(Z32 - lse) - T) / B
    G32 = f32(G)
    gW2 = r(G32).T @ r(H32)
    gb2 = G.sum(0)
    dH = r(G32) @ r(f32(W2))
    dH = dH * (H32 > 0)
    dH32 = f32(dH)
    gW1 = r(dH32).T @ r(Xf)
    gb1 = dH.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2])


def scores(params, X, d, h, c, bf16=True, prior=None):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2 = split(params, d, h, c)
    H = np.maximum(r(np.asarray(X, np.float32)) @ r(W1.astype(np.float32)).T + b1, 0.0)
    Z = r(H.astype(np.float32)) @ r(W2.astype(np.float32)).T + b2
    return Z if prior is None else Z + prior
