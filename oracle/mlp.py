"""CPU oracle of the learned probe MLP -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the network, loss and gradients that the device kernels compute
(d -> hidden (ReLU) -> nlist, soft cross-entropy against the clusters of the exact top-L
neighbours), with bf16 rounding at the same points as the device: the inputs and weights of every
product, the hidden activations, the logit gradient and the masked hidden gradient. Products
accumulate in fp64 here (the device: fp32), so the two agree to fp32 rounding of the sums.

There is no reference counterpart (the reference has no learned probe); the gradients here are
pinned by central finite differences of `loss` in fp64 (tests/test_learned_probe.py).
"""

import numpy as np


def bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params, d, hid, ncls):
    n1, n2 = hid * d, ncls * hid
    W1 = params[:n1].reshape(hid, d)
    b1 = params[n1:n1 + hid]
    W2 = params[n1 + hid:n1 + hid + n2].reshape(ncls, hid)
    b2 = params[n1 + hid + n2:]
    return W1, b1, W2, b2


def targets(labels, ncls):
    B, L = labels.shape
    T = np.zeros((B, ncls), np.float64)
    for j in range(L):
        ok = (labels[:, j] >= 0) & (labels[:, j] < ncls)
        np.add.at(T, (np.nonzero(ok)[0], labels[ok, j]), 1.0 / L)
    return T


def forward(params, x, d, hid, ncls):
    W1, b1, W2, b2 = unpack(params, d, hid, ncls)
    Xb = bf16(x).astype(np.float64)
    Hpre = Xb @ bf16(W1).astype(np.float64).T + b1
    H = bf16(np.maximum(Hpre, 0).astype(np.float32)).astype(np.float64)
    logits = H @ bf16(W2).astype(np.float64).T + b2
    return Xb, H, logits


def loss_and_grads(params, x, labels, d, hid, ncls):
    """(batch-mean loss, gradient vector in the parameter layout) with the device's bf16 points."""
    W1, b1, W2, b2 = unpack(params, d, hid, ncls)
    B = x.shape[0]
    Xb, H, logits = forward(params, x, d, hid, ncls)
    T = targets(labels, ncls)
    m = logits.max(1, keepdims=True)
    lse = m + np.log(np.exp(logits - m).sum(1, keepdims=True))
    P = np.exp(logits - lse)
    loss = float((T * (lse - logits)).sum() / B)
    dL = bf16(((P - T) / B).astype(np.float32)).astype(np.float64)
    dH = bf16(((dL @ bf16(W2).astype(np.float64)) * (H > 0)).astype(np.float32)).astype(np.float64)
    gW2 = dL.T @ H
    gW1 = dH.T @ Xb
    g = np.concatenate([gW1.ravel(), dH.sum(0), gW2.ravel(), dL.sum(0)])
    return loss, g


def loss_exact(params, x, labels, d, hid, ncls):
    """The same loss without any bf16 rounding, in fp64 (for finite differences)."""
    W1, b1, W2, b2 = unpack(params.astype(np.float64), d, hid, ncls)
    H = np.maximum(x.astype(np.float64) @ W1.T + b1, 0)
    logits = H @ W2.T + b2
    T = targets(labels, ncls)
    m = logits.max(1, keepdims=True)
    lse = m + np.log(np.exp(logits - m).sum(1, keepdims=True))
    return float((T * (lse - logits)).sum() / x.shape[0])


def grads_exact(params, x, labels, d, hid, ncls):
    """Analytic fp64 gradient of loss_exact."""
    W1, b1, W2, b2 = unpack(params.astype(np.float64), d, hid, ncls)
    B = x.shape[0]
    X = x.astype(np.float64)
    H = np.maximum(X @ W1.T + b1, 0)
    logits = H @ W2.T + b2
    T = targets(labels, ncls)
    m = logits.max(1, keepdims=True)
    P = np.exp(logits - m)
    P /= P.sum(1, keepdims=True)
    dL = (P - T) / B
    dH = (dL @ W2) * (H > 0)
    return np.concatenate([(dH.T @ X).ravel(), dH.sum(0), (dL.T @ H).ravel(), dL.sum(0)])
