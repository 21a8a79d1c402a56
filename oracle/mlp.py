"""FP64 reference of the probe network (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference (the reference probes a
fixed number of lists, pq_idx.rs:424), so its oracle is the mathematics:
  H = relu(X W1 + b1), Z = H W2 + b2, p = softmax(Z),
  loss = mean_rows( -sum_j T_j log p_j ),
and its gradients.  `round_like_gpu=True` rounds X, W1, W2, H, G and dH to
bf16 (round to nearest even) where the GPU kernels do, so that the GPU
gradients can be compared tightly; the plain FP64 form is pinned by
central finite differences in tests/test_probe_gpu.py.
"""
from __future__ import annotations

import numpy as np


def bf16(x) -> np.ndarray:
    """Round float32 values to bf16 (RNE), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(W1, b1, W2, b2, X, round_like_gpu=True):
    rnd = bf16 if round_like_gpu else (lambda a: np.asarray(a, np.float64))
    x, w1, w2 = rnd(X), rnd(W1), rnd(W2)
    h = np.maximum(x @ w1 + np.asarray(b1, np.float64), 0.0)
    hq = rnd(h.astype(np.float32)) if round_like_gpu else h
    z = hq @ w2 + np.asarray(b2, np.float64)
    return x, w1, w2, h, hq, z


def loss_grad(W1, b1, W2, b2, X, T, round_like_gpu=True):
    rnd = bf16 if round_like_gpu else (lambda a: np.asarray(a, np.float64))
    x, w1, w2, h, hq, z = forward(W1, b1, W2, b2, X, round_like_gpu)
    n = x.shape[0]
    T = np.asarray(T, np.float64)
    zmax = z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(z - zmax).sum(axis=1, keepdims=True)) + zmax
    p = np.exp(z - lse)
    loss = -(T * (z - lse)).sum(axis=1)
    g = (p - T) / n
    gq = rnd(g.astype(np.float32)) if round_like_gpu else g
    gW2 = hq.T @ gq
    gb2 = gq.sum(axis=0)
    dh = gq @ w2.T
    dhq = rnd(dh.astype(np.float32)) if round_like_gpu else dh
    dhq = dhq * (hq > 0)
    gW1 = x.T @ dhq
    gb1 = dhq.sum(axis=0)
    return {"loss": loss, "gW1": gW1, "gb1": gb1, "gW2": gW2, "gb2": gb2, "logits": z}


def mean_loss(W1, b1, W2, b2, X, T) -> float:
    return float(loss_grad(W1, b1, W2, b2, X, T, round_like_gpu=False)["loss"].mean())
