"""CPU oracle for the parameter-proposal MLP (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module; the product (lenia_amd/) never does.

float64 restatement of the network of include/lenia_mlp.h: 81 -> 256 -> 256 -> 256 -> 3,
GELU (tanh form) on the hidden layers, loss = BCE-with-logits(out0, survival) +
(out1 - concentration)^2 + (out2 - drift)^2, mean over the batch, and its gradient by
backpropagation.  The reference has no learned component; this restates the network's
own definition.  With ``bf16=True`` every matrix operand is rounded to bfloat16 (round to
nearest even) where the GPU rounds it -- inputs, weights, hidden activations, output and
pre-activation gradients -- so the GPU's fp32 accumulation is the only remaining
difference.  Pinned by central finite differences (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

IN, INP, HID, OUT, OUTP = 81, 128, 256, 3, 64
SHAPES = [("W1", (INP, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID, OUTP)), ("b4", (OUTP,))]
C_GELU = np.sqrt(2.0 / np.pi)


def param_count() -> int:
    return int(sum(np.prod(s) for _, s in SHAPES))


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def to_bf16(a) -> np.ndarray:
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) << np.uint64(16)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def forward(p: np.ndarray, x: np.ndarray, bf16=False):
    """Outputs [n, 3] and the saved activations."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    X = np.zeros((x.shape[0], INP))
    X[:, :IN] = x
    X = r(X)
    acts, zs = [X], []
    h = X
    for l in (1, 2, 3):
        z = h @ r(w[f"W{l}"]) + w[f"b{l}"]
        zs.append(z)
        h = r(gelu(z))
        acts.append(h)
    o = h @ r(w["W4"]) + w["b4"]
    return o[:, :OUT], (acts, zs, o)


def loss_and_grad(p: np.ndarray, x: np.ndarray, y: np.ndarray, bf16=False):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    _, (acts, zs, o) = forward(p, x, bf16)
    n = x.shape[0]
    xo, s = o[:, 0], y[:, 0]
    l = np.maximum(xo, 0) - xo * s + np.log1p(np.exp(-np.abs(xo)))
    d1, d2 = o[:, 1] - y[:, 1], o[:, 2] - y[:, 2]
    loss = float(np.mean(l + d1 * d1 + d2 * d2))
    dO = np.zeros_like(o)
    dO[:, 0] = (1.0 / (1.0 + np.exp(-xo)) - s) / n
    dO[:, 1] = 2.0 * d1 / n
    dO[:, 2] = 2.0 * d2 / n
    g = {"W4": acts[3].T @ r(dO), "b4": dO.sum(axis=0)}
    dz16 = r(dO)
    W_next = w["W4"]
    for l in (3, 2, 1):
        dh = dz16 @ r(W_next).T
        dz = dh * gelu_grad(zs[l - 1])
        dz16 = r(dz)
        g[f"W{l}"] = acts[l - 1].T @ dz16
        g[f"b{l}"] = dz.sum(axis=0)
        W_next = w[f"W{l}"]
    return loss, pack(g)
