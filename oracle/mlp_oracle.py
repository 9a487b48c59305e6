"""FP64 restatement of the warm-start network (test infrastructure only).

Only the tests and bench.py's CPU-baseline leg use this module, as the oracle. The product path
runs the HIP kernels of isar_imaging_cuda_amd/csrc/kernels/nn.hip through include/isar_learned.h.

The network is a stack of dense layers with ReLU after every layer but the last, with parameters
packed per layer as W (in x out, row-major) then b (out). The reference has no such network; it
seeds the fast-entropy autofocus of phase_adjustment.cu:282-324. With `bf16=True` every matrix
product rounds both operands to bfloat16 (round to nearest even) first, as the matrix-core
kernels do, and accumulates in FP64. The device accumulates in FP32, so the two agree to FP32
accumulation error. With `bf16=False` this is the exact FP64 network. Its gradients are pinned
against central finite differences in tests/test_learned.py.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even) and return them as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, widths: list[int]) -> list[tuple[np.ndarray, np.ndarray]]:
    out, off = [], 0
    for i, o in zip(widths[:-1], widths[1:]):
        W = params[off:off + i * o].reshape(i, o)
        off += i * o
        b = params[off:off + o]
        off += o
        out.append((W.astype(np.float64), b.astype(np.float64)))
    assert off == params.size
    return out


def _mm(a: np.ndarray, b: np.ndarray, bf16: bool) -> np.ndarray:
    if bf16:
        return bf16_round(a) @ bf16_round(b)
    return a.astype(np.float64) @ b.astype(np.float64)


def forward(params: np.ndarray, widths: list[int], x: np.ndarray, bf16: bool = True):
    """Returns the output and the list of layer outputs (activations) for backward."""
    layers = unpack(params, widths)
    acts = [np.asarray(x, dtype=np.float64)]
    a = acts[0]
    for li, (W, b) in enumerate(layers):
        z = _mm(a, W, bf16) + b
        a = np.maximum(z, 0.0) if li + 1 < len(layers) else z
        acts.append(a)
    return a, acts


def backward(params: np.ndarray, widths: list[int], acts: list[np.ndarray], grad_out: np.ndarray,
             bf16: bool = True) -> np.ndarray:
    """Parameter gradients (packed like the parameters) for the loss gradient `grad_out`."""
    layers = unpack(params, widths)
    grads = []
    dz = np.asarray(grad_out, dtype=np.float64)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        a_prev = acts[li]
        gW = _mm(a_prev.T, dz, bf16)
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            dz = _mm(dz, W.T, bf16) * (a_prev > 0)
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb.ravel()]
    return np.concatenate(flat)


def mse(pred: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
    d = np.asarray(pred, np.float64) - np.asarray(target, np.float64)
    return float(np.mean(d * d)), 2.0 * d / d.size


def adam(params, grads, m1, m2, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step; returns new (params, m1, m2)."""
    m1 = beta1 * m1 + (1 - beta1) * grads
    m2 = beta2 * m2 + (1 - beta2) * grads * grads
    p = params - lr * (m1 / (1 - beta1 ** step)) / (np.sqrt(m2 / (1 - beta2 ** step)) + eps)
    return p, m1, m2


def init_params(widths: list[int], seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    flat = []
    for i, o in zip(widths[:-1], widths[1:]):
        flat.append((rng.standard_normal((i, o)) * np.sqrt(2.0 / i)).ravel())
        flat.append(np.zeros(o))
    return np.concatenate(flat).astype(np.float32)
