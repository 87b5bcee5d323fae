"""NumPy restatement of the learned-guess MLP forward/backward/Adam (TEST
INFRASTRUCTURE ONLY): the CPU checker for pychem_amd's FP32 HIP kernels.
Computed in float64."""
import numpy as np


def forward(Ws, bs, X):
    acts = [X]
    for i, (W, b) in enumerate(zip(Ws, bs)):
        Z = acts[-1] @ W + b
        acts.append(np.tanh(Z) if i < len(Ws) - 1 else Z)
    return acts


def mse(Y, T):
    d = Y - T
    return float((d * d).mean()), 2.0 * d / d.size


def backward(Ws, acts, dY):
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    dA = dY
    for i in reversed(range(len(Ws))):
        dZ = dA * (1.0 - acts[i + 1] ** 2) if i < len(Ws) - 1 else dA
        gW[i] = acts[i].T @ dZ
        gb[i] = dZ.sum(0)
        dA = dZ @ Ws[i].T
    return gW, gb


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
