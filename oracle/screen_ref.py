"""ORACLE — test infrastructure only.

FP64 NumPy restatement of the learned pair screen (an MLP 16 -> 64 -> 64 -> 1 with ReLU, scored by the mean binary
cross-entropy with logits). It has no reference counterpart; it pins the device's BF16 matrix-core kernels:
  * forward_backward(..., bf16=False) is the exact FP64 function and gradient, itself checked against central finite
    differences in the CPU tests;
  * forward_backward(..., bf16=True) rounds to bfloat16 exactly the operands the device feeds to its matrix cores
    (inputs, W1, W2, the hidden activation H1, and the back-propagated dH2 and dH1), so the device must agree with it
    to FP32 accumulation error.
Parameter order: W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3.
"""
import numpy as np

N_IN, N_HID = 16, 64
N_PARAMS = N_IN * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID + 1


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + N_IN * N_HID].reshape(N_IN, N_HID); o += N_IN * N_HID
    b1 = p[o:o + N_HID]; o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = p[o:o + N_HID]; o += N_HID
    W3 = p[o:o + N_HID]; o += N_HID
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def init_params(seed):
    """He-normal weights, zero biases (float32)."""
    r = np.random.default_rng(seed)
    W1 = r.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_HID))
    W2 = r.normal(0, np.sqrt(2.0 / N_HID), (N_HID, N_HID))
    W3 = r.normal(0, np.sqrt(1.0 / N_HID), N_HID)
    return np.concatenate([W1.ravel(), np.zeros(N_HID), W2.ravel(), np.zeros(N_HID), W3, [0.0]]).astype(np.float32)


def forward(p, X, bf16_ops=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    H1 = np.maximum(r(X) @ r(W1) + b1, 0.0)
    H2 = np.maximum(r(H1) @ r(W2) + b2, 0.0)
    return H2 @ W3 + b3


def forward_backward(p, X, y, bf16_ops=False):
    """(loss, gradient) of the mean binary cross-entropy with logits; gradient in the parameter order."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    n = len(X)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Xr = r(X)
    H1 = np.maximum(Xr @ r(W1) + b1, 0.0)
    H1r = r(H1)
    H2 = np.maximum(H1r @ r(W2) + b2, 0.0)
    z = H2 @ W3 + b3
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    gW3 = H2.T @ dz
    gb3 = dz.sum()
    dH2 = np.outer(dz, W3) * (H2 > 0)
    gb2 = dH2.sum(0)
    dH2r = r(dH2)
    gW2 = H1r.T @ dH2r
    dH1 = (dH2r @ r(W2).T) * (H1 > 0)
    gb1 = dH1.sum(0)
    gW1 = Xr.T @ r(dH1)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
