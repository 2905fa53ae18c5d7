"""The FP64 restatement of the learned screen (oracle/screen_ref.py) against the mathematics it must satisfy: its
gradient equals central finite differences of its loss, its bfloat16 rounding is round-to-nearest-even, and the loss
of a known case has its closed form. A plausible slip (a transposed weight, a missing ReLU mask, a wrong sign of the
logistic term) fails the finite-difference check."""
import numpy as np

from oracle import screen_ref as sr


def test_gradient_matches_finite_differences(rng):
    p = sr.init_params(3).astype(np.float64)
    p[sr.N_PARAMS - 1] = 0.3
    X = rng.normal(size=(24, sr.N_IN))
    y = (rng.uniform(size=24) < 0.5).astype(float)
    loss, g = sr.forward_backward(p, X, y)
    idx = np.concatenate([rng.choice(sr.N_PARAMS, 60, replace=False), [sr.N_PARAMS - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (sr.forward_backward(p + e, X, y)[0] - sr.forward_backward(p - e, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_closed_form():
    p = np.zeros(sr.N_PARAMS)
    p[-1] = 0.7  # every logit is b3
    X = np.ones((5, sr.N_IN))
    y = np.array([1, 0, 1, 1, 0], float)
    loss, g = sr.forward_backward(p, X, y)
    z = 0.7
    expect = np.mean(np.where(y == 1, np.log1p(np.exp(-z)), np.log1p(np.exp(z))))
    assert abs(loss - expect) < 1e-15
    assert abs(g[-1] - np.mean(1 / (1 + np.exp(-z)) - y)) < 1e-15


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # spacing of bfloat16 at 1
    vals = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + ulp / 2 + 2 ** -20, -3.14159, 0.0], np.float32)
    out = sr.bf16(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                 # tie, to even
    assert out[2] == 1 + 2 * ulp         # tie, to even (odd neighbour 1 + ulp)
    assert out[3] == 1 + ulp             # above the tie
    assert out[4] == -3.140625
    assert out[5] == 0.0
    assert one == 1
