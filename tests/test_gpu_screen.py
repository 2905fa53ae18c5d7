"""The learned pair screen on the GPU: its BF16 matrix-core forward pass and gradient against the FP64 restatement
(oracle/screen_ref.py), training on the GPU, and screened force computations against plain ones.

Tolerances: logits 1e-4 + 1e-3 |z| against the restatement that rounds the same operands to bfloat16 (the rest is
FP32 accumulation order); gradients 1e-2 of each tensor's largest entry against that restatement (a bfloat16
rounding of an FP32 intermediate may fall on the other side of a tie) and 0.2 in relative 2-norm against exact FP64
(bfloat16 rounding, amplified where a gradient entry is a sum of few terms that cancel: up to 0.105 over 80 random
draws of the restatement alone); loss 1e-5 relative. Screened forces equal plain forces to 1e-6 of the largest force
(the sum order of the kept pairs may differ, and two windows that reach a target at the same FP64 distance may leave
either's FP32 start tangent), energies to 1e-12, and no dropped pair lies within the cutoff.
"""
import numpy as np
import pytest

import curvedspacesim_amd as csa
from curvedspacesim_amd import screen, workloads
from oracle import screen_ref as sr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


def _blocks(g):
    o = 0
    out = []
    for n in (16 * 64, 64, 64 * 64, 64, 64, 1):
        out.append(g[o:o + n])
        o += n
    return out


def test_forward_matches_restatement(ctx, rng):
    print("case screen forward", flush=True)
    p = sr.init_params(7)
    p[sr.N_PARAMS - 1] = 0.25
    X = rng.normal(size=(1000, 16)).astype(np.float32)
    z = ctx.screen_forward(p, X)
    ref = sr.forward(p, X, bf16_ops=True)
    np.testing.assert_allclose(z, ref, atol=1e-4, rtol=1e-3)
    exact = sr.forward(p, X)
    assert np.max(np.abs(z - exact)) < 5e-2 * np.max(np.abs(exact))


@pytest.mark.parametrize("n", [64, 1000, 4099])
def test_gradient_matches_restatement(ctx, rng, n):
    print(f"case screen gradient n={n}", flush=True)
    p = sr.init_params(11)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.float32)
    g, loss = ctx.screen_gradient(p, X, y)
    lref, gref = sr.forward_backward(p, X, y, bf16_ops=True)
    lex, gex = sr.forward_backward(p, X, y)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    for a, b, c in zip(_blocks(g), _blocks(gref), _blocks(gex)):
        scale = np.max(np.abs(b)) + 1e-12
        assert np.max(np.abs(a - b)) <= 1e-2 * scale
        assert np.linalg.norm(a - c) <= 0.2 * (np.linalg.norm(c) + 1e-12)


def test_training_reduces_loss(ctx, rng):
    print("case screen training", flush=True)
    X = rng.normal(size=(20000, 16)).astype(np.float32)
    y = ((X[:, 0] + 0.5 * X[:, 1] ** 2 - 0.3 * X[:, 2] * X[:, 3]) > 0.3).astype(np.float32)
    p, losses = ctx.screen_train(screen.init_params(3), X, y, steps=400, batch=2048, lr=3e-3, seed=5)
    assert losses[-20:].mean() < 0.4 * losses[:5].mean()
    acc = np.mean((ctx.screen_forward(p, X) > 0) == (y > 0.5))
    assert acc > 0.95


@pytest.fixture(scope="module")
def full_screen(ctx):
    """The screen of the benchmark: trained on MD snapshots of two other N = 20,000 bumpy spheres (seeds 13378, 13379),
    calibrated on snapshots of two more (13380, 13381)."""
    print("case build the N = 20,000 screen", flush=True)
    return screen.build(ctx, config=3)


def test_screened_forces_match_plain_full_size_md(ctx, full_screen):
    """The benchmark workload (N = 20,000, seed 13377) after 25, 50 and 100 Nose-Hoover steps from its placement: no
    pair within the cutoff is dropped and the screened forces equal the plain ones."""
    w = workloads.config(3)
    sysm = workloads.make_system(ctx, w)
    done = 0
    for n in (25, 50, 100):
        workloads.step(sysm, w, n - done)
        done = n
        print(f"case screened vs plain, N = 20,000 after {n} NVT steps", flush=True)
        sysm.record_pairs(True)
        F0, E0 = sysm.compute_forces()
        d = sysm.pair_distances()
        X = sysm.screen_features()
        sysm.record_pairs(False)
        drop = screen.dropped(ctx, full_screen["params"], full_screen["threshold"], X)
        assert not np.any(drop & (d < w.range)), f"{int(np.sum(drop & (d < w.range)))} within-cutoff pairs would be dropped"
        sysm.set_screen(full_screen["params"], full_screen["threshold"])
        F1, E1 = sysm.compute_forces()
        st = sysm.screen_stats()
        sysm.set_screen(None)
        assert st["candidates"] == len(d) and st["kept"] == int(np.sum(~drop))
        assert np.max(np.abs(F1 - F0)) <= 1e-6 * np.abs(F0).max()
        assert abs(E1 - E0) <= 1e-12 * abs(E0)
        print(f"screen after {n} steps: {int(drop.sum())} of {len(d)} pairs dropped, beyond cutoff {np.mean(d >= w.range):.4%}",
              flush=True)


def test_screened_forces_match_plain(ctx):
    """A screen trained on other bumpy spheres (seeds 13378, 13379), calibrated on two more (13380, 13381), screening the
    N = 4,000 bumpy sphere of seed 13377: the forces equal the plain ones, no within-cutoff pair is dropped, and the
    pairs it does drop are reported (what the trained network saves over solving every candidate)."""
    print("case screened vs plain forces", flush=True)
    built = screen.build(ctx, config=3, n=4000, steps=600, batch=4096)
    w = workloads.config(3, n=4000)
    X, y, sysm = screen.collect(ctx, w)
    F0, E0 = sysm.compute_forces()
    drop = screen.dropped(ctx, built["params"], built["threshold"], X)
    assert not np.any(drop & (y > 0.5)), "a within-cutoff pair would be dropped"
    sysm.set_screen(built["params"], built["threshold"])
    F1, E1 = sysm.compute_forces()
    st = sysm.screen_stats()
    assert st["candidates"] == len(y) and st["kept"] == int(np.sum(~drop))
    scale = np.abs(F0).max()
    assert np.max(np.abs(F1 - F0)) <= 1e-6 * scale
    assert abs(E1 - E0) <= 1e-12 * abs(E0)
    print(f"screen: {built['train_pairs']} training pairs, threshold {built['threshold']:.3f}, "
          f"{np.sum(drop)} of {len(y)} pairs dropped ({np.mean(drop):.4%}); beyond cutoff {np.mean(y < 0.5):.4%}",
          flush=True)
