"""Learned pair screen: training data from the exact solver, training on the GPU, and margin calibration.

The screen (an MLP 16 -> 64 -> 64 -> 1 run in BF16 on the matrix cores, see include/csg.h) scores how likely a
candidate pair of the cell list lies within the potential's range. Pairs scoring below a threshold are not solved;
every other pair is solved exactly, with patches still built from every candidate, so the kept pairs get exactly the
distances of the plain computation. The threshold is the lowest score of any within-range pair of held-out systems,
lowered by a margin, so that no within-range pair of those systems would have been dropped.
"""
import numpy as np

from . import _native as nat
from . import workloads

N_IN, N_HID = nat.SCREEN_FEATURES, 64


def init_params(seed):
    """He-normal weights and zero biases, in the parameter order of the C-ABI."""
    r = np.random.default_rng(seed)
    W1 = r.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_HID))
    W2 = r.normal(0, np.sqrt(2.0 / N_HID), (N_HID, N_HID))
    W3 = r.normal(0, np.sqrt(1.0 / N_HID), N_HID)
    p = np.concatenate([W1.ravel(), np.zeros(N_HID), W2.ravel(), np.zeros(N_HID), W3, [0.0]]).astype(np.float32)
    assert p.size == nat.SCREEN_PARAMS
    return p


def collect(ctx, w, md_steps=(0,)):
    """Features and labels (1: the exact geodesic distance is below the range) of every candidate pair of workload w,
    taken after each of the given numbers of its own MD steps from the placement (pairs of all snapshots stacked),
    plus the system. The screen meets states that the dynamics evolved, whose pair statistics differ from those of
    a fresh placement (repulsion thins out close pairs), so its data come from such states too."""
    sysm = workloads.make_system(ctx, w)
    Xs, ys = [], []
    done = 0
    for n in sorted(md_steps):
        if n > done:
            workloads.step(sysm, w, n - done)
            done = n
        sysm.record_pairs(True)
        sysm.compute_forces()
        d = sysm.pair_distances()
        Xs.append(sysm.screen_features())
        sysm.record_pairs(False)
        ys.append((d < pair_ranges(sysm, w)).astype(np.float32))
    return np.concatenate(Xs), np.concatenate(ys), sysm


def pair_ranges(sysm, w):
    """The cutoff of every candidate pair of the last force computation: the range, or the source's group range."""
    groups = w.meta.get("groups")
    if groups is None:
        return np.float64(w.range)
    off, _ = sysm.pairs()
    return np.repeat(np.asarray(w.meta["ranges"], float)[groups], np.diff(off))


def balance(X, y, fraction=0.25):
    """Repeats the rare beyond-range rows until they make up about `fraction` of the training set (on these
    workloads fewer than 1 in 400 candidate pairs lie beyond the range, too few for the loss to see them)."""
    neg = np.where(y < 0.5)[0]
    if len(neg) == 0:
        return X, y
    reps = max(1, int(fraction * (len(y) - len(neg)) / ((1 - fraction) * len(neg))))
    extra = np.repeat(neg, reps - 1)
    return np.concatenate([X, X[extra]]), np.concatenate([y, y[extra]])


def train(ctx, X, y, steps=2000, batch=8192, lr=2e-3, seed=1, init_seed=7):
    """Adam on the GPU from a He-normal start; returns (params, per-step minibatch losses)."""
    return ctx.screen_train(init_params(init_seed), X, y, steps, batch, lr, seed)


GUARD = 1.5  # pairs with 100 (1 - |d|/rc) above this are always kept (kScreenGuard of the library)


def dropped(ctx, params, threshold, X):
    """The pairs a screen drops: score below the threshold and separation at least 0.985 of the range."""
    return (ctx.screen_forward(params, X) < threshold) & (X[:, 0] <= GUARD)


def calibrate(ctx, params, X, y, margin=12.0):
    """Threshold below every within-range pair's logit of the held-out rows, by `margin` (logit units)."""
    z = ctx.screen_forward(params, X)
    pos = z[y > 0.5]
    return float(pos.min() - margin) if len(pos) else float(z.min() - margin)


TRAIN_MD_STEPS = (0, 40, 80)
HOLDOUT_MD_STEPS = (25, 50, 100, 150)


def build(ctx, config=3, n=None, train_seeds=(13378, 13379), holdout_seeds=(13380, 13381), steps=2000, batch=8192, lr=2e-3,
          margin=12.0, train_md_steps=TRAIN_MD_STEPS, holdout_md_steps=HOLDOUT_MD_STEPS):
    """Train and calibrate a screen on systems of the given configuration with other seeds than the one it will
    screen, from snapshots taken after the given numbers of MD steps. Returns dict(params, threshold, losses,
    train_pairs, holdout_pairs, holdout_drop, holdout_beyond)."""
    Xs, ys = [], []
    for s in train_seeds:
        X, y, _ = collect(ctx, workloads.config(config, seed=s, n=n), train_md_steps)
        Xs.append(X)
        ys.append(y)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    Xt, yt = balance(X, y)
    params, losses = train(ctx, Xt, yt, steps=steps, batch=batch, lr=lr)
    held = [collect(ctx, workloads.config(config, seed=s, n=n), holdout_md_steps)[:2] for s in holdout_seeds]
    Xh = np.concatenate([h[0] for h in held])
    yh = np.concatenate([h[1] for h in held])
    thr = calibrate(ctx, params, Xh, yh, margin)
    return dict(params=params, threshold=thr, losses=losses, train_pairs=len(y), holdout_pairs=len(yh),
                holdout_drop=float(np.mean(dropped(ctx, params, thr, Xh))), holdout_beyond=float(np.mean(yh < 0.5)))
