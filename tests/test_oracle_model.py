"""Pins the oracle's candidate lists and potentials (callers of the geodesic solve)."""
import numpy as np
import pytest

from oracle import pyoracle


def test_cell_candidates_match_bruteforce(rng):
    N = 800
    pos = rng.uniform(-1.2, 1.3, size=(N, 3))
    lo = np.minimum(pos.min(0), 0.0)
    hi = np.maximum(pos.max(0), 0.0)
    r = 0.23
    off, idx, md = pyoracle.cell_candidates(pos, lo, hi, r, r)
    D = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    for i in range(N):
        got = idx[off[i]:off[i + 1]]
        want = np.where((D[i] ** 2 < r * r) & (np.arange(N) != i))[0]
        assert sorted(got.tolist()) == want.tolist()
        assert abs(md[i] - (D[i, want].max() if len(want) else 0.0)) < 1e-15


def test_cell_candidate_order_is_reference_order():
    """Cells x-outermost then y then z, each cell by particle index (hyperRectangularCellList.cpp:133-145, :82-125)."""
    lo, hi = np.zeros(3), np.array([3.0, 3.0, 3.0])
    # one particle per cell centre of a 3x3x3 grid, inserted in a scrambled order; source in the middle cell
    centres = np.array([[x + 0.5, y + 0.5, z + 0.5] for x in range(3) for y in range(3) for z in range(3)])
    perm = np.random.default_rng(3).permutation(27)
    pos = centres[perm]
    src = int(np.where(perm == 13)[0][0])
    off, idx, md = pyoracle.cell_candidates(pos, lo, hi, 1.0, 10.0)
    got = idx[off[src]:off[src + 1]]
    cell_of = {int(i): int(perm[i]) for i in range(27)}
    order = [cell_of[int(j)] for j in got]
    # cell id x*9+y*3+z in centres; the reference loops xx, yy, zz
    assert order == [c for c in range(27) if c != 13]


@pytest.mark.parametrize("d", [0.0, 0.1, 0.29, 0.3, 0.31])
def test_harmonic_closed_form(d):
    k, s = 11.11, 0.3
    e, sc = pyoracle.pair_force(0, [k, s, 0], d)
    assert sc == pytest.approx(-k * (s - d) if d <= s else 0.0, abs=1e-15)
    assert e == pytest.approx(0.5 * k * (s - d) ** 2 if d < s else 0.0, abs=1e-15)


@pytest.mark.parametrize("d", [0.0, 0.05, 0.149, 0.15, 0.2])
def test_truncated_gaussian_closed_form(d):
    eps, s, rc = 1.0, 0.05, 0.15
    e, sc = pyoracle.pair_force(2, [eps, s, rc], d)
    if d < rc:
        assert e == pytest.approx(eps * np.exp(-d * d / (2 * s * s)), rel=1e-14)
        # the force magnitude is -dU/dd
        h = 1e-7
        e1, _ = pyoracle.pair_force(2, [eps, s, rc], d + h) if d + h < rc else (None, None)
        if e1 is not None and d > 0:
            e0, _ = pyoracle.pair_force(2, [eps, s, rc], d - h)
            assert sc == pytest.approx(-(e0 - e1) / (2 * h), rel=1e-6)
    else:
        assert e == 0 and sc == 0


def test_reference_gaussian_quirk():
    """gaussianRepulsion's force uses sigma^(3/2) (gaussianRepulsion.h:23), so it is not -dE/dd unless sigma = 1."""
    a, s, d = 1.0, 0.5, 0.3
    e, sc = pyoracle.pair_force(1, [a, s, 0], d)
    assert e == pytest.approx(a * np.exp(-d * d / (2 * s * s)) / (np.sqrt(2 * np.pi) * s), rel=1e-15)
    assert sc == pytest.approx(-d * a * np.exp(-d * d / (2 * s * s)) / (np.sqrt(2 * np.pi) * s ** 1.5), rel=1e-15)
