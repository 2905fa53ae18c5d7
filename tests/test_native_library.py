"""The C-ABI libraries build, load and export every symbol their headers declare (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csgh?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("csg.h", "libcsg.so"), ("csg_host.h", "libcsg_host.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, "curvedspacesim_amd", lib)
    assert os.path.exists(path), f"{lib} not built"
    if lib == "libcsg_host.so":
        ctypes.CDLL(os.path.join(ROOT, "curvedspacesim_amd", "libcsg.so"))
    L = ctypes.CDLL(path)
    names = _declared(header)
    assert names
    for n in names:
        assert hasattr(L, n), f"{lib} does not export {n}"


def test_context_creation_fails_cleanly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import curvedspacesim_amd as csa

    with pytest.raises(csa.CsgError):
        csa.Context(0)


def test_mesh_generators():
    import numpy as np

    import curvedspacesim_amd as csa
    from oracle.pyoracle import OracleMesh

    for level, nv, nf in [(0, 12, 20), (4, 2562, 5120), (6, 40962, 81920)]:
        X, T = csa.icosphere(level)
        assert X.shape == (nv, 3) and T.shape == (nf, 3)
        assert np.allclose(np.linalg.norm(X, axis=1), 1.0)
    X, T = csa.jittered_torus(3.0, 1.0, 200, 100, 0.02, 13377)
    assert X.shape == (20000, 3) and T.shape == (40000, 3)
    c = OracleMesh(X, T).counts()  # consistent orientation, closed
    assert c["border"] == 0
    n = np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]])
    ring = X.copy()
    ring[:, 2] = 0
    ring = 3.0 * ring / np.linalg.norm(ring, axis=1)[:, None]
    outward = ((X[T].mean(1) - ring[T].mean(1)) * n).sum(1) > 0
    assert outward.mean() > 0.999
    X, T = csa.bumpy_sphere(6, 20, 0.15, 0.3, 13377)
    assert T.shape == (81920, 3)
    r = np.linalg.norm(X, axis=1)
    assert r.min() >= 1.0 - 1e-12 and r.max() <= 1.0 + 0.15 * 20
    # Euler characteristic 2 (sphere) and 0 (torus)
    assert len(X) - 3 * len(T) // 2 + len(T) == 2
