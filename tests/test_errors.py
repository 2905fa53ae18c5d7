"""The reference's throw sites on the CPU side: the OFF reader of the host layer (triangulatedMeshSpace::
loadMeshFromFile, src/models/triangulatedMeshSpace.cpp:42-72: unreadable file, non-triangular mesh) and the oracle's
mesh checks and walk failures (src/models/triangulatedMeshSpace.cpp:522-548, src/utility/meshUtilities.cpp:184-195)."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
from oracle.pyoracle import OracleMesh


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_off_reader_accepts_a_triangle_mesh(tmp_path):
    X, T = csa.read_off(_write(tmp_path, "tet.off", "OFF\n4 4 0\n1 1 1\n1 -1 -1\n-1 1 -1\n-1 -1 1\n"
                                                     "3 0 1 2\n3 0 3 1\n3 0 2 3\n3 1 3 2\n"))
    assert X.shape == (4, 3) and T.shape == (4, 3) and T[1].tolist() == [0, 3, 1]


@pytest.mark.parametrize("text,what", [
    ("OFF\n4 1 0\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n", "non-triangular"),  # a quad: is_triangle_mesh fails
    ("PLY\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n", "not an OFF"),
    ("OFF\n3 2 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n", "truncated"),
])
def test_off_reader_rejects(tmp_path, text, what):
    with pytest.raises(csa.CsgError, match=what):
        csa.read_off(_write(tmp_path, "bad.off", text))


def test_off_reader_missing_file(tmp_path):
    with pytest.raises(csa.CsgError, match="cannot open"):
        csa.read_off(str(tmp_path / "absent.off"))


@pytest.mark.parametrize("case", ["out_of_range", "degenerate", "non_manifold", "flipped"])
def test_oracle_mesh_checks(case):
    X, T = sm.tetrahedron()
    T = T.copy()
    if case == "out_of_range":
        T[0, 0] = 9
    elif case == "degenerate":
        T[0, 1] = T[0, 0]
    elif case == "non_manifold":  # a fifth face on edge (0, 1): three faces share it
        X = np.vstack([X, [[2.0, 0.0, 0.0]]])
        T = np.vstack([T, [[1, 0, 4]]])
    else:  # one face flipped: its edges run the same way as its neighbours'
        T[0] = T[0][[0, 2, 1]]
    with pytest.raises(ValueError):
        OracleMesh(X, T)


def test_oracle_walk_failure_codes():
    """Border edge and border vertex of a closed-surface walk, and a NaN target, give the oracle's codes and leave
    the particle where it was."""
    X, T = sm.grid(3, 3)
    m = OracleMesh(X, T)
    f0, b0 = np.array([8, 8, 8], np.int32), np.array([[0.5, 0.25, 0.25]] * 3)
    P0 = m.positions(f0[:1], b0[:1])[0]  # (1.75, 1.25) in square (1, 1)
    disp = [[5.0, 0.1, 0.0],                                    # leaves through the border edge x = 3
            list(2.0 * (np.array([3.0, 3.0, 0.0]) - P0)),        # aimed through the border corner (3, 3)
            [np.nan, 0.0, 0.0]]
    f, b, d, v, st = m.transport(f0, b0, disp)
    assert st[0] == -11 and st[2] == -4
    assert st[1] in (-9, -11)  # a corner hit at the border is a border vertex, or the border edge it rounds onto
    assert np.array_equal(f, f0) and np.allclose(b, b0)
