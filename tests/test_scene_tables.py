"""The builtin quad and cube of the scene format (Mesh.fs:126-155), pinned by
their geometry rather than by their tables: the quad is the square [-1, 1]^2
in the plane y = 0 with its normal cross(p1 - p0, p2 - p0) along -Y; the cube
is [-1, 1]^3, closed (every edge shared by two triangles in opposite
directions), with every normal pointing outwards.  A swapped index flips a
normal or opens an edge and fails."""
import numpy as np

from barnacle_amd import scene


def tris(v, ix):
    return v[np.asarray(ix).reshape(-1, 3)].astype(np.float64)


def test_builtin_quad():
    t = tris(scene.QUAD_VERTICES, scene.QUAD_INDICES)
    assert t.shape == (2, 3, 3)
    assert np.all(t[..., 1] == 0) and np.all(np.abs(t[..., [0, 2]]) == 1)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.allclose(n / np.linalg.norm(n, axis=1, keepdims=True), [0, -1, 0])
    assert np.isclose(0.5 * np.linalg.norm(n, axis=1).sum(), 4.0)


def test_builtin_cube():
    t = tris(scene.CUBE_VERTICES, scene.CUBE_INDICES)
    assert t.shape == (12, 3, 3)
    assert np.all(np.abs(t) == 1)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(axis=1)) > 0)          # outward
    assert np.isclose(0.5 * np.linalg.norm(n, axis=1).sum(), 24.0)
    ix = np.asarray(scene.CUBE_INDICES).reshape(-1, 3)
    edges = {}
    for a, b, c in ix:
        for e in ((a, b), (b, c), (c, a)):
            edges[e] = edges.get(e, 0) + 1
    assert all(n == 1 and edges.get((e[1], e[0])) == 1 for e, n in edges.items())
