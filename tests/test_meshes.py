"""The reference's procedural meshes (swegl/data/model.hpp:181-466) through the
library's generators: closed forms (strip layout, texture coordinates, the
triangle) and the geometry's invariants (points on the torus and the
sphere, normals of unit length pointing out of the surface), and
calculate_face_normals on shapes whose normals are known."""
import numpy as np
import pytest

from swegl_amd import _abi as A
from swegl_amd import host, scenes

F = np.float32


@pytest.mark.parametrize("precision", [3, 10, 100])
def test_tore_layout_and_geometry(precision):
    node = scenes.make_tore(precision, 4)
    n = precision + 1
    assert len(node.primitives) == n
    for bg, p in enumerate(node.primitives):
        assert (p.mode, p.material_id) == (A.TRIANGLE_STRIP, 4)
        assert len(p.positions) == 2 * n
        want = np.empty(2 * n, np.uint32)
        want[0::2] = n + np.arange(n)
        want[1::2] = np.arange(n)
        assert np.array_equal(p.indices, want)
        for row in range(2):
            uv = p.tex_coords[row * n:(row + 1) * n]
            assert np.array_equal(uv[:, 0], np.full(n, F((bg + row) / precision)))
            assert np.array_equal(uv[:, 1], (np.arange(n) / precision).astype(F))
        pos = p.positions.astype(np.float64)
        nor = p.normals.astype(np.float64)
        # on the torus of radii 2 (about the y axis) and 0.8
        ring = np.hypot(pos[:, 0], pos[:, 2])
        assert np.allclose((ring - 2) ** 2 + pos[:, 1] ** 2, 0.64, atol=2e-5)
        assert np.allclose(np.linalg.norm(nor, axis=1), 1, atol=1e-6)
        # the tube centre p - 0.8 n lies on the major circle
        c = pos - 0.8 * nor
        assert np.allclose(c[:, 1], 0, atol=2e-5)
        assert np.allclose(np.hypot(c[:, 0], c[:, 2]), 2, atol=2e-5)
    # the first vertex: the tube's outer point on +x
    assert np.allclose(node.primitives[0].positions[0], [2.8, 0, 0], atol=1e-6)


@pytest.mark.parametrize("precision,radius", [(4, 1.0), (100, 2.0)])
def test_sphere_layout_and_geometry(precision, radius):
    node = scenes.make_sphere(precision, radius, 2)
    n = precision + 1
    assert len(node.primitives) == n
    for bg, p in enumerate(node.primitives):
        assert (p.mode, p.material_id) == (A.TRIANGLE_STRIP, 2)
        for row in range(2):
            uv = p.tex_coords[row * n:(row + 1) * n]
            assert np.array_equal(uv[:, 0], (np.arange(n) / precision).astype(F))
            assert np.array_equal(uv[:, 1], np.full(n, F((bg + row) / precision)))
        pos = p.positions.astype(np.float64)
        assert np.allclose(np.linalg.norm(pos, axis=1), radius, atol=2e-5 * radius)
        assert np.allclose(np.linalg.norm(p.normals, axis=1), 1, atol=1e-6)
        assert np.allclose(np.sum(pos * p.normals, axis=1), radius, atol=2e-5 * radius)
    # each row starts at the north pole and turns through half a circle
    p0 = node.primitives[0]
    assert np.array_equal(p0.positions[0], [0, F(radius), 0])
    assert np.allclose(p0.positions[n - 1], [0, -radius, 0], atol=1e-5)


def test_make_tri():
    node = scenes.make_tri(2, 5)
    p = node.primitives[0]
    assert np.array_equal(p.positions, [[0, 0, 0], [2, 0, 0], [0, 2, 0]])
    assert np.array_equal(p.tex_coords, [[0, 0], [0, 1], [1, 0]])
    assert np.array_equal(p.normals, np.tile([0, 0, 1], (3, 1)))
    assert (p.mode, p.material_id, list(p.indices)) == (A.TRIANGLES, 5, [0, 1, 2])


TETRA = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], F)


def prim(mode, idx):
    return host.primitive_t(TETRA, np.full((4, 3), 7, F), np.zeros((4, 2)), idx, mode, 0)


def test_face_normals_strip():
    """Strip 0,1,2,3: the first triangle's normal (+z) for 0, 1, 2.  The loop
    moves i0 and i1 only at the end of an iteration, so vertex 3 comes from
    (0, 1, 3), odd, operands swapped: cross(p3 - p0, p1 - p0) = +y."""
    p = prim(A.TRIANGLE_STRIP, [0, 1, 2, 3])
    scenes.calculate_face_normals(p)
    assert np.array_equal(p.normals[:3], np.tile([0, 0, 1], (3, 1)))
    assert np.array_equal(p.normals[3], [0, 1, 0])


def test_face_normals_fan():
    """Fan 0,1,2,3: +z for 0, 1, 2, then vertex 3 from (0, 1, 3) (i1 moves at
    the end of the iteration): cross(p1 - p0, p3 - p0) = -y; the reference's
    last loop (written for triangle lists, run for fans) sets 0, 1, 2 again
    from (0, 1, 2)."""
    p = prim(A.TRIANGLE_FAN, [0, 1, 2, 3])
    scenes.calculate_face_normals(p)
    assert np.array_equal(p.normals[:3], np.tile([0, 0, 1], (3, 1)))
    assert np.array_equal(p.normals[3], [0, -1, 0])
    # Fan 1,0,2,3,1,2: the fan pass gives -z to 1, 0, 2 and +y to 3, then
    # zero vectors (degenerate triangles; a zero length is left alone) to 1
    # and 2.  The last loop takes index triples (1,0,2) -> -z for 1, 0, 2
    # and (3,1,2) -> cross(p1 - p3, p2 - p3) = (1,1,1) normalised for 3, 1, 2.
    q = prim(A.TRIANGLE_FAN, [1, 0, 2, 3, 1, 2])
    scenes.calculate_face_normals(q)
    d = np.array([1, 0, -1], F), np.array([0, 1, -1], F)
    c = np.cross(d[0], d[1]).astype(F)
    c = c / np.sqrt(F(np.sum(c * c, dtype=F)), dtype=F)
    assert np.array_equal(q.normals[0], [0, 0, -1])
    for v in (1, 2, 3):
        assert np.allclose(q.normals[v], c, atol=1e-7)


def test_face_normals_triangle_list_untouched():
    p = prim(A.TRIANGLES, [0, 1, 2, 0, 2, 3])
    scenes.calculate_face_normals(p)
    assert np.array_equal(p.normals, np.full((4, 3), 7, F))


def test_demo_scene_builds():
    s = scenes.demo_scene(tore_precision=10, sphere_precision=10)
    assert len(s.nodes) == 8 and s.root_nodes == list(range(8))
    assert [n.primitives[0].material_id for n in s.nodes] == [1, 0, 2, 5, 4, 4, 6, 7]
    assert s.n_triangles() == 2 * 11 * 20 + 3 * 12 + 3
