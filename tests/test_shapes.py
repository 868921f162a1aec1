"""Pose math and shape generators for simulator callers (neraulgraphics_amd/shapes.py, the surface of the
reference's python/neraulgraphics/shapes.py; its tests/test_shapes.py:16-71 states the properties checked here).
The rotations are checked against scipy's, an independent implementation of the same (x, y, z, w) convention."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from neraulgraphics_amd import shapes


def _poses(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    return rng.normal(size=(n, 3)), q / np.linalg.norm(q, axis=1, keepdims=True)


def test_quat_to_matrix_is_scipys_rotation():
    _, q = _poses(64, 0)
    r = shapes.quat_to_matrix(q)
    assert np.allclose(r, Rotation.from_quat(q).as_matrix(), atol=1e-12)
    assert np.allclose(r @ r.transpose(0, 2, 1), np.eye(3), atol=1e-12) and np.allclose(np.linalg.det(r), 1.0)


def test_a_quarter_turn_about_z_takes_x_to_y_and_length_does_not_matter():
    s = np.sin(np.pi / 4)
    for scale in (1.0, 2.0, 0.1):
        q = scale * np.array([[0.0, 0.0, s, s]])
        assert np.allclose(shapes.quat_to_matrix(q)[0] @ [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])


def test_compose_poses_is_the_product_of_the_rigid_motions():
    p1, q1 = _poses(40, 1)
    p2, q2 = _poses(40, 2)
    p, q = shapes.compose_poses(p1, q1, p2, q2)
    want = Rotation.from_quat(q1) * Rotation.from_quat(q2)
    assert np.allclose(shapes.quat_to_matrix(q), want.as_matrix(), atol=1e-12)
    assert np.allclose(p, p1 + Rotation.from_quat(q1).apply(p2), atol=1e-12)
    x = np.random.default_rng(3).normal(size=(40, 3))  # applying the composition = applying one after the other
    one_by_one = p1 + Rotation.from_quat(q1).apply(p2 + Rotation.from_quat(q2).apply(x))
    assert np.allclose(p + Rotation.from_quat(q).apply(x), one_by_one, atol=1e-12)


def test_pose_matrices_scale_along_the_meshs_own_axes():
    p, q = _poses(5, 4)
    scale = np.tile([2.0, 3.0, 4.0], (5, 1))
    m = shapes.pose_matrices(p, q, scale)
    assert m.shape == (5, 3, 4) and m.dtype == np.float32
    x = np.array([1.0, 1.0, 1.0])
    want = Rotation.from_quat(q).apply(scale * x) + p
    assert np.allclose(np.einsum("nij,j->ni", m[:, :, :3], x) + m[:, :, 3], want, atol=1e-5)


@pytest.mark.parametrize("make, triangles", [(lambda: shapes.cylinder(1, 1, 16), 64), (lambda: shapes.cone(1, 1, 16), 32),
                                             (lambda: shapes.plane(2, 3), 2), (lambda: shapes.uv_sphere(1, 12, 6), 12 * 2 * 5),
                                             (lambda: shapes.capsule(0.3, 0.7, 12, 4), 2 * 12 + 7 * 24)])
def test_triangle_counts(make, triangles):
    assert make().triangle_count == triangles


@pytest.mark.parametrize("mesh", [shapes.capsule(0.3, 0.7, 12, 4), shapes.uv_sphere(1.0, 12, 6), shapes.cylinder(1, 1, 9),
                                  shapes.cone(1, 1, 7)])
def test_closed_surfaces(mesh):
    """V - E + F = 2 with every edge shared by exactly two triangles, and the vertices where the sizes put them."""
    t = mesh.indices.reshape(-1, 3).astype(np.int64)
    edges = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert mesh.vertex_count - len(counts) + len(t) == 2


def test_sizes():
    c = shapes.capsule(0.5, 1.0, 16, 4).vertices
    assert np.isclose(c[:, 2].max(), 1.5) and np.isclose(c[:, 2].min(), -1.5)
    assert np.isclose(np.hypot(c[:, 0], c[:, 1]).max(), 0.5, atol=1e-6)
    s = shapes.uv_sphere(2.0, 16, 8).vertices
    assert np.allclose(np.linalg.norm(s, axis=1), 2.0, atol=1e-5)


def test_look_at_frames_the_target():
    m = shapes.look_at([[3.0, 0.0, 1.0], [0.0, -2.0, 0.5]], [0.0, 0.0, 0.5])
    for cam, eye in zip(m, ([3.0, 0.0, 1.0], [0.0, -2.0, 0.5])):
        fwd = -cam[:, 2]
        to_target = np.array([0.0, 0.0, 0.5]) - eye
        assert np.allclose(fwd, to_target / np.linalg.norm(to_target), atol=1e-6)
        assert abs(cam[2, 0]) < 1e-6  # right is horizontal for a Z-up world
        assert np.allclose(cam[:, 3], eye)
    with pytest.raises(ValueError, match="up axis"):
        shapes.look_at([0.0, 0.0, 5.0], [0.0, 0.0, 0.0])


def test_generators_reject_too_few_segments():
    with pytest.raises(ValueError, match="segments"):
        shapes.cylinder(1, 1, 2)
