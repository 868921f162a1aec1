"""The simulator front end (neraulgraphics_amd/simulators.py): shapes to meshes, instances and a body binding, as the
reference's adapters build them (python/neraulgraphics/newton.py:48-89), and the per-step flow of body poses.
Newton, MuJoCo and Warp are not installed, so the simulators themselves are not exercised; the pose math is checked
against the oracle's FP64 body composition (the reference host's, src/instances.cpp:126-164)."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import shapes
from neraulgraphics_amd.simulators import Shape, SimBatchRenderer, build_scene
from oracle.oracle import forward


def _shapes():
    rng = np.random.default_rng(5)

    def quat():
        q = rng.normal(size=4)
        return tuple(q / np.linalg.norm(q))
    return [Shape("sphere", (0.3,), (0.0, 0.0, 0.2), quat(), 0, (0.9, 0.1, 0.1)),
            Shape("sphere", (0.5,), (0.1, 0.0, 0.0), quat(), 1),
            Shape("box", (0.2, 0.4, 0.1), (0.0, 0.3, 0.0), quat(), 1, (0.1, 0.9, 0.1)),
            Shape("capsule", (0.1, 0.3), (0.0, 0.0, 0.0), quat(), 2),
            Shape("capsule", (0.1, 0.3), (0.2, 0.0, 0.0), quat(), 3),
            Shape("capsule", (0.2, 0.3), (0.0, 0.2, 0.0), quat(), 3),
            Shape("cylinder", (0.15, 0.4), (0.0, 0.0, 0.0), quat(), 0),
            Shape("cone", (0.2, 0.3), (0.0, 0.0, 0.1), quat(), 2),
            Shape("ellipsoid", (0.1, 0.2, 0.3), (0.0, 0.0, 0.0), quat(), -1),
            Shape("plane", (0.0, 0.0), (0.0, 0.0, -1.5), (0.0, 0.0, 0.0, 1.0), -1)]


def test_shapes_share_meshes_and_scale_their_instances():
    sc = build_scene(_shapes())
    mesh_of = [i.mesh for i in sc.instances]
    assert mesh_of[0] == mesh_of[1] == mesh_of[8]          # spheres and the ellipsoid: one unit sphere
    assert mesh_of[3] == mesh_of[4] != mesh_of[5]          # capsules: one mesh per size
    assert len(sc.meshes) == 7                              # sphere, box, 2 capsules, cylinder, cone, plane
    assert np.allclose(sc.scale[[0, 1, 2, 6, 7, 8]], [[0.3] * 3, [0.5] * 3, [0.2, 0.4, 0.1], [0.15, 0.15, 0.4],
                                                      [0.2, 0.2, 0.3], [0.1, 0.2, 0.3]])
    assert np.allclose(sc.scale[[3, 4, 5, 9]], 1.0)
    assert np.isclose(sc.meshes[mesh_of[9]].vertices[:, 0].max(), 500.0)  # an unbounded plane is 1000 wide
    assert sc.body_of_instance.tolist() == [0, 1, 1, 2, 3, 3, 0, 2, -1, -1]
    assert np.allclose(np.linalg.norm(sc.local_pose[:, 3:], axis=1), 1.0, atol=1e-6)
    assert np.allclose(sc.instances[0].material.as_array()[:3], (0.9, 0.1, 0.1))
    with pytest.raises(ValueError, match="not drawn"):
        build_scene([Shape("heightfield")])


def test_body_binding_places_shapes_as_the_pose_math_does(oracle):
    """What set_body_poses computes for a binding (the oracle's FP64 composition) equals composing each shape's pose
    on its body and scaling along the mesh's axes (shapes.compose_poses, shapes.pose_matrices)."""
    sc = build_scene(_shapes())
    worlds, bodies = 3, 4
    rng = np.random.default_rng(9)
    q = rng.normal(size=(worlds, bodies, 4))
    poses = np.concatenate([rng.uniform(-2, 2, size=(worlds, bodies, 3)), q / np.linalg.norm(q, axis=-1, keepdims=True)], -1)
    poses = poses.astype(np.float32)
    n = len(sc.instances)
    world_off = np.arange(worlds + 1, dtype=np.int32) * n
    got = oracle.body_poses_to_matrices(world_off, bodies, sc.body_of_instance, sc.local_pose, sc.scale, poses).reshape(worlds, n, 3, 4)
    for w in range(worlds):
        for i in range(n):
            b = sc.body_of_instance[i]
            lp, lq = sc.local_pose[i:i + 1, :3], sc.local_pose[i:i + 1, 3:]
            if b >= 0:
                p, qq = shapes.compose_poses(poses[w, b:b + 1, :3], poses[w, b:b + 1, 3:], lp, lq)
            else:
                p, qq = lp, lq
            want = shapes.pose_matrices(p, qq, sc.scale[i:i + 1])[0]
            assert np.allclose(got[w, i], want, atol=2e-6), (w, i)


@pytest.mark.gpu
def test_a_simulated_batch_renders_as_the_oracle(gpu, oracle):
    from tests.helpers import compare_images
    worlds, bodies = 3, 4
    shp = _shapes()
    rng = np.random.default_rng(2)
    q = rng.normal(size=(worlds, bodies, 4))
    poses = np.concatenate([rng.uniform(-1, 1, size=(worlds, bodies, 3)), q / np.linalg.norm(q, axis=-1, keepdims=True)], -1)
    poses = poses.astype(np.float32)
    cams = np.stack([shapes.look_at([3.0, -3.0, 1.5], [0.0, 0.0, 0.0])[0].reshape(12)] * worlds)
    sim = SimBatchRenderer(shp, worlds, bodies, 48, 40)
    out = sim.render(poses, cams, 1.0, depth=True, normals=True, object_ids=True)
    sc = sim.scene
    cfg = dict(width=48, height=40, worlds=worlds, cameras_per_world=1, meshes=sc.meshes, instances=sc.instances, counts=None,
               mode="bodies", bodies=bodies, body_of_instance=sc.body_of_instance, local_pose=sc.local_pose, scale=sc.scale,
               body_poses=poses, cameras=cams, fov=1.0)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "simulated batch")
    assert rep["hits"] > 0.3 * rep["pixels"]
