"""The simulator front end (neraulgraphics_amd/simulators.py): shapes to meshes, instances and a body binding, as the
reference's adapters build them (python/neraulgraphics/newton.py:48-89), and the per-step flow of body poses; MuJoCo
models to shapes and geom frames to instance matrices (python/neraulgraphics/mujoco.py:68-136, :177-302).
Newton, MuJoCo and Warp are not installed, so the simulators themselves are not exercised: a MuJoCo model is stood for
by an object holding the MjModel fields the adapter reads. The pose math is checked against the oracle's FP64 body
composition (the reference host's, src/instances.cpp:126-164)."""
from types import SimpleNamespace

import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import shapes
from neraulgraphics_amd.simulators import (MJWarpBatchRenderer, NewtonBatchRenderer, Shape, SimBatchRenderer, build_scene, mujoco_shapes,
                                          newton_shapes, scene_bounds)
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


def _mujoco_model():
    """The MjModel fields the MuJoCo adapter reads, for nine geoms: a plane, a sphere taking its material's colour,
    a capsule, a box in group 3 (hidden), an invisible ellipsoid, a cylinder, a mesh, a heightfield (not drawn) and the
    same mesh again."""
    ico = ng.make_icosphere(0.25, 1)
    grey = [0.5, 0.5, 0.5, 1.0]
    return SimpleNamespace(
        ngeom=9,
        geom_type=np.array([0, 2, 3, 6, 4, 5, 7, 1, 7]),
        geom_size=np.array([[2.0, 3.0, 0.1], [0.3, 0, 0], [0.1, 0.25, 0], [0.2, 0.3, 0.1], [0.1, 0.2, 0.3], [0.15, 0.2, 0],
                            [0, 0, 0], [1, 1, 1], [0, 0, 0]]),
        geom_rgba=np.array([[0.3, 0.3, 0.3, 1], grey, [0.1, 0.8, 0.1, 1], grey, [0.2, 0.2, 0.9, 0.0], [0.7, 0.7, 0.1, 1],
                            [0.2, 0.4, 0.9, 1], grey, [0.9, 0.5, 0.1, 1]], np.float32),
        geom_group=np.array([0, 0, 1, 3, 0, 2, 0, 0, 2]),
        geom_matid=np.array([-1, 0, -1, -1, -1, -1, -1, -1, -1]),
        mat_rgba=np.array([[0.9, 0.2, 0.2, 1.0]], np.float32),
        geom_dataid=np.array([-1, -1, -1, -1, -1, -1, 0, 0, 0]),
        mesh_vertadr=np.array([0]), mesh_vertnum=np.array([len(ico.vertices)]),
        mesh_faceadr=np.array([0]), mesh_facenum=np.array([len(ico.indices) // 3]),
        mesh_vert=ico.vertices, mesh_face=ico.indices.reshape(-1, 3))


def test_mujoco_geoms_become_shapes():
    model = _mujoco_model()
    shp, geoms = mujoco_shapes(model)
    assert geoms.tolist() == [0, 1, 2, 5, 6, 8]
    assert [s.kind for s in shp] == ["plane", "sphere", "capsule", "cylinder", "mesh", "mesh"]
    assert shp[0].size == (4.0, 6.0)                      # half-extents to full sizes
    assert shp[1].size == (0.3,) and shp[2].size == (0.1, 0.25) and shp[3].size == (0.15, 0.2)
    assert np.allclose(shp[1].color, (0.9, 0.2, 0.2))     # the default colour gives way to the material's
    assert np.allclose(shp[2].color, (0.1, 0.8, 0.1))
    assert shp[4].mesh is shp[5].mesh                     # one Mesh per MuJoCo mesh
    shp3, geoms3 = mujoco_shapes(model, geom_groups=(3,))
    assert geoms3.tolist() == [3] and shp3[0].kind == "box" and shp3[0].size == (0.2, 0.3, 0.1)
    sc = build_scene(shp)
    assert np.allclose(sc.scale[[1, 3]], [[0.3] * 3, [0.15, 0.15, 0.2]])


def test_scene_bounds():
    c, e = scene_bounds([[0, 0, 0], [2, 0, 0]], [1.0, 1.0])
    assert np.allclose(c, [1, 0, 0]) and np.isclose(e, np.sqrt(16 + 4 + 4) / 2)
    c, e = scene_bounds(np.zeros((0, 3)), np.zeros(0))
    assert np.allclose(c, 0) and e == 1.0


def _geom_frames(worlds, ngeom, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(worlds * ngeom, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    xmat = shapes.quat_to_matrix(q).astype(np.float32).reshape(worlds, ngeom, 3, 3)
    xpos = rng.uniform(-0.8, 0.8, size=(worlds, ngeom, 3)).astype(np.float32)
    xpos[:, 0] = (0.0, 0.0, -1.0)                     # the plane, below
    xmat[:, 0] = np.eye(3, dtype=np.float32)
    return xpos, xmat, q.reshape(worlds, ngeom, 4)


@pytest.mark.gpu
def test_geom_frames_place_instances_and_render_as_the_oracle(gpu, oracle):
    """ngx_set_geom_frames composes R diag(s) | p in FP32 (the Warp compose kernel's arithmetic, bit for bit); the
    same placement from the quaternions through the oracle's FP64 body composition agrees to rounding (a transposed
    rotation or a scale on the wrong axis would not); and the images equal the oracle's for the placed matrices."""
    import torch
    from tests.helpers import compare_images
    model, worlds = _mujoco_model(), 3
    xpos, xmat, q = _geom_frames(worlds, model.ngeom, 11)
    sim = MJWarpBatchRenderer(model, worlds, 48, 40, data=SimpleNamespace(geom_xpos=xpos))
    cams = sim.default_cameras().reshape(worlds, 12)
    host = sim.render(SimpleNamespace(geom_xpos=xpos, geom_xmat=xmat), cams, rgb=False, depth=True, normals=True,
                      object_ids=True)
    m_host = sim.renderer.read_build()["matrices"].reshape(worlds, -1, 3, 4)
    dev = SimpleNamespace(geom_xpos=torch.from_numpy(xpos).to(gpu), geom_xmat=torch.from_numpy(xmat).to(gpu))
    out = sim.render(dev, cams, rgb=False, depth=True, normals=True, object_ids=True)
    torch.cuda.synchronize()
    m_dev = sim.renderer.read_build()["matrices"].reshape(worlds, -1, 3, 4)
    g, sc = sim._geoms, sim.scene.scale
    want = np.empty_like(m_host)
    want[..., :3] = xmat[:, g] * sc[None, :, None, :]
    want[..., 3] = xpos[:, g]
    assert np.array_equal(m_host.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(m_dev.view(np.uint32), want.view(np.uint32))
    for k in ("depth", "object_ids", "normals"):
        assert np.array_equal(out[k], host[k]), k
    # the same placement from the quaternions, composed in FP64 as a body pose with an identity local pose
    n = len(g)
    poses = np.concatenate([xpos[:, g], q[:, g]], axis=-1).astype(np.float32).reshape(worlds, n, 7)
    poses[:, 0, 3:] = (0.0, 0.0, 0.0, 1.0)
    local = np.tile(np.array([0, 0, 0, 0, 0, 0, 1], np.float32), (n, 1))
    world_off = np.arange(worlds + 1, dtype=np.int32) * n
    via_q = oracle.body_poses_to_matrices(world_off, n, np.arange(n, dtype=np.int32), local, sc, poses)
    assert np.allclose(via_q.reshape(want.shape), want, atol=2e-6)
    cfg = dict(width=48, height=40, worlds=worlds, cameras_per_world=1, meshes=sim.scene.meshes, instances=sim.scene.instances,
               counts=None, mode="matrices", matrices=m_host.reshape(worlds, n, 12), cameras=cams, fov=1.0)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "mujoco geoms")
    assert rep["hits"] > 0.3 * rep["pixels"]
    seen = set(np.unique(out["object_ids"]).tolist()) - {ng.OBJECT_NONE}
    assert seen <= set(sim.object_id_of_geom.values()) and len(seen) >= 3


@pytest.mark.gpu
def test_geom_binding_is_checked(gpu):
    model = _mujoco_model()
    sim = MJWarpBatchRenderer(model, 2, 16, 16)
    with pytest.raises(ValueError, match="geom"):
        sim.renderer.bind_geoms(np.full(6, 9), sim.scene.scale, model.ngeom)   # geom 9 of 9
    bad = np.zeros((2, model.ngeom, 3), np.float32)
    bad[1, 4, 2] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        sim.renderer.set_geom_frames(bad, np.tile(np.eye(3, dtype=np.float32), (2, model.ngeom, 1, 1)))
    r = ng.BatchRenderer(16, 16, 1, meshes=[ng.make_box((1, 1, 1))], instances=[ng.Instance(0)])
    with pytest.raises(ValueError, match="bind_geoms"):
        r.set_geom_frames(np.zeros((1, 1, 3), np.float32), np.zeros((1, 1, 9), np.float32))


@pytest.mark.gpu
def test_the_reference_benchmark_scene_renders_as_the_oracle(gpu, oracle):
    """The falling-bodies scene of the reference's batch-renderer benchmark (bench/scene.py), drawn through
    MJWarpBatchRenderer as bench/run_ours.py draws it (geom frames from the GPU, one hard-shadow light, RGB, one
    sample), against the oracle: depth and IDs to the usual bar, the picture to tests/test_batch.cpp:177's bar of
    0.05 of a level mean per channel."""
    import torch
    from tests.helpers import compare_images
    from neraulgraphics_amd import workloads
    fb = workloads.falling_bodies(8, worlds=6, frames=3)
    light = ng.Light(fb["light"], intensity=1.0, radius=0.0)
    sim = MJWarpBatchRenderer(fb["model"], 6, 48, 48, lights=[light], fov=fb["fov"])
    cams = np.tile(fb["camera"].reshape(1, 12), (6, 1))
    dev_cams = torch.from_numpy(cams).to(gpu)
    for f in (0, 2):
        data = SimpleNamespace(geom_xpos=torch.from_numpy(fb["geom_xpos"][f]).to(gpu),
                               geom_xmat=torch.from_numpy(fb["geom_xmat"][f]).to(gpu))
        out = sim.render(data, dev_cams, rgb=True, depth=True, object_ids=True, zero_copy=True)
        got = {k: v.cpu().numpy() for k, v in out.items()}
        mats = sim.renderer.read_build()["matrices"].reshape(6, 8, 12)
        sc = sim.scene
        cfg = dict(width=48, height=48, worlds=6, cameras_per_world=1, meshes=sc.meshes, instances=sc.instances, counts=None,
                   mode="matrices", matrices=mats, cameras=cams, fov=fb["fov"], floor=np.zeros(4, np.float32),
                   shading=ng.shading_arrays(sc.instances, [], ng.Plane(), [light], 0.2, None), spp=1, frame=0)
        ref = forward(cfg, oracle, outputs=("depth", "normal", "id", "rgb"))
        rep = compare_images(dict(depth=got["depth"], object_id=got["object_ids"]), ref, f"falling bodies frame {f}")
        assert rep["hits"] > 0.03 * rep["pixels"]
        diff = np.abs(got["rgb"].astype(np.int32) - ref["rgb"].astype(np.int32))
        assert float(diff.mean()) < 0.05, float(diff.mean())


# Newton's GeoType names with codes of this test's own (Newton is not installed; the adapter reads the names).
_GEO = SimpleNamespace(PLANE=0, SPHERE=2, CAPSULE=3, ELLIPSOID=4, CYLINDER=5, BOX=6, MESH=7, CONE=9, CONVEX_MESH=10)


def _newton_model(worlds=3, break_world=None):
    """The Newton Model fields the adapter reads: a ground plane in no world, then per world a sphere on body 0, a box
    and a capsule on body 1 (with an offset and a turn) and an invisible box."""
    q = np.array([0.0, 0.38268343, 0.0, 0.92387953])
    types, scales, flags, sworld, sbody, tfs, colors = [0], [[0.0, 0.0, 0.0]], [1], [-1], [-1], [[0, 0, -0.3, 0, 0, 0, 1]], [[0.3] * 3]
    for w in range(worlds):
        box = [0.1, 0.2, 0.3] if w != break_world else [0.1, 0.2, 0.35]
        for t, sc, fl, b, tf, c in [(2, [0.2, 0, 0], 1, 0, [0, 0, 0.1, 0, 0, 0, 1], [0.9, 0.1, 0.1]),
                                    (6, box, 1, 1, [0.1, 0, 0, *q], [0.1, 0.9, 0.1]),
                                    (3, [0.05, 0.2, 0], 1, 1, [0, 0.2, 0, 0, 0, 0, 1], [0.1, 0.1, 0.9]),
                                    (6, [1, 1, 1], 0, 1, [0, 0, 0, 0, 0, 0, 1], [1, 1, 1])]:
            types.append(t); scales.append(sc); flags.append(fl); sworld.append(w); sbody.append(2 * w + b); tfs.append(tf)
            colors.append(c)
    rng = np.random.default_rng(4)
    qb = rng.normal(size=(2 * worlds, 4))
    body_q = np.concatenate([rng.uniform(-0.5, 0.5, (2 * worlds, 3)) + [0, 0, -3.0], qb / np.linalg.norm(qb, axis=1, keepdims=True)], 1)
    return SimpleNamespace(world_count=worlds, body_count=2 * worlds, body_world=np.repeat(np.arange(worlds), 2),
                           shape_type=np.array(types), shape_scale=np.array(scales, np.float32), shape_flags=np.array(flags),
                           shape_world=np.array(sworld), shape_body=np.array(sbody), shape_transform=np.array(tfs, np.float32),
                           shape_color=np.array(colors, np.float32), shape_source=[None] * len(types), up_axis=2,
                           body_q=body_q.astype(np.float32))


def test_newton_models_become_a_template_on_bodies():
    shp, per_world = newton_shapes(_newton_model(), _GEO, 1)
    assert per_world == 2
    assert [s.kind for s in shp] == ["plane", "sphere", "box", "capsule"]
    assert [s.body for s in shp] == [-1, 0, 1, 1]
    assert shp[2].size == pytest.approx((0.1, 0.2, 0.3)) and shp[2].position == pytest.approx((0.1, 0.0, 0.0))
    assert np.allclose(shp[1].color, (0.9, 0.1, 0.1))
    with pytest.raises(ValueError, match="world 2 differs"):
        newton_shapes(_newton_model(break_world=2), _GEO, 1)
    m = _newton_model()
    m.body_world = np.array([0, 1, 0, 1, 2, 2])
    with pytest.raises(ValueError, match="world by world"):
        newton_shapes(m, _GEO, 1)


@pytest.mark.gpu
def test_a_newton_batch_renders_as_the_oracle(gpu, oracle):
    """NewtonBatchRenderer with state.body_q as a GPU tensor of all worlds' bodies, against the oracle's FP64 body
    composition and render of the same template and poses."""
    import torch
    from tests.helpers import compare_images
    model = _newton_model()
    view = NewtonBatchRenderer(model, 40, 32, geo_types=_GEO, visible_flag=1)
    cams = view.default_cameras().reshape(-1, 12)
    state = SimpleNamespace(body_q=torch.from_numpy(model.body_q).to(gpu))
    out = view.render(state, cams, rgb=False, depth=True, normals=True, object_ids=True)
    sc = view.scene
    cfg = dict(width=40, height=32, worlds=3, cameras_per_world=1, meshes=sc.meshes, instances=sc.instances, counts=None,
               mode="bodies", bodies=2, body_of_instance=sc.body_of_instance, local_pose=sc.local_pose, scale=sc.scale,
               body_poses=model.body_q.reshape(3, 2, 7), cameras=cams, fov=1.0)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "newton batch")
    assert rep["hits"] > 0.1 * rep["pixels"]


def test_the_falling_bodies_scene_is_the_reference_benchmarks():
    """bench/scene.py's layout: rows of 8 bodies 0.3 apart, even ones 320-triangle balls of radius 0.12 and odd ones
    boxes of half-size 0.1, falling from 0.3 + 0.02 i (+ up to 0.08) to rest on the ground; rotations stay rotations."""
    from neraulgraphics_amd import workloads
    fb = workloads.falling_bodies(12, worlds=5, frames=24)
    shp, geoms = mujoco_shapes(fb["model"])
    assert geoms.tolist() == list(range(12))
    sc = build_scene(shp)
    assert [m.triangle_count for m in sc.meshes] == [320, 12]
    assert np.isclose(np.linalg.norm(sc.meshes[0].vertices, axis=1).max(), 0.12, atol=1e-6)
    assert np.isclose(np.abs(sc.meshes[1].vertices).max(), 0.1)
    x, z = fb["geom_xpos"][..., 0], fb["geom_xpos"][..., 2]
    assert np.allclose(x[0, :, :8].mean(axis=0), (np.arange(8) - 3.5) * 0.3, atol=0.04)
    rest = np.where(np.arange(12) % 2 == 0, 0.12, 0.1)
    assert np.all(z >= rest - 1e-6) and np.allclose(z[-1], rest)              # at rest by the last frame
    assert np.all(z[0] <= 0.3 + 0.02 * np.arange(12) + 0.08 + 1e-6)
    r = fb["geom_xmat"].reshape(-1, 3, 3).astype(np.float64)
    assert np.allclose(r @ np.swapaxes(r, 1, 2), np.eye(3), atol=1e-6) and np.allclose(np.linalg.det(r), 1.0, atol=1e-6)
    cam = fb["camera"]
    assert np.allclose(cam[:, 3], (0.0, -1.9, 1.1)) and np.isclose(fb["fov"], np.radians(45.0))


@pytest.mark.gpu
def test_geom_frame_steps_replay_from_graphs(gpu):
    """The MuJoCo Warp step (geom frames into the top-level build, then the RGB picture and depth from persistent
    launches) captured in HIP graphs, one per frame, replayed in turn with direct steps between the replays: every
    replay gives the images the same step gives launched directly."""
    import torch
    from neraulgraphics_amd import workloads
    fb = workloads.falling_bodies(8, worlds=16, frames=3)
    sim = MJWarpBatchRenderer(fb["model"], 16, 32, 32, lights=[ng.Light(fb["light"], radius=0.0)], fov=fb["fov"])
    cams = torch.from_numpy(np.tile(fb["camera"].reshape(1, 12), (16, 1))).to(gpu)
    frames = [SimpleNamespace(geom_xpos=torch.from_numpy(fb["geom_xpos"][f]).to(gpu),
                              geom_xmat=torch.from_numpy(fb["geom_xmat"][f]).to(gpu)) for f in range(3)]

    def step(k):
        return sim.render(frames[k], cams, rgb=True, depth=True, zero_copy=True)

    want = []
    for k in range(3):
        out = step(k)
        torch.cuda.synchronize()
        want.append({n: out[n].clone() for n in ("rgb", "depth")})
    assert not torch.equal(want[0]["depth"], want[2]["depth"])
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(0)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graphs = []
    for k in range(3):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = step(k)
        graphs.append((g, out))
    torch.cuda.synchronize()
    for rep in range(3):
        for k in (2, 0, 1):
            g, out = graphs[k]
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out["depth"], want[k]["depth"]), (rep, k)
            assert torch.equal(out["rgb"], want[k]["rgb"]), (rep, k)
            step((k + 1) % 3)
            torch.cuda.synchronize()
