"""GPU parity of normals, spheres and the floor: the reference's own known answers (tests/test_renderer.cpp:722-775)
rendered through the C ABI, and meshes with vertex normals under shearing instance maps, against the oracle, whose
normals tests/test_oracle_normals.py pins."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd.geometry import Instance, Mesh, look_at, make_box, make_grid, make_icosphere
from oracle.oracle import forward

from .helpers import compare_images
from .test_oracle_normals import FOV, H, IDENTITY_CAMERA, W, _no_mesh_render, _quad, _random_affine, _scene

pytestmark = pytest.mark.gpu

WALKS = {"packet": dict(packet=True, lane_min_triangles=1 << 30), "hybrid": dict(packet=True, lane_min_triangles=0),
         "lanes": dict(packet=False)}


def _say(*a):
    print(*a, flush=True)


def _gpu_images(r, cams=IDENTITY_CAMERA, fov=FOV):
    out = r.render(cams, fov, depth=True, normals=True, object_ids=True)
    return dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"])


@pytest.mark.parametrize("walk", list(WALKS))
def test_the_references_known_answers_on_the_device(gpu, oracle, walk):
    _say(f"known answers, {walk} walk")
    mat = ng.Material((0.7, 0.5, 0.3), 0.9, 0.2, 0.0, 40.0)
    # a sphere 8 away at its nearest point (tests/test_renderer.cpp:739-749)
    r = ng.BatchRenderer(W, H, 1, spheres=[ng.Sphere((0.0, 0.0, -10.0), 2.0, mat)])
    r.set_trace_config(**WALKS[walk])
    g = _gpu_images(r)
    compare_images(g, _no_mesh_render(oracle, spheres=np.array([[0, 0, -10, 2]], np.float32)), "sphere")
    assert abs(g["depth"][0, H // 2, W // 2] - 8.0) < 0.02 and g["object_id"][0, H // 2, W // 2] == 1
    assert g["normal"][0, H // 2, W // 2, 2] > 0.99
    assert np.isposinf(g["depth"][0, 3, 3]) and g["object_id"][0, 3, 3] == 0 and not g["normal"][0, 3, 3].any()
    # a wall 5 away, wound either way (:751-758)
    for flip in (False, True):
        q = _quad(3.0)
        if flip:
            q.indices = q.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1).copy()
        r = ng.BatchRenderer(W, H, 1, meshes=[q], instances=[Instance(0)])
        r.set_trace_config(**WALKS[walk])
        g = _gpu_images(r)
        compare_images(g, forward(_scene([q], [np.eye(3, 4)]), oracle), f"wall, flipped {flip}")
        hit = np.isfinite(g["depth"][0])
        assert abs(g["depth"][0, H // 2 + 20, W // 2 + 40] - 5.0) < 1e-3 and g["object_id"][0, H // 2, W // 2] == 34
        assert np.all(g["normal"][0][hit] == [0.0, 0.0, 1.0])
    # the floor 2 below the eye (:760-768)
    floor = np.array([-2.0, 50.0, -100.0, 100.0], np.float32)
    r = ng.BatchRenderer(W, H, 1, floor=ng.Plane(y=-2.0, half_extent_x=50.0, z_min=-100.0, z_max=100.0))
    r.set_trace_config(**WALKS[walk])
    g = _gpu_images(r)
    compare_images(g, _no_mesh_render(oracle, floor=floor), "floor")
    row = H * 3 // 4
    expect = 2.0 / (2.0 * (row + 0.5) / H - 1.0)
    assert g["object_id"][0, row, W // 2] == 33 and abs(g["depth"][0, row, W // 2] - expect) < 0.01 * expect
    assert abs(g["normal"][0, row, W // 2, 1] - 1.0) < 1e-5
    # the nearer of two spheres (:770-775)
    r = ng.BatchRenderer(W, H, 1, spheres=[ng.Sphere((0.0, 0.0, -3.0), 1.0), ng.Sphere((0.0, 0.0, -8.0), 2.0)])
    r.set_trace_config(**WALKS[walk])
    g = _gpu_images(r)
    compare_images(g, _no_mesh_render(oracle, spheres=np.array([[0, 0, -3, 1], [0, 0, -8, 2]], np.float32)), "two spheres")
    assert g["object_id"][0, H // 2, W // 2] == 1 and abs(g["depth"][0, H // 2, W // 2] - 2.0) < 0.02


def _smooth_meshes():
    ico = make_icosphere(1.0, 2)
    grid = make_grid(40, 25, 2.0, 1.25, 0.05, seed=9)
    rng = np.random.default_rng(4)
    gn = np.array([0.0, 0.0, 1.0]) + rng.normal(0.0, 0.2, size=grid.vertices.shape)
    gn /= np.linalg.norm(gn, axis=1, keepdims=True)
    return [Mesh(ico.vertices, ico.indices, normals=ico.vertices.copy()), make_box((0.5, 0.5, 0.5)), make_icosphere(0.8, 1),
            Mesh(grid.vertices, grid.indices, normals=gn.astype(np.float32))]


@pytest.mark.parametrize("walk", list(WALKS))
def test_vertex_normals_and_sheared_instances_match_the_oracle(gpu, oracle, walk):
    """Four meshes (a unit icosphere with its vertices as normals, a box, an icosphere without normals, a 2000-triangle
    grid with tilted vertex normals) under maps that rotate, scale along oblique axes and translate; 6 worlds."""
    _say(f"vertex normals and sheared instances, {walk} walk")
    meshes = _smooth_meshes()
    worlds, per = 6, 24
    rng = np.random.default_rng(12)
    mats = _random_affine(rng, worlds * per, 0.5, 1.6).reshape(worlds, per, 12)
    centres = mats.reshape(worlds, per, 3, 4)[..., 3].mean(axis=1)
    cams = np.stack([look_at(rng.uniform(-1.0, 1.0, 3), c).reshape(12) for c in centres]).astype(np.float32)
    cfg = dict(width=96, height=64, worlds=worlds, cameras_per_world=1, meshes=meshes, instances=[Instance(i % 4) for i in range(per)],
               counts=None, mode="matrices", matrices=mats, cameras=cams, fov=1.0)
    r = ng.BatchRenderer(96, 64, worlds, meshes=meshes, instances=cfg["instances"])
    r.set_instance_transforms(mats)
    r.set_trace_config(**WALKS[walk])
    g = _gpu_images(r, cams, 1.0)
    ref = forward(cfg, oracle)
    rep = compare_images(g, ref, f"vertex normals, {walk}")
    _say(rep)
    assert rep["hits"] > 0.1 * rep["pixels"]
    # every mesh was seen, the smooth ones included
    seen_tmpl = np.unique(g["object_id"][g["object_id"] >= 34]) - 34
    assert {0, 1, 2, 3} <= set(int(t) % 4 for t in seen_tmpl)
    bits = np.mean(g["normal"].view(np.uint32)[np.isfinite(ref["depth"])] == ref["normal"].view(np.uint32)[np.isfinite(ref["depth"])])
    _say(f"normals bit-identical to the oracle's: {bits:.6f}")
