"""Deforming meshes (ngx_set_mesh_vertices, no counterpart in the reference): a mesh's bottom level refitted on the
GPU to new vertices, keeping its tree. The closest hit does not depend on the tree, so images after a refit must
match the oracle's, which builds the deformed mesh's tree afresh (equal distances; IDs equal but for ties); and a
refit back to the original vertices must give back the original images bit for bit, since a refit computes every
box as the host build does."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import workloads
from oracle.oracle import forward

from .helpers import compare_images, compare_rays

pytestmark = pytest.mark.gpu


def _wave(mesh, amplitude=0.4):
    v = mesh.vertices.copy()
    v[:, 2] += amplitude * np.sin(3.0 * v[:, 0]) * np.cos(2.0 * v[:, 1])
    return v


@pytest.mark.parametrize("walk", [dict(packet=True, lane_min_triangles=1 << 30), dict(packet=True, lane_min_triangles=0),
                                  dict(packet=False)], ids=["packet", "hybrid", "lanes"])
def test_a_refitted_mesh_renders_as_the_oracle_builds_it(gpu, oracle, walk):
    import torch
    cfg = workloads.make("instanced_worlds", worlds=8)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**walk)
    before = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    moved = _wave(cfg["meshes"][2])
    r.set_mesh_vertices(2, torch.from_numpy(moved).cuda())
    with pytest.raises(ng._lib.NgxError, match="mesh moved"):
        r.render(cfg["cameras"], cfg["fov"], depth=True)
    r.set_instance_transforms(cfg["matrices"])
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    deformed = dict(cfg)
    deformed["meshes"] = [cfg["meshes"][0], cfg["meshes"][1], ng.Mesh(moved, cfg["meshes"][2].indices)]
    ref = forward(deformed, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "deformed")
    assert rep["hits"] > 0.2 * rep["pixels"]
    assert not np.array_equal(out["depth"], before["depth"])
    # back to the original vertices (from the host this time): the original images, bit for bit
    r.set_mesh_vertices(2, cfg["meshes"][2].vertices)
    r.set_instance_transforms(cfg["matrices"])
    again = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    for k in ("depth", "normals", "object_ids"):
        assert np.array_equal(again[k].view(np.uint32) if again[k].dtype == np.float32 else again[k],
                              before[k].view(np.uint32) if before[k].dtype == np.float32 else before[k]), k


def test_a_refitted_lidar_mesh_answers_rays_as_the_oracle(gpu, oracle):
    cfg = workloads.make("lidar_sweeps", worlds=4, rays_per_world=4096)
    r = workloads.make_renderer(cfg)
    moved = _wave(cfg["meshes"][0], 0.2)
    r.set_mesh_vertices(0, moved)
    r.set_instance_transforms(cfg["matrices"])
    out = r.cast_rays(cfg["origins"], cfg["directions"], cfg["max_distance"])
    deformed = dict(cfg)
    deformed["meshes"] = [ng.Mesh(moved, cfg["meshes"][0].indices)]
    st = forward(deformed, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], cfg["origins"], cfg["directions"], cfg["max_distance"])
    rep = compare_rays(out, ref, "deformed lidar")
    assert rep["hits"] > 0


def test_refits_refuse_bad_input(gpu):
    cfg = workloads.make("instanced_worlds", worlds=2)
    r = workloads.make_renderer(cfg)
    with pytest.raises(ValueError, match="not one of"):
        r.set_mesh_vertices(3, cfg["meshes"][0].vertices)
    bad = cfg["meshes"][1].vertices.copy()
    bad[0, 0] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        r.set_mesh_vertices(1, bad)
    with pytest.raises(ValueError, match="values"):
        r.set_mesh_vertices(1, bad[:-1])
    r.render(cfg["cameras"], cfg["fov"], depth=True)  # nothing moved: still renders
