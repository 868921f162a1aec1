"""The allocation-free depth backward, ngx_pose_backward_depth_noalloc, called through the C ABI: its gradients
equal ngx_pose_backward_depth's bit for bit (the same kernels and the same fixed-order sums, only the partial
sums live elsewhere) on two calls in a row, the second on a smaller batch, so what the first left in the shared
array cannot leak into the second; a call whose partial sums would not fit is refused before anything changes."""
import ctypes

import numpy as np
import pytest

from neraulgraphics_amd import _lib, workloads
from neraulgraphics_amd.batch import BatchRenderer

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _renderer(name, worlds, cameras_per_world, seed):
    """A batch of the named body-pose workload with cameras_per_world cameras a world (the workload's camera, then
    copies whose eyes are moved a little)."""
    cfg = workloads.make(name, worlds=worlds, seed=seed)
    r = BatchRenderer(cfg["width"], cfg["height"], worlds, cameras_per_world, meshes=cfg["meshes"], instances=cfg["instances"],
                      instances_in_world=cfg["counts"])
    r.bind_bodies(cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["bodies"])
    r.set_body_poses(cfg["body_poses"])
    rng = np.random.default_rng(seed)
    cams = np.repeat(cfg["cameras"].reshape(worlds, 1, 12), cameras_per_world, axis=1).copy()
    cams[:, 1:, [3, 7, 11]] += rng.normal(0.0, 0.2, (worlds, cameras_per_world - 1, 3)).astype(np.float32)
    return r, cfg, cams.reshape(-1, 12).astype(np.float32)


def _gradients(r, cfg, cams, seed):
    """(allocation-free, allocating) gradients of one random dL/d depth, the allocation-free entry called first."""
    import torch
    dev = torch.device("cuda", 0)
    poses = torch.from_numpy(cfg["body_poses"]).to(dev).contiguous()
    c = torch.from_numpy(cams).to(dev).contiguous()
    r.set_body_poses(poses)
    depth = r.render(c, cfg["fov"], depth=True, zero_copy=True)["depth"].clone()
    g = torch.from_numpy(np.random.default_rng(seed).normal(size=tuple(depth.shape)).astype(np.float32)).to(dev)
    out_free = torch.full((r.world_count, cfg["bodies"], 7), float("nan"), device=dev)
    out_alloc = torch.full_like(out_free, float("nan"))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(r._lib.ngx_pose_backward_depth_noalloc(r._h, poses.data_ptr(), c.data_ptr(), float(cfg["fov"]),
                                                      depth.data_ptr(), g.data_ptr(), out_free.data_ptr(), stream))
    _lib.check(r._lib.ngx_pose_backward_depth(r._h, poses.data_ptr(), c.data_ptr(), float(cfg["fov"]), depth.data_ptr(),
                                              g.data_ptr(), out_alloc.data_ptr(), stream))
    torch.cuda.synchronize()
    return out_free.cpu().numpy(), out_alloc.cpu().numpy()


def test_no_alloc_backward_equals_the_allocating_one_on_two_calls(gpu):
    # first call: 96 worlds of 32 bodies at 64x64, 3 cameras a world (288 views x 8 parts x 32 bodies x 7 floats)
    _say("no-alloc backward: posed_bodies, 96 worlds x 3 cameras")
    r, cfg, cams = _renderer("posed_bodies", 96, 3, seed=31)
    free, alloc = _gradients(r, cfg, cams, seed=32)
    assert np.isfinite(alloc).all() and np.abs(alloc).sum() > 0
    assert np.array_equal(free.view(np.uint32), alloc.view(np.uint32))
    del r
    # second call, smaller: 8 worlds, 1 camera; the array still holds the first call's partial sums
    _say("no-alloc backward: posed_bodies, 8 worlds x 1 camera")
    r, cfg, cams = _renderer("posed_bodies", 8, 1, seed=33)
    free, alloc = _gradients(r, cfg, cams, seed=34)
    assert np.isfinite(alloc).all() and np.abs(alloc).sum() > 0
    assert np.array_equal(free.view(np.uint32), alloc.view(np.uint32))


def test_no_alloc_backward_refuses_what_does_not_fit(gpu):
    import torch
    # 1200 worlds at 128x128 (16 parts a view) of 64 bodies: 1200 x 16 x 64 x 7 = 8,601,600 floats > 8,388,608
    _say("no-alloc backward: body_worlds, 1200 worlds (refused)")
    r, cfg, cams = _renderer("body_worlds", 1200, 1, seed=35)
    dev = torch.device("cuda", 0)
    poses = torch.from_numpy(cfg["body_poses"]).to(dev).contiguous()
    c = torch.from_numpy(cams).to(dev).contiguous()
    pixels = torch.zeros((r.view_count, 128, 128), device=dev)
    out = torch.full((1200, cfg["bodies"], 7), 7.0, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    status = r._lib.ngx_pose_backward_depth_noalloc(r._h, poses.data_ptr(), c.data_ptr(), float(cfg["fov"]), pixels.data_ptr(),
                                                    pixels.data_ptr(), out.data_ptr(), stream)
    assert status == _lib.NGX_INVALID_ARGUMENT
    assert "8601600" in r._lib.ngx_last_error().decode()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
