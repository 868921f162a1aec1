"""torch.autograd for depth and LiDAR range with respect to body poses.

Forward: the poses place every instance (one kernel rebuilds every world's top-level tree) and the batch is
rendered. Backward: a hand-written kernel re-traces each pixel (bounded by its forward distance, so only
the boxes in front of the hit are visited), takes the fixed hit triangle and returns
dL/d(position, quaternion) of every body; hits are recomputed, not stored.

Poses, cameras and rays may be given on the host (NumPy or CPU tensors) or on the GPU: the forward keeps device
copies of what the backward re-reads, and the gradient comes back on the poses' own device."""
import torch


def _on_device(a, device):
    return torch.as_tensor(a, dtype=torch.float32, device=device).contiguous()


class _PoseDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, body_poses, renderer, cameras, fov):
        dev = torch.device("cuda", renderer.device_id)
        poses = _on_device(body_poses.detach(), dev)
        cams = _on_device(cameras, dev)
        renderer.set_body_poses(poses)
        depth = renderer.render(cams, fov, depth=True, zero_copy=True)["depth"].clone()
        ctx.renderer, ctx.fov, ctx.pose_device = renderer, fov, body_poses.device
        ctx.save_for_backward(poses, cams, depth)
        return depth

    @staticmethod
    def backward(ctx, grad_depth):
        poses, cams, depth = ctx.saved_tensors
        grad = ctx.renderer.pose_backward_depth(poses, cams, ctx.fov, depth, _on_device(grad_depth, poses.device))
        return grad.to(ctx.pose_device), None, None, None


class _PoseRange(torch.autograd.Function):
    @staticmethod
    def forward(ctx, body_poses, renderer, origins, directions, max_distance):
        dev = torch.device("cuda", renderer.device_id)
        poses = _on_device(body_poses.detach(), dev)
        o, d = _on_device(origins, dev), _on_device(directions, dev)
        renderer.set_body_poses(poses)
        dist = renderer.cast_rays(o, d, max_distance, zero_copy=True)["distance"].clone()
        ctx.renderer, ctx.max_distance, ctx.pose_device = renderer, max_distance, body_poses.device
        ctx.save_for_backward(poses, o, d, dist)
        return dist

    @staticmethod
    def backward(ctx, grad_distance):
        poses, o, d, dist = ctx.saved_tensors
        grad = ctx.renderer.pose_backward_range(poses, o, d, ctx.max_distance, dist, _on_device(grad_distance, poses.device))
        return grad.to(ctx.pose_device), None, None, None, None


def render_depth(renderer, body_poses, cameras, fov):
    """Depth (views, H, W) of the batch placed by body_poses (world_count, bodies_per_world, 7), differentiable
    with respect to body_poses."""
    return _PoseDepth.apply(body_poses, renderer, cameras, fov)


def cast_range(renderer, body_poses, origins, directions, max_distance=float("inf")):
    """Ray distances (world_count, rays_per_world), differentiable with respect to body_poses."""
    return _PoseRange.apply(body_poses, renderer, origins, directions, max_distance)
