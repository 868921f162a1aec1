"""torch.autograd for depth and LiDAR range with respect to body poses.

Forward: the poses place every instance (one kernel rebuilds every world's top-level tree) and the batch is
rendered. Backward: a hand-written kernel re-traces each pixel (bounded by its forward distance, so only
the boxes in front of the hit are visited), takes the fixed hit triangle and returns
dL/d(position, quaternion) of every body; hits are recomputed, not stored."""
import torch


class _PoseDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, body_poses, renderer, cameras, fov):
        renderer.set_body_poses(body_poses)
        depth = renderer.render(cameras, fov, depth=True, zero_copy=True)["depth"].clone()
        ctx.renderer, ctx.fov = renderer, fov
        ctx.save_for_backward(body_poses, cameras, depth)
        return depth

    @staticmethod
    def backward(ctx, grad_depth):
        body_poses, cameras, depth = ctx.saved_tensors
        grad = ctx.renderer.pose_backward_depth(body_poses, cameras, ctx.fov, depth, grad_depth)
        return grad, None, None, None


class _PoseRange(torch.autograd.Function):
    @staticmethod
    def forward(ctx, body_poses, renderer, origins, directions, max_distance):
        renderer.set_body_poses(body_poses)
        dist = renderer.cast_rays(origins, directions, max_distance, zero_copy=True)["distance"].clone()
        ctx.renderer, ctx.max_distance = renderer, max_distance
        ctx.save_for_backward(body_poses, origins, directions, dist)
        return dist

    @staticmethod
    def backward(ctx, grad_distance):
        body_poses, origins, directions, dist = ctx.saved_tensors
        grad = ctx.renderer.pose_backward_range(body_poses, origins, directions, ctx.max_distance, dist, grad_distance)
        return grad, None, None, None, None


def render_depth(renderer, body_poses, cameras, fov):
    """Depth (views, H, W) of the batch placed by body_poses (world_count, bodies_per_world, 7), differentiable
    with respect to body_poses."""
    return _PoseDepth.apply(body_poses, renderer, cameras, fov)


def cast_range(renderer, body_poses, origins, directions, max_distance=float("inf")):
    """Ray distances (world_count, rays_per_world), differentiable with respect to body_poses."""
    return _PoseRange.apply(body_poses, renderer, origins, directions, max_distance)
