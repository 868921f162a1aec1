"""Shape generators and pose math for feeding simulators to the renderer, with the names and conventions of the
reference's python/neraulgraphics/shapes.py: meshes along the Z axis (the MuJoCo and Newton convention), quaternions
as (x, y, z, w), poses as (N, 3) positions and (N, 4) quaternions, matrices as the (N, 3, 4) float32 arrays that
BatchRenderer.set_instance_transforms takes. Cylinders and cones are made at unit size and scaled per instance;
a capsule's round caps need a mesh per size. Pose math is done in float64."""
from __future__ import annotations

import numpy as np

from .geometry import Mesh

__all__ = ["capsule", "compose_poses", "cone", "cylinder", "look_at", "matrices_from_rotations", "plane", "pose_matrices",
           "quat_to_matrix", "uv_sphere"]


def _revolve(profile, segments: int) -> Mesh:
    """A closed surface of revolution about Z from a list of (radius, z) rows, top to bottom. A row of radius 0 is one
    vertex on the axis (a pole or the centre of a cap); any other row is a ring of `segments` vertices. Neighbouring
    rings are joined by quads of two triangles, a ring and a point by a fan."""
    if segments < 3:
        raise ValueError(f"a surface of revolution needs 3 or more segments around its axis, not {segments}")
    phi = 2.0 * np.pi * np.arange(segments) / segments
    ring_xy = np.stack([np.cos(phi), np.sin(phi)], axis=1)
    verts, first = [], []
    for radius, z in profile:
        first.append(sum(len(v) for v in verts))
        if radius == 0.0:
            verts.append(np.array([[0.0, 0.0, z]]))
        else:
            verts.append(np.concatenate([radius * ring_xy, np.full((segments, 1), z)], axis=1))
    tris = []
    nxt = np.roll(np.arange(segments), -1)
    cur = np.arange(segments)
    for row in range(len(profile) - 1):
        top_pt, bot_pt = profile[row][0] == 0.0, profile[row + 1][0] == 0.0
        a, b = first[row], first[row + 1]
        if top_pt and bot_pt:
            continue
        if top_pt:
            tris.append(np.stack([np.full(segments, a), b + cur, b + nxt], axis=1))
        elif bot_pt:
            tris.append(np.stack([a + cur, a + nxt, np.full(segments, b)], axis=1))
        else:
            tris.append(np.stack([a + cur, a + nxt, b + nxt], axis=1))
            tris.append(np.stack([a + cur, b + nxt, b + cur], axis=1))
    return Mesh(np.concatenate(verts).astype(np.float32), np.concatenate(tris).astype(np.uint32))


def uv_sphere(radius: float = 1.0, segments: int = 32, rings: int = 16) -> Mesh:
    """A latitude/longitude sphere about the origin, poles on Z."""
    lat = np.linspace(0.5 * np.pi, -0.5 * np.pi, rings + 1)
    profile = [(radius * np.cos(t), radius * np.sin(t)) for t in lat]
    profile[0], profile[-1] = (0.0, radius), (0.0, -radius)
    return _revolve(profile, segments)


def capsule(radius: float, half_height: float, segments: int = 32, rings: int = 8) -> Mesh:
    """A cylinder of half-height `half_height` between two hemispheres of `radius`: 2 (half_height + radius) long
    along Z, as MuJoCo and Newton size capsules."""
    up = np.linspace(0.5 * np.pi, 0.0, rings + 1)
    down = np.linspace(0.0, -0.5 * np.pi, rings + 1)
    profile = [(radius * np.cos(t), half_height + radius * np.sin(t)) for t in up]
    profile += [(radius * np.cos(t), -half_height + radius * np.sin(t)) for t in down]
    profile[0], profile[-1] = (0.0, half_height + radius), (0.0, -half_height - radius)
    return _revolve(profile, segments)


def cylinder(radius: float = 1.0, half_height: float = 1.0, segments: int = 32) -> Mesh:
    """A capped cylinder along Z; at unit size, scale an instance to any size."""
    return _revolve([(0.0, half_height), (radius, half_height), (radius, -half_height), (0.0, -half_height)], segments)


def cone(radius: float = 1.0, half_height: float = 1.0, segments: int = 32) -> Mesh:
    """A cone along Z, apex at +half_height, flat base at -half_height."""
    return _revolve([(0.0, half_height), (radius, -half_height), (0.0, -half_height)], segments)


def plane(width: float, length: float, tile: float | None = None) -> Mesh:
    """A rectangle in the XY plane, `width` along X and `length` along Y (full sizes), two triangles. `tile` sets the
    texture repeat in the reference; this backend draws no textures, so it is accepted and has no effect."""
    x, y = 0.5 * width, 0.5 * length
    return Mesh(np.array([[-x, -y, 0.0], [x, -y, 0.0], [x, y, 0.0], [-x, y, 0.0]], np.float32),
                np.array([0, 1, 2, 0, 2, 3], np.uint32))


def look_at(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """(N, 3, 4) camera-to-world matrices (OpenGL: x right, y up, looking along -z) of cameras at `eye` looking at
    `target`; both (3,) or (N, 3), broadcast together. `up` is the world's up axis (Z for the simulators).

    The frame is built from the camera's back axis b (from the target to the eye): right = up x b and the camera's
    up = b x right, both unit length; the columns are right, up, b and the eye."""
    eye, target = np.broadcast_arrays(np.atleast_2d(np.asarray(eye, np.float64)), np.atleast_2d(np.asarray(target, np.float64)))
    back = eye - target
    back = back / np.linalg.norm(back, axis=1, keepdims=True)
    side = np.cross(np.broadcast_to(np.asarray(up, np.float64), back.shape), back)
    length = np.linalg.norm(side, axis=1)
    bad = np.flatnonzero(length < 1e-9)
    if bad.size:
        raise ValueError(f"look_at: camera {int(bad[0])} views along the up axis, which leaves its horizontal unset; "
                         "pick another eye, target or up")
    side = side / length[:, None]
    frame = np.empty((len(back), 3, 4))
    frame[:, :, 0] = side
    frame[:, :, 1] = np.cross(back, side)
    frame[:, :, 2] = back
    frame[:, :, 3] = eye
    return frame.astype(np.float32)


def quat_to_matrix(q) -> np.ndarray:
    """(N, 4) quaternions (x, y, z, w), normalized first, as (N, 3, 3) rotation matrices."""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], axis=1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], axis=1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], axis=1),
    ], axis=1)


def compose_poses(p1, q1, p2, q2):
    """The pose (p1, q1) applied after (p2, q2): position p1 + R(q1) p2 and quaternion q1 q2 (Hamilton product)."""
    p1, q1, p2, q2 = (np.asarray(a, np.float64) for a in (p1, q1, p2, q2))
    pos = p1 + np.einsum("nij,nj->ni", quat_to_matrix(q1), p2)
    v1, w1 = q1[:, :3], q1[:, 3:]
    v2, w2 = q2[:, :3], q2[:, 3:]
    vec = w1 * v2 + w2 * v1 + np.cross(v1, v2)
    w = w1 * w2 - np.sum(v1 * v2, axis=1, keepdims=True)
    return pos, np.concatenate([vec, w], axis=1)


def matrices_from_rotations(rotation, position, scale=None) -> np.ndarray:
    """(N, 3, 4) float32 matrices [R diag(scale) | p] from (N, 3, 3) rotations (as MuJoCo reports them)."""
    rotation = np.asarray(rotation, np.float64)
    out = np.empty((len(rotation), 3, 4), np.float32)
    out[:, :, :3] = rotation if scale is None else rotation * np.asarray(scale, np.float64)[:, None, :]
    out[:, :, 3] = np.asarray(position, np.float64)
    return out


def pose_matrices(position, quat, scale=None) -> np.ndarray:
    """(N, 3, 4) float32 matrices of poses; `scale` (N, 3) acts first, along the mesh's own axes."""
    return matrices_from_rotations(quat_to_matrix(quat), position, scale)
