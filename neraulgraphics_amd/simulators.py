"""A simulator's shapes and body poses to the batched renderer: the part the reference's simulator adapters
(python/neraulgraphics/newton.py:48-89, :309-348; python/neraulgraphics/_adapter.py:30-85) share, without
importing any simulator. A simulator describes its shapes once: kind, size, pose on its body, body, colour. This
module turns them into meshes (one per kind and size that needs its own, unit meshes scaled per instance
otherwise), instances with a matte material of the shape's colour, and a body binding; afterwards each step only
hands the bodies' poses (worlds, bodies, 7: x, y, z, qx, qy, qz, qw) to BatchRenderer.set_body_poses, one kernel.

Sizes follow Newton's shape_scale convention: sphere (radius), box (half-extents x, y, z), ellipsoid (semi-axes),
cylinder and cone (radius, half-height, along Z), capsule (radius, half-height of the straight part), plane (full
width and length, 0 for unbounded), mesh (scale along the mesh's own axes)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .batch import BatchRenderer
from .geometry import Instance, Material, Mesh, make_box, make_icosphere
from .shapes import capsule, cone, cylinder, plane

__all__ = ["Shape", "SimScene", "SimBatchRenderer", "build_scene"]

KINDS = ("sphere", "box", "ellipsoid", "cylinder", "cone", "capsule", "plane", "mesh")
UNBOUNDED_PLANE = 1000.0  # the size the Newton viewer gives a plane of size 0
_FALLBACK_COLOR = (0.62, 0.64, 0.70)


@dataclass
class Shape:
    kind: str
    size: tuple = (1.0, 1.0, 1.0)
    position: tuple = (0.0, 0.0, 0.0)        # pose on its body (or in the world when body is -1)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)   # x, y, z, w
    body: int = -1                           # index of the body within its world; -1: fixed in the world
    color: tuple = _FALLBACK_COLOR
    mesh: Mesh | None = None                 # for kind "mesh"


@dataclass
class SimScene:
    meshes: list = field(default_factory=list)
    instances: list = field(default_factory=list)
    body_of_instance: np.ndarray = None
    local_pose: np.ndarray = None
    scale: np.ndarray = None


def _material(color):
    rgb = tuple(float(np.clip(c, 0.0, 1.0)) for c in list(color)[:3])
    return Material(rgb, 0.9, 0.2, 0.0, 40.0)


def build_scene(shapes) -> SimScene:
    """Meshes, instances and the body binding for a world's shapes (every world holds the same shapes)."""
    scene = SimScene()
    cache = {}

    def mesh(key, make):
        if key not in cache:
            cache[key] = len(scene.meshes)
            scene.meshes.append(make())
        return cache[key]

    body, local, scales = [], [], []
    for s in shapes:
        if s.kind not in KINDS:
            raise ValueError(f"shape kind {s.kind!r} is not drawn; kinds are {', '.join(KINDS)}")
        z = [float(v) for v in (s.size if np.ndim(s.size) else (s.size,))]
        one = np.ones(3)
        if s.kind == "sphere":
            m, sc = mesh("unit_sphere", lambda: make_icosphere(1.0, 3)), one * z[0]
        elif s.kind == "box":
            m, sc = mesh("unit_box", lambda: make_box((1.0, 1.0, 1.0))), np.array(z[:3]) if len(z) >= 3 else one * z[0]
        elif s.kind == "ellipsoid":
            m, sc = mesh("unit_sphere", lambda: make_icosphere(1.0, 3)), np.array(z[:3])
        elif s.kind == "cylinder":
            m, sc = mesh("unit_cylinder", cylinder), np.array([z[0], z[0], z[1]])
        elif s.kind == "cone":
            m, sc = mesh("unit_cone", cone), np.array([z[0], z[0], z[1]])
        elif s.kind == "capsule":
            m, sc = mesh(("capsule", z[0], z[1]), lambda: capsule(z[0], z[1])), one
        elif s.kind == "plane":
            w = z[0] if z[0] > 0.0 else UNBOUNDED_PLANE
            length = z[1] if len(z) > 1 and z[1] > 0.0 else UNBOUNDED_PLANE
            m, sc = mesh(("plane", w, length), lambda: plane(w, length)), one
        else:
            if s.mesh is None:
                raise ValueError("a mesh shape needs its mesh")
            m, sc = mesh(("mesh", id(s.mesh)), lambda: s.mesh), np.array(z[:3]) if len(z) >= 3 else one * z[0]
        scene.instances.append(Instance(m, material=_material(s.color)))
        body.append(int(s.body))
        q = np.asarray(s.rotation, np.float64)
        local.append([*s.position, *(q / np.linalg.norm(q))])
        scales.append(sc)
    scene.body_of_instance = np.asarray(body, np.int32)
    scene.local_pose = np.asarray(local, np.float32).reshape(-1, 7)
    scene.scale = np.asarray(scales, np.float32).reshape(-1, 3)
    return scene


class SimBatchRenderer:
    """Renders every world of a batched simulation: built once from the shapes, then moved by body poses.

    The flow of the reference's NewtonBatchRenderer (python/neraulgraphics/newton.py:309-348): bind the bodies once,
    then per step set_body_poses(body_q) and render. body_q may be a NumPy array or a GPU tensor, used in place."""

    def __init__(self, shapes, world_count, bodies_per_world, width, height, cameras_per_world=1, lights=(), floor=None,
                 ambient=0.2, device_id=None):
        self.scene = build_scene(shapes)
        self.renderer = BatchRenderer(width, height, world_count, cameras_per_world, lights=list(lights), floor=floor,
                                      meshes=self.scene.meshes, instances=self.scene.instances, device_id=device_id)
        self.renderer.set_ambient(ambient)
        self.renderer.bind_bodies(self.scene.body_of_instance, self.scene.local_pose, self.scene.scale, int(bodies_per_world))
        self._bodies = int(bodies_per_world)
        if self._bodies == 0:  # everything fixed: placed once
            self.renderer.set_body_poses(np.zeros((world_count, 0, 7), np.float32))

    def update(self, body_q) -> None:
        """Places every instance of every world from the bodies' poses (worlds, bodies, 7)."""
        if self._bodies:
            self.renderer.set_body_poses(body_q)

    def render(self, body_q, cameras, fov=1.0, samples_per_pixel=1, **outputs):
        """update(body_q), then BatchRenderer.render(cameras, fov, samples_per_pixel, **outputs)."""
        self.update(body_q)
        return self.renderer.render(cameras, fov, samples_per_pixel, **outputs)

    def cast_rays(self, origins, directions, max_distance=float("inf"), **kw):
        return self.renderer.cast_rays(origins, directions, max_distance, **kw)
