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
from .geometry import OBJECT_INSTANCE, Instance, Light, Material, Mesh, make_box, make_icosphere
from .shapes import capsule, compose_poses, cone, cylinder, look_at, plane

__all__ = ["Shape", "SimScene", "SimBatchRenderer", "MJWarpBatchRenderer", "NewtonBatchRenderer", "build_scene", "mujoco_shapes",
           "newton_shapes", "scene_bounds"]

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


# MuJoCo's geom types (mjtGeom), the codes MjModel.geom_type holds.
MJ_PLANE, MJ_HFIELD, MJ_SPHERE, MJ_CAPSULE, MJ_ELLIPSOID, MJ_CYLINDER, MJ_BOX, MJ_MESH = range(8)
_MJ_DEFAULT_RGBA = np.array([0.5, 0.5, 0.5, 1.0], np.float32)


def mujoco_shapes(model, geom_groups=(0, 1, 2)):
    """The drawn geoms of a MuJoCo model as Shapes, and their geom indices. Reads only the model's arrays (ngeom,
    geom_type, geom_size, geom_rgba, geom_group, geom_matid, mat_rgba, geom_dataid and the mesh_* arrays), so an
    MjModel or anything holding the same fields will do. The selection and sizes of the reference's MuJoCo adapter
    (python/neraulgraphics/mujoco.py:68-88, :104-136): geoms of the groups asked for with a visible colour; planes,
    spheres, capsules, ellipsoids, cylinders, boxes and meshes (heightfields and SDFs are not drawn); a geom with a
    material and the default colour takes its material's colour. Planes have their colour, not the reference's
    checkerboard texture (textures are not drawn by this backend)."""
    shapes, geoms, meshes = [], [], {}
    for g in range(int(model.ngeom)):
        kind = int(model.geom_type[g])
        rgba = np.asarray(model.geom_rgba[g], np.float32)
        if int(model.geom_group[g]) not in geom_groups or rgba[3] <= 0.0:
            continue
        size = [float(v) for v in model.geom_size[g]]
        if kind == MJ_SPHERE:
            shape = Shape("sphere", (size[0],))
        elif kind == MJ_ELLIPSOID:
            shape = Shape("ellipsoid", tuple(size[:3]))
        elif kind == MJ_BOX:
            shape = Shape("box", tuple(size[:3]))
        elif kind == MJ_CYLINDER:
            shape = Shape("cylinder", (size[0], size[1]))
        elif kind == MJ_CAPSULE:
            shape = Shape("capsule", (size[0], size[1]))
        elif kind == MJ_PLANE:  # MuJoCo sizes a plane by its half-extents; 0 extends it forever
            shape = Shape("plane", (2.0 * size[0] if size[0] > 0.0 else 0.0, 2.0 * size[1] if size[1] > 0.0 else 0.0))
        elif kind == MJ_MESH:
            m = int(model.geom_dataid[g])
            if m not in meshes:  # one Mesh per MuJoCo mesh, shared by the geoms that draw it
                v0, nv = int(model.mesh_vertadr[m]), int(model.mesh_vertnum[m])
                f0, nf = int(model.mesh_faceadr[m]), int(model.mesh_facenum[m])
                meshes[m] = Mesh(np.asarray(model.mesh_vert[v0:v0 + nv], np.float32),
                                 np.asarray(model.mesh_face[f0:f0 + nf], np.uint32))
            shape = Shape("mesh", (1.0, 1.0, 1.0), mesh=meshes[m])
        else:
            continue
        matid = int(model.geom_matid[g])
        if matid >= 0 and np.allclose(rgba, _MJ_DEFAULT_RGBA):
            rgba = np.asarray(model.mat_rgba[matid], np.float32)
        shape.color = tuple(float(c) for c in rgba[:3])
        shapes.append(shape)
        geoms.append(g)
    return shapes, np.asarray(geoms, np.int32)


def _reach(shape: Shape, scale) -> float:
    """Radius about the shape's origin that holds it (planes: 0, they are the world)."""
    if shape.kind == "plane":
        return 0.0
    if shape.kind == "capsule":
        return float(shape.size[0] + shape.size[1])
    if shape.kind == "mesh":
        return float(np.linalg.norm(shape.mesh.vertices, axis=1).max() * np.abs(scale).max())
    unit = {"sphere": 1.0, "ellipsoid": 1.0, "box": np.sqrt(3.0), "cylinder": np.sqrt(2.0), "cone": np.sqrt(2.0)}[shape.kind]
    return float(unit * np.abs(scale).max())


def scene_bounds(positions, reach):
    """Centre and half-size of spheres of radius `reach` at `positions` (none: a unit scene at the origin), as the
    reference's adapters frame a scene (python/neraulgraphics/_adapter.py:71-76)."""
    positions, reach = np.asarray(positions, np.float64).reshape(-1, 3), np.asarray(reach, np.float64).reshape(-1)
    if len(positions) == 0:
        return np.zeros(3), 1.0
    lo, hi = (positions - reach[:, None]).min(axis=0), (positions + reach[:, None]).max(axis=0)
    return (lo + hi) / 2.0, float(np.linalg.norm(hi - lo)) / 2.0


class MJWarpBatchRenderer:
    """Draws MuJoCo Warp worlds, thousands at once, each with its own cameras: the reference's MJWarpBatchRenderer
    (python/neraulgraphics/mujoco.py:177-323). The model's geoms become instances once; each render reads the
    data's geom_xpos (world_count, ngeom, 3) and geom_xmat (world_count, ngeom, 3, 3) as float32 GPU tensors (or
    anything torch.from_dlpack takes, or NumPy arrays) and places every geom of every world while rebuilding the top
    levels, one kernel (ngx_set_geom_frames); the reference composes the matrices in a Warp kernel first.

    The reference frames its default camera and lights from the model after mj_forward; here `data` at construction
    (anything with the model's geom_xpos, for one world or all) plays that part, and without it the frame is a
    unit scene at the origin."""

    def __init__(self, model, world_count, width=128, height=128, cameras_per_world=1, *, geom_groups=(0, 1, 2),
                 lights=None, fov=1.0, ambient=0.2, data=None, device_id=None):
        shapes, self._geoms = mujoco_shapes(model, geom_groups)
        if not shapes:
            raise ValueError("the model has no visible geoms to draw")
        self.scene = build_scene(shapes)
        self._worlds, self._cameras_per_world = int(world_count), int(cameras_per_world)
        self.width, self.height, self.fov = int(width), int(height), float(fov)
        self.up = np.array([0.0, 0.0, 1.0])
        reach = np.array([_reach(sh, sc) for sh, sc in zip(shapes, self.scene.scale)])
        if data is not None:
            xpos = np.asarray(_to_numpy(data.geom_xpos), np.float64).reshape(-1, int(model.ngeom), 3)[0][self._geoms]
            solid = reach > 0
            self.center, extent = scene_bounds(xpos[solid], reach[solid])
        else:
            self.center, extent = np.zeros(3), 1.0
        self.extent = max(extent, 1e-3)
        if lights is None:
            lights = self.default_lights()
        self.renderer = BatchRenderer(self.width, self.height, self._worlds, self._cameras_per_world, lights=list(lights),
                                      meshes=self.scene.meshes, instances=self.scene.instances, device_id=device_id)
        self.renderer.set_ambient(ambient)
        self.renderer.bind_geoms(self._geoms, self.scene.scale, int(model.ngeom))

    @property
    def view_count(self) -> int:
        return self._worlds * self._cameras_per_world

    def default_lights(self):
        """A key light high and to one side and a dimmer fill from the other, scaled to the scene, as the
        reference's adapters place them (python/neraulgraphics/_adapter.py:128-134), for Z up."""
        side, across = np.array([0.0, 1.0, 0.0]), np.array([-1.0, 0.0, 0.0])
        key = self.center + self.up * 3.0 * self.extent + side * 1.5 * self.extent + across * 1.0 * self.extent
        fill = self.center + self.up * 2.0 * self.extent - side * 2.0 * self.extent - across * 1.5 * self.extent
        return [Light(tuple(key), intensity=1.0, radius=0.0), Light(tuple(fill), intensity=0.4, radius=0.0)]

    def default_cameras(self):
        """One camera per view, framing the scene from the side, slightly above: (views, 3, 4)."""
        eye = self.center - np.array([0.0, 1.0, 0.0]) * 2.6 * self.extent + self.up * 0.7 * self.extent
        return np.tile(look_at(eye, self.center, self.up), (self.view_count, 1, 1))

    def update(self, data) -> None:
        """Places every geom of every world from data.geom_xpos and data.geom_xmat (step the simulation first)."""
        self.renderer.set_geom_frames(_as_frames(data.geom_xpos), _as_frames(data.geom_xmat))

    def render(self, data, cameras=None, *, fov=None, rgb=True, depth=False, normals=False, object_ids=False, zero_copy=False,
               frame=0, samples_per_pixel=1):
        """update(data), then the images asked for as a dict (see BatchRenderer.render). GPU cameras need zero_copy."""
        self.update(data)
        return self.renderer.render(self.default_cameras() if cameras is None else cameras, self.fov if fov is None else fov,
                                    samples_per_pixel, rgb=rgb, depth=depth, normals=normals, object_ids=object_ids,
                                    frame=frame, zero_copy=zero_copy)

    @property
    def object_id_of_geom(self) -> dict:
        """Which object ID each drawn MuJoCo geom index has in an object_ids image or a ray query."""
        return {int(g): OBJECT_INSTANCE + i for i, g in enumerate(self._geoms)}

    def cast_rays(self, origins, directions, max_distance=float("inf"), **kw):
        """Ray queries in every world: (world_count, rays_per_world, 3) arrays (see BatchRenderer.cast_rays)."""
        return self.renderer.cast_rays(origins, directions, max_distance, **kw)


def _as_frames(a):
    """A GPU tensor, a NumPy array, or anything torch.from_dlpack takes (a Warp array), as the renderer's input."""
    if isinstance(a, np.ndarray) or not hasattr(a, "__dlpack__"):
        return np.asarray(a, np.float32)
    import torch
    return a if isinstance(a, torch.Tensor) else torch.from_dlpack(a)


def _to_numpy(a):
    if isinstance(a, np.ndarray) or not hasattr(a, "__dlpack__"):
        return np.asarray(a)
    import torch
    t = a if isinstance(a, torch.Tensor) else torch.from_dlpack(a)
    return t.detach().cpu().numpy()


# ---- Newton

_NEWTON_UP = {0: (1.0, 0.0, 0.0), 1: (0.0, 1.0, 0.0), 2: (0.0, 0.0, 1.0)}


def _newton_enums(geo_types, visible_flag):
    """Newton's GeoType and the VISIBLE shape flag: from `newton` when it is importable, else as given."""
    if geo_types is None or visible_flag is None:
        try:
            import newton
        except ImportError as e:
            raise ImportError("newton is not installed: pass geo_types (an object with Newton's GeoType names as "
                              "attributes) and visible_flag") from e
        geo_types = newton.GeoType if geo_types is None else geo_types
        visible_flag = int(newton.ShapeFlags.VISIBLE) if visible_flag is None else visible_flag
    return geo_types, int(visible_flag)


def _host(a):
    """A Warp array (.numpy()), a tensor or a NumPy array as NumPy."""
    if hasattr(a, "numpy") and not isinstance(a, np.ndarray):
        a = a.detach().cpu() if hasattr(a, "detach") else a
        return np.asarray(a.numpy())
    return np.asarray(a)


def newton_shapes(model, geo_types=None, visible_flag=None):
    """One world's template of a batched Newton model, as the reference's NewtonBatchRenderer reads it
    (python/neraulgraphics/newton.py:250-296): the visible shapes that belong to no world (a ground plane), then those
    of world 0, each on its body's index within the world; every world must hold the same shapes on the same bodies,
    as Newton's replicate() makes them (checked). Returns (shapes, bodies_per_world). Reads world_count, body_count,
    body_world, shape_type, shape_scale, shape_flags, shape_world, shape_body, shape_transform (x, y, z, qx, qy, qz,
    qw), shape_color and shape_source. Sizes are Newton's shape_scale, the convention Shape takes; planes have their
    colour, not the reference's checkerboard texture."""
    geo, visible = _newton_enums(geo_types, visible_flag)
    worlds, bodies = int(model.world_count), int(model.body_count)
    if bodies % worlds:
        raise ValueError(f"the model has {bodies} bodies in {worlds} worlds, which is not the same number in each")
    per_world = bodies // worlds
    if bodies and not np.array_equal(_host(model.body_world), np.repeat(np.arange(worlds), per_world)):
        raise ValueError("the bodies are not stored world by world (all of world 0's, then world 1's, and so on)")
    shape_type, shape_scale = _host(model.shape_type), _host(model.shape_scale)
    flags, shape_world = _host(model.shape_flags), _host(model.shape_world)
    shape_body, transform = _host(model.shape_body), _host(model.shape_transform)
    colors = _host(model.shape_color) if getattr(model, "shape_color", None) is not None else None
    kinds = {int(geo.SPHERE): "sphere", int(geo.BOX): "box", int(geo.ELLIPSOID): "ellipsoid", int(geo.CYLINDER): "cylinder",
             int(geo.CONE): "cone", int(geo.CAPSULE): "capsule", int(geo.PLANE): "plane", int(geo.MESH): "mesh"}
    if hasattr(geo, "CONVEX_MESH"):
        kinds[int(geo.CONVEX_MESH)] = "mesh"
    entries, signatures, meshes = {}, {}, {}
    for i in range(len(shape_type)):
        if not int(flags[i]) & visible:
            continue
        kind = kinds.get(int(shape_type[i]))
        source = model.shape_source[i] if kind == "mesh" else None
        if kind is None or (kind == "mesh" and source is None):
            continue
        w, body = int(shape_world[i]), int(shape_body[i])
        local_body = body - w * per_world if (body >= 0 and w >= 0) else body
        color = tuple(float(c) for c in colors[i][:3]) if colors is not None else _FALLBACK_COLOR
        mesh = None
        if kind == "mesh":  # one Mesh per source, shared by the shapes that draw it
            if id(source) not in meshes:
                meshes[id(source)] = Mesh(np.asarray(source.vertices, np.float32), np.asarray(source.indices, np.uint32))
            mesh = meshes[id(source)]
        size = tuple(float(v) for v in shape_scale[i])
        tf = [float(v) for v in transform[i][:7]]
        entries.setdefault(w, []).append(Shape(kind, size, tuple(tf[:3]), tuple(tf[3:7]), local_body, color, mesh))
        signatures.setdefault(w, []).append((kind, tuple(np.round(size, 5)), local_body, tuple(np.round(tf, 5)),
                                             tuple(np.round(color, 4)), None if mesh is None else id(source)))
    template = signatures.get(0, [])
    for w in range(1, worlds):
        if signatures.get(w, []) != template:
            raise ValueError(f"world {w} differs from world 0 (its shapes, sizes, colours, bodies or offsets), "
                             "but a batch needs identical worlds, as Newton's replicate() makes them")
    shapes = entries.get(-1, []) + entries.get(0, [])
    if not shapes:
        raise ValueError("the model has no visible shapes to draw")
    return shapes, per_world


class NewtonBatchRenderer:
    """Draws a batched Newton simulation, each world seen by its own cameras: the reference's NewtonBatchRenderer
    (python/neraulgraphics/newton.py:235-370). The shapes of one world become instances bound to their bodies once;
    each render hands state.body_q ((world_count x bodies_per_world) poses x, y, z, qx, qy, qz, qw; a GPU tensor,
    a Warp array through DLPack, or NumPy) to set_body_poses, one kernel. The default cameras and lights frame world
    0's shapes in the model's initial pose (model.body_q), as the reference does."""

    def __init__(self, model, width=128, height=128, cameras_per_world=1, *, lights=None, fov=1.0, ambient=0.2,
                 geo_types=None, visible_flag=None, device_id=None):
        shapes, per_world = newton_shapes(model, geo_types, visible_flag)
        self._worlds, self._cameras_per_world = int(model.world_count), int(cameras_per_world)
        self._bodies = per_world
        self.width, self.height, self.fov = int(width), int(height), float(fov)
        self.up = np.asarray(_NEWTON_UP[int(getattr(model, "up_axis", 2))], np.float64)
        scene = build_scene(shapes)
        initial = _host(model.body_q).reshape(-1, 7)[:per_world] if per_world else np.zeros((0, 7))
        positions, reach = [], []
        for sh, sc in zip(shapes, scene.scale):
            p = np.asarray(sh.position, np.float64)[None]
            if sh.body >= 0:
                p, _ = compose_poses(initial[sh.body:sh.body + 1, :3], initial[sh.body:sh.body + 1, 3:7], p,
                                     np.asarray(sh.rotation, np.float64)[None])
            positions.append(p[0])
            reach.append(_reach(sh, sc))
        reach = np.asarray(reach)
        solid = reach > 0
        self.center, extent = scene_bounds(np.asarray(positions).reshape(-1, 3)[solid], reach[solid])
        self.extent = max(extent, 1e-3)
        self._initial_body_q = initial
        self.sim = SimBatchRenderer(shapes, self._worlds, per_world, self.width, self.height, self._cameras_per_world,
                                    lights=self.default_lights() if lights is None else lights, ambient=ambient,
                                    device_id=device_id)
        self.renderer = self.sim.renderer
        self.scene = self.sim.scene

    @property
    def view_count(self) -> int:
        return self._worlds * self._cameras_per_world

    def default_lights(self):
        """A key light high and to one side and a dimmer fill from the other, scaled to the scene
        (python/neraulgraphics/_adapter.py:128-134)."""
        side = _horizontal(self.up)
        across = np.cross(self.up, side)
        key = self.center + self.up * 3.0 * self.extent + side * 1.5 * self.extent + across * 1.0 * self.extent
        fill = self.center + self.up * 2.0 * self.extent - side * 2.0 * self.extent - across * 1.5 * self.extent
        return [Light(tuple(key), intensity=1.0, radius=0.0), Light(tuple(fill), intensity=0.4, radius=0.0)]

    def default_cameras(self):
        """One camera per view, framing world 0's shapes from the side, slightly above: (views, 3, 4)."""
        eye = self.center - _horizontal(self.up) * 2.6 * self.extent + self.up * 0.7 * self.extent
        return np.tile(look_at(eye, self.center, self.up), (self.view_count, 1, 1))

    def update(self, state=None) -> None:
        """Places every instance of every world from state.body_q (default: the model's initial pose)."""
        if not self._bodies:
            return
        body_q = _as_frames(state.body_q) if state is not None else np.tile(self._initial_body_q, (self._worlds, 1))
        self.renderer.set_body_poses(body_q)

    def render(self, state, cameras=None, *, fov=None, rgb=True, depth=False, normals=False, object_ids=False,
               zero_copy=False, frame=0, samples_per_pixel=1):
        """update(state), then the images asked for as a dict (see BatchRenderer.render)."""
        self.update(state)
        return self.renderer.render(self.default_cameras() if cameras is None else cameras, self.fov if fov is None else fov,
                                    samples_per_pixel, rgb=rgb, depth=depth, normals=normals, object_ids=object_ids,
                                    frame=frame, zero_copy=zero_copy)

    def cast_rays(self, origins, directions, max_distance=float("inf"), **kw):
        return self.renderer.cast_rays(origins, directions, max_distance, **kw)


def _horizontal(up):
    """A unit vector perpendicular to `up`, the direction the default camera looks along (_adapter.py:79-83)."""
    probe = np.array([0.0, 1.0, 0.0]) if abs(up[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    h = np.cross(up, np.cross(probe, up))
    return h / np.linalg.norm(h)
