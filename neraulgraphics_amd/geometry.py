"""Scene types and mesh makers of the Python edge, mirroring the reference's (include/neraulgraphics/mesh.h,
scene.h; src/mesh.cpp). Meshes are built in float32 with the reference's operation order, so a mesh made
here has the same vertices, bit for bit, as the reference's make_icosphere / make_box."""
from dataclasses import dataclass, field

import numpy as np

F = np.float32

OBJECT_NONE = 0
OBJECT_SPHERE = 1
MAX_SPHERES = 32
OBJECT_FLOOR = OBJECT_SPHERE + MAX_SPHERES
OBJECT_INSTANCE = OBJECT_FLOOR + 1


@dataclass
class Mesh:
    vertices: np.ndarray                 # (V, 3) float32
    indices: np.ndarray                  # (T*3,) uint32
    normals: np.ndarray | None = None    # (V, 3) float32, or None

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1, 3)

    @property
    def triangle_count(self):
        return self.indices.size // 3

    @property
    def vertex_count(self):
        return len(self.vertices)


@dataclass
class Transform:
    """The top three rows of a 4x4 affine transform, row-major (include/neraulgraphics/mesh.h)."""
    matrix: np.ndarray = field(default_factory=lambda: np.eye(3, 4, dtype=np.float32))

    def __post_init__(self):
        m = np.asarray(self.matrix, dtype=np.float32)
        self.matrix = np.ascontiguousarray(m[:3, :4] if m.shape == (4, 4) else m.reshape(3, 4))


@dataclass
class Material:
    """The classic fields of a material (include/neraulgraphics/scene.h:35-40), with the reference Python
    binding's defaults (python/bindings.cpp:393-395). The physically based fields only serve the reference's path
    tracer, which the batched path does not use."""
    diffuse_color: tuple = (0.8, 0.8, 0.8)
    diffuse_albedo: float = 0.9
    specular_albedo: float = 0.1
    reflectivity: float = 0.0
    specular_exponent: float = 50.0

    def as_array(self):
        return np.array([*self.diffuse_color, self.diffuse_albedo, self.specular_albedo, self.reflectivity,
                         self.specular_exponent, 0.0], np.float32)


@dataclass
class Instance:
    mesh: int
    transform: Transform = field(default_factory=Transform)
    material: Material = field(default_factory=Material)


@dataclass
class Sphere:
    center: tuple
    radius: float
    material: Material = field(default_factory=Material)


@dataclass
class Plane:
    """A checkerboard floor at constant y; half_extent_x 0 is no floor (include/neraulgraphics/scene.h:94-103)."""
    y: float = 0.0
    half_extent_x: float = 0.0
    z_min: float = 0.0
    z_max: float = 0.0
    checker_scale: float = 0.5
    material: Material = field(default_factory=lambda: Material((1.0, 1.0, 1.0), 1.0, 0.0, 0.0, 1.0))


@dataclass
class Light:
    """A sphere light (include/neraulgraphics/scene.h:78-92, defaults of python/bindings.cpp:511-512): radius 0 gives
    hard shadows; falloff > 0 divides the intensity by 1 + (distance / falloff)^2; spot_angle > 0 makes it a spotlight
    along direction, fading over the cone's outer spot_softness; color tints it."""
    position: tuple
    intensity: float = 1.0
    radius: float = 0.0
    falloff: float = 0.0
    direction: tuple = (0.0, 0.0, -1.0)
    spot_angle: float = 0.0
    spot_softness: float = 0.25
    color: tuple = (1.0, 1.0, 1.0)

    def as_array(self):
        return np.array([*self.position, self.intensity, self.radius, self.falloff, *self.direction, self.spot_angle,
                         self.spot_softness, *self.color], np.float32)


@dataclass
class Environment:
    """The sky's gradient (include/neraulgraphics/scene.h:109-121): brighter or darker along up. Environment maps
    are not supported by this backend."""
    up: tuple = (0.0, 1.0, 0.0)
    color_up: tuple = (0.10, 0.26, 0.60)
    color_down: tuple = (0.35, 0.52, 0.70)
    intensity: float = 1.0

    def as_array(self):
        return np.array([*self.up, *self.color_up, *self.color_down, self.intensity], np.float32)


def shading_arrays(instances, spheres=(), floor=None, lights=(), ambient=0.0, environment=None):
    """The RGB picture's inputs as flat float32 arrays, in the layout of the C ABI's ngx_shading (and of the
    oracle): materials 8 floats each, lights 14, sky 10."""
    floor = floor if floor is not None else Plane()
    env = environment if environment is not None else Environment()
    im = np.stack([i.material.as_array() for i in instances]) if instances else np.zeros((0, 8), np.float32)
    sm = np.stack([s.material.as_array() for s in spheres]) if spheres else np.zeros((0, 8), np.float32)
    ls = np.stack([l.as_array() for l in lights]) if lights else np.zeros((0, 14), np.float32)
    return dict(instance_materials=im, sphere_materials=sm, floor_material=floor.material.as_array(),
                floor_checker_scale=float(floor.checker_scale), lights=ls, ambient=float(ambient), sky=env.as_array())


def make_box(h=(1.0, 1.0, 1.0)):
    """An axis-aligned box of half-extents h, 12 triangles (src/mesh.cpp:36-50)."""
    hx, hy, hz = (F(v) for v in h)
    verts = [((hx if i & 1 else -hx), (hy if i & 2 else -hy), (hz if i & 4 else -hz)) for i in range(8)]
    idx = [0, 2, 1, 1, 2, 3, 4, 5, 6, 5, 7, 6, 0, 1, 4, 1, 5, 4, 2, 6, 3, 3, 6, 7, 0, 4, 2, 2, 4, 6, 1, 3, 5, 3, 7, 5]
    return Mesh(np.array(verts, dtype=np.float32), np.array(idx, dtype=np.uint32))


def make_icosphere(radius=1.0, subdivisions=2):
    """A unit icosahedron subdivided n times and scaled: 20 * 4^n triangles (src/mesh.cpp:52-95)."""
    if not radius > 0 or not np.isfinite(radius):
        raise ValueError(f"icosphere radius must be positive, got {radius}")
    if not 0 <= subdivisions <= 8:
        raise ValueError(f"icosphere subdivisions must be 0 to 8, got {subdivisions}")
    verts = []

    def add(x, y, z):
        x, y, z = F(x), F(y), F(z)
        ln = np.sqrt(F(F(x * x) + F(y * y)) + F(z * z))
        verts.append((x / ln, y / ln, z / ln))
        return len(verts) - 1

    t = (F(1.0) + np.sqrt(F(5.0))) / F(2.0)
    for p in [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
              (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]:
        add(*p)
    idx = [0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
           3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]
    for _ in range(subdivisions):
        cache = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key in cache:
                return cache[key]
            va, vb = verts[a], verts[b]
            cache[key] = add(va[0] + vb[0], va[1] + vb[1], va[2] + vb[2])
            return cache[key]

        nxt = []
        for i in range(0, len(idx), 3):
            a, b, c = idx[i], idx[i + 1], idx[i + 2]
            ab = midpoint(a, b)
            bc = midpoint(b, c)
            ca = midpoint(c, a)
            nxt += [a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca]
        idx = nxt
    r = F(radius)
    v = np.array([(x * r, y * r, z * r) for x, y, z in verts], dtype=np.float32)
    return Mesh(v, np.array(idx, dtype=np.uint32))


def make_grid(nx, ny, size_x=2.0, size_y=2.0, displacement=0.05, seed=0):
    """A nx x ny quad grid in the x-y plane, centred on the origin, with z displaced by N(0, displacement):
    2 * nx * ny triangles. Not a reference function: the displaced-grid meshes of the benchmark scenes."""
    rng = np.random.default_rng(seed)
    xs = -0.5 * size_x + size_x * np.arange(nx + 1) / nx
    ys = -0.5 * size_y + size_y * np.arange(ny + 1) / ny
    gx, gy = np.meshgrid(xs, ys)
    gz = rng.normal(0.0, displacement, size=gx.shape)
    v = np.stack([gx, gy, gz], axis=-1).reshape(-1, 3).astype(np.float32)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v00 = (j * (nx + 1) + i).reshape(-1)
    v10, v01 = v00 + 1, v00 + nx + 1
    v11 = v01 + 1
    tris = np.stack([v00, v10, v11, v00, v11, v01], axis=-1).reshape(-1).astype(np.uint32)
    return Mesh(v, tris)


def make_transform(position=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0, 1.0), scale=1.0):
    """Translation, rotation (a quaternion x, y, z, w, normalized) and uniform scale (src/mesh.cpp:20-34)."""
    x, y, z, w = (F(c) for c in rotation)
    norm = np.sqrt(F(F(F(x * x) + F(y * y)) + F(z * z)) + F(w * w))
    if not norm > 0 or not np.isfinite(norm):
        raise ValueError("rotation quaternion must be finite and not zero")
    x, y, z, w = x / norm, y / norm, z / norm, w / norm
    s, one, two = F(scale), F(1.0), F(2.0)
    px, py, pz = (F(c) for c in position)
    m = [[s * (one - two * (y * y + z * z)), s * (two * (x * y - z * w)), s * (two * (x * z + y * w)), px],
         [s * (two * (x * y + z * w)), s * (one - two * (x * x + z * z)), s * (two * (y * z - x * w)), py],
         [s * (two * (x * z - y * w)), s * (two * (y * z + x * w)), s * (one - two * (x * x + y * y)), pz]]
    return Transform(np.array(m, dtype=np.float32))


def make_transforms(positions, rotations, scales):
    """make_transform over arrays: positions (..., 3), rotations (..., 4), scales (...) -> (..., 12) float32,
    each element with the same float32 operations as make_transform."""
    p = np.asarray(positions, np.float32)
    q = np.asarray(rotations, np.float32)
    s = np.asarray(scales, np.float32)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    norm = np.sqrt(((x * x + y * y) + z * z) + w * w)
    x, y, z, w = x / norm, y / norm, z / norm, w / norm
    one, two = F(1.0), F(2.0)
    m = np.stack([s * (one - two * (y * y + z * z)), s * (two * (x * y - z * w)), s * (two * (x * z + y * w)), p[..., 0],
                  s * (two * (x * y + z * w)), s * (one - two * (x * x + z * z)), s * (two * (y * z - x * w)), p[..., 1],
                  s * (two * (x * z - y * w)), s * (two * (y * z + x * w)), s * (one - two * (x * x + y * y)), p[..., 2]],
                 axis=-1)
    return m.astype(np.float32)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """A camera-to-world matrix (3 x 4, OpenGL convention: x right, y up, looking along -z) at eye looking at
    target, computed in float64 and rounded."""
    eye, target, up = (np.asarray(a, dtype=np.float64) for a in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return np.stack([r, u, -f, eye], axis=1).astype(np.float32)
