"""The oracle's RGB picture (a restatement of render_core.h:146-283) against answers fixed by the mathematics:
a sky of one colour is that colour through the sRGB curve; a floor lit straight from above by a point light
is its diffuse colour times intensity times albedo (Lambert at normal incidence); a floor point hidden from
the only light by a sphere gets the ambient term alone; a mirror floor shows the sky; a glossy floor's Phong
highlight, a coloured spotlight's cone and a light's distance falloff follow their formulas pixel by pixel, each
pixel's ray jittered by the first two draws of its PCG stream (which pins the stream's seeding and draw order). The
reference's own
RGB code cannot be run here (its surface shading calls the CPU backend's texture sampler, which does not
build), so these closed forms are what pins it."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from oracle.oracle import forward


def srgb_byte(v):
    """IEC 61966-2-1's transfer curve, then rounding to a byte."""
    v = np.clip(np.float32(v), 0, 1)
    s = np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(v, 1 / 2.4) - 0.055)
    return np.floor(np.clip(s, 0, 1) * 255 + 0.5).astype(np.int32)


def scene(lights=(), spheres=(), floor=None, ambient=0.0, env=None, cam=None, size=32, spp=1, floor_material=None):
    # one box far behind the camera keeps the trees non-empty without being seen
    inst = [ng.Instance(0, ng.make_transform((0.0, 0.0, 500.0)), ng.Material())]
    cfg = dict(width=size, height=size, worlds=1, cameras_per_world=1, meshes=[ng.make_box((0.5, 0.5, 0.5))], instances=inst,
               counts=None, mode="matrices", matrices=np.stack([i.transform.matrix.reshape(12) for i in inst])[None],
               cameras=(cam if cam is not None else ng.look_at((0, 0, 0), (0, 0, -1))).reshape(1, 12), fov=1.0, spp=spp)
    fl = floor if floor is not None else ng.Plane()
    cfg["spheres"] = np.array([[*s.center, s.radius] for s in spheres], np.float32) if spheres else None
    cfg["floor"] = np.array([fl.y, fl.half_extent_x, fl.z_min, fl.z_max], np.float32)
    cfg["shading"] = ng.shading_arrays(inst, list(spheres), fl, list(lights), ambient, env)
    return cfg


def test_a_sky_of_one_colour_is_that_colour(oracle):
    c = (0.2, 0.5, 0.8)
    cfg = scene(env=ng.Environment(color_up=c, color_down=c, intensity=0.75))
    rgb = forward(cfg, oracle, outputs=("rgb",))["rgb"]
    want = srgb_byte(np.array(c, np.float32) * np.float32(0.75))
    assert np.all(rgb.reshape(-1, 3) == want)


def test_a_floor_lit_from_straight_above(oracle):
    # camera 1 above the floor looking straight down; a point light 4 above; diffuse albedo 0.8, no specular
    fm = ng.Material((1.0, 1.0, 1.0), 0.8, 0.0, 0.0, 1.0)
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=fm)
    light = ng.Light((0.0, 3.0, 0.0), intensity=1.5)
    cam = ng.look_at((0, 0, 0), (0, -1, 0), up=(0, 0, -1))
    cfg = scene(lights=[light], floor=floor, cam=cam, size=33)
    rgb = forward(cfg, oracle, outputs=("rgb",))["rgb"][0]
    # the checker cell (0, 0) is the tinted one: diffuse colour 0.3 x (1, 0.7, 0.3)
    want = srgb_byte(np.float32(0.3) * np.array([1.0, 0.7, 0.3], np.float32) * np.float32(1.5 * 0.8))
    centre = rgb[16, 16].astype(np.int32)
    assert np.all(np.abs(centre - want) <= 1), (centre, want)


def test_a_shadowed_point_gets_only_ambient(oracle):
    fm = ng.Material((1.0, 1.0, 1.0), 0.5, 0.0, 0.0, 1.0)
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=fm)
    light = ng.Light((0.0, 3.0, 0.0), intensity=2.0)
    blocker = ng.Sphere((0.0, 1.5, 0.0), 0.5)
    cam = ng.look_at((0, 0, 0), (0, -1, 0), up=(0, 0, -1))
    dark = forward(scene(lights=[light], spheres=[blocker], floor=floor, cam=cam, size=33), oracle, outputs=("rgb",))["rgb"][0]
    assert np.all(dark[16, 16] == 0), dark[16, 16]
    lit = forward(scene(lights=[light], spheres=[blocker], floor=floor, cam=cam, size=33, ambient=0.4), oracle, outputs=("rgb",))["rgb"][0]
    want = srgb_byte(np.float32(0.3) * np.array([1.0, 0.7, 0.3], np.float32) * np.float32(0.4 * 0.5))
    assert np.all(np.abs(lit[16, 16].astype(np.int32) - want) <= 1), (lit[16, 16], want)


def test_a_mirror_floor_shows_the_sky(oracle):
    # a perfect mirror with no diffuse albedo: what the camera sees is the sky in the reflected direction
    fm = ng.Material((1.0, 1.0, 1.0), 0.0, 0.0, 1.0, 1.0)
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=fm)
    c = (0.3, 0.6, 0.1)
    cam = ng.look_at((0, 0, 0), (0, -1, 0), up=(0, 0, -1))
    rgb = forward(scene(floor=floor, cam=cam, env=ng.Environment(color_up=c, color_down=c)), oracle, outputs=("rgb",))["rgb"][0]
    assert np.all(rgb.reshape(-1, 3) == srgb_byte(np.array(c, np.float32)))


def test_an_area_light_is_sampled_per_pixel_and_frame(oracle):
    """A sphere light of radius 0.6 over a floor with a blocker: each sample lights a point from a random point of the
    light, so lit floor varies from pixel to pixel, differently in another frame, less with 16 samples a pixel;
    floor points that a single sample found in shadow are, on average over 16 samples, darker than open floor."""
    fm = ng.Material((1.0, 1.0, 1.0), 0.8, 0.0, 0.0, 1.0)
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=fm)
    light = ng.Light((0.0, 3.0, 0.0), intensity=1.5, radius=0.6)
    blocker = ng.Sphere((0.0, 1.0, 0.0), 0.35)
    cam = ng.look_at((0, 5, 0), (0, -1, 0), up=(0, 0, -1))  # sees the floor out to about 3 from the light's foot
    imgs, ids = {}, None
    for spp, frame in [(1, 0), (1, 1), (16, 0)]:
        cfg = scene(lights=[light], spheres=[blocker], floor=floor, cam=cam, size=33, spp=spp)
        cfg["frame"] = frame
        out = forward(cfg, oracle, outputs=("rgb", "id"))
        imgs[(spp, frame)] = out["rgb"][0].astype(np.int32)
        ids = out["object_id"][0]
    a, b, c = imgs[(1, 0)], imgs[(1, 1)], imgs[(16, 0)]
    ring = np.zeros(ids.shape, bool)
    ring[:4, :] = ring[-4:, :] = ring[:, :4] = ring[:, -4:] = True  # far from the light's foot: lit floor
    assert np.all(ids[ring] == ng.OBJECT_FLOOR)
    assert not np.array_equal(a, b)
    assert c[ring].std() < a[ring].std()
    shaded = (ids == ng.OBJECT_FLOOR) & (a[..., 0] == 0)
    assert shaded.sum() > 0 and c[shaded].mean() < c[ring].mean()


# ---- direct light over the floor, in closed form, with the jittered camera rays worked out independently ----------

M32 = np.uint64(0xFFFFFFFF)


def pcg_hash(x):
    """The PCG-family integer hash (Jarzynski and Olano, "Hash Functions for GPU Rendering", 2020), on uint32 arrays."""
    x = np.asarray(x, np.uint64) & M32
    state = (x * np.uint64(747796405) + np.uint64(2891336453)) & M32
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & M32
    return ((word >> np.uint64(22)) ^ word) & M32


def jitter(width, height, frame):
    """Per pixel, the first two uniforms of its stream: seeded by hash(pixel index ^ hash(frame + 1)), each draw
    re-hashing the state and scaling it by 2^-32; x's offset is drawn first."""
    idx = np.arange(width * height, dtype=np.uint64)
    state = pcg_hash(idx ^ pcg_hash(np.uint64(frame + 1)))
    s1 = pcg_hash(state)
    s2 = pcg_hash(s1)
    to_unit = np.float32(2.3283064365386963e-10)
    jx = s1.astype(np.float32) * to_unit
    jy = s2.astype(np.float32) * to_unit
    return jx.reshape(height, width), jy.reshape(height, width)


def camera_dirs(cam12, fov, width, height, jx, jy):
    """Unit directions (float64) through (x + jx, y + jy) of an OpenGL camera (x right, y up, looking along -z)."""
    m = np.asarray(cam12, np.float64).reshape(3, 4)
    right, up, back = (m[:, k] / np.linalg.norm(m[:, k]) for k in range(3))
    fs = np.tan(fov / 2)
    ys, xs = np.mgrid[0:height, 0:width]
    sx = (2.0 * (xs + jx) / width - 1.0) * fs * (width / height)
    sy = -(2.0 * (ys + jy) / height - 1.0) * fs
    d = sx[..., None] * right + sy[..., None] * up - back
    return m[:, 3], d / np.linalg.norm(d, axis=-1, keepdims=True)


def srgb_float(v):
    v = np.clip(v, 0.0, 1.0)
    s = np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(v, 1 / 2.4) - 0.055)
    return np.floor(np.clip(s, 0, 1) * 255 + 0.5).astype(np.int32)


def floor_direct(eye, dirs, floor_y, light, mat, tint):
    """Radiance of the floor (normal +y) seen along dirs from eye, lit by one point light with nothing in the way: the
    diffuse colour times the Lambert term times the diffuse albedo, plus the Phong term max(0, reflect(l, n) . d)^e
    times the specular albedo, the light scaled by 1 / (1 + (distance / falloff)^2) and by the spotlight's smoothstep
    from the cone's outer edge to its inner one."""
    t = (floor_y - eye[1]) / dirs[..., 1]
    p = eye + dirs * t[..., None]
    to_l = np.asarray(light.position, np.float64) - p
    dist = np.linalg.norm(to_l, axis=-1)
    l = to_l / dist[..., None]
    f = np.ones_like(dist)
    if light.falloff > 0:
        f /= 1.0 + (dist / light.falloff) ** 2
    if light.spot_angle > 0:
        axis = np.asarray(light.direction, np.float64)
        axis /= np.linalg.norm(axis)
        c = -(l @ axis)
        outer, inner = np.cos(light.spot_angle), np.cos(light.spot_angle * (1.0 - light.spot_softness))
        s = np.clip((c - outer) / max(inner - outer, 1e-6), 0.0, 1.0)
        f *= s * s * (3.0 - 2.0 * s)
    strength = light.intensity * f
    lam = np.maximum(0.0, l[..., 1])
    r = l.copy()
    r[..., 1] = -r[..., 1]  # reflect(l, n) with n = +y
    phong = np.maximum(0.0, np.sum(r * dirs, axis=-1)) ** mat.specular_exponent
    col = np.asarray(light.color, np.float64)
    lit = strength[..., None] * col
    return np.asarray(tint, np.float64) * lit * lam[..., None] * mat.diffuse_albedo + lit * phong[..., None] * mat.specular_albedo


def _floor_scene(light, mat, size=48, frame=0, eye=(0.0, 0.0, 0.0)):
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=mat)
    cam = ng.look_at(eye, (0.3, -1.0, -0.2), up=(0, 0, -1))
    cfg = scene(lights=[light], floor=floor, cam=cam, size=size)
    cfg["frame"] = frame
    return cfg, cam.reshape(12)


TINT = np.float32(0.3) * np.array([1.0, 0.7, 0.3])  # checker cell (0, 0): the floor's tinted colour


def test_phong_highlight_and_lambert_term_follow_the_closed_form(oracle):
    """A glossy floor (diffuse albedo 0.5, specular albedo 0.7, exponent 40) under a point light, seen at an angle:
    every pixel equals the closed form within one level, with the jitter of each pixel's ray taken from its PCG stream.
    The highlight is steep enough that pixel-centre rays (no jitter) or the jitter's two draws swapped miss it."""
    mat = ng.Material((1.0, 1.0, 1.0), 0.5, 0.7, 0.0, 40.0)
    light = ng.Light((0.9, 1.5, -0.5), intensity=1.4)
    cfg, cam = _floor_scene(light, mat, frame=2)
    got = forward(cfg, oracle, outputs=("rgb",))["rgb"][0].astype(np.int32)
    jx, jy = jitter(48, 48, 2)
    eye, d = camera_dirs(cam, 1.0, 48, 48, jx, jy)
    want = srgb_float(floor_direct(eye, d, -1.0, light, mat, TINT))
    assert np.abs(got - want).max() <= 1, np.abs(got - want).max()
    assert got.max() >= 200 and len(np.unique(got[..., 0])) > 40  # a highlight and a gradient, not a flat colour
    # the check of the check: without the jitter, or with its draws swapped, the same closed form misses
    for a, b in ((np.full_like(jx, 0.5), np.full_like(jy, 0.5)), (jy, jx)):
        _, d2 = camera_dirs(cam, 1.0, 48, 48, a, b)
        assert np.abs(got - srgb_float(floor_direct(eye, d2, -1.0, light, mat, TINT))).max() > 1


def test_spotlight_cone_and_distance_falloff_follow_the_closed_form(oracle):
    """A coloured spotlight (half-angle 0.45 rad, softness 0.5) with falloff 1.2 over a matte floor: full strength
    inside the inner cone, nothing outside the outer one, a smoothstep between, all divided by 1 + (d / 1.2)^2;
    every pixel within one level of the closed form."""
    mat = ng.Material((1.0, 1.0, 1.0), 0.9, 0.0, 0.0, 1.0)
    light = ng.Light((0.3, 1.0, -0.4), intensity=4.0, falloff=1.2, direction=(0.1, -1.0, 0.05), spot_angle=0.45,
                     spot_softness=0.5, color=(1.0, 0.8, 0.6))
    cfg, cam = _floor_scene(light, mat, frame=5, eye=(0.0, 1.5, 0.0))
    got = forward(cfg, oracle, outputs=("rgb",))["rgb"][0].astype(np.int32)
    jx, jy = jitter(48, 48, 5)
    eye, d = camera_dirs(cam, 1.0, 48, 48, jx, jy)
    want = srgb_float(floor_direct(eye, d, -1.0, light, mat, TINT))
    assert np.abs(got - want).max() <= 1
    # the cone is in view: lit pixels, dark pixels outside it, and a soft rim
    assert (got[..., 0] == 0).sum() > 100 and (got[..., 0] > 60).sum() > 50
    rim = (got[..., 0] > 0) & (got[..., 0] < 0.5 * got[..., 0].max())
    assert rim.sum() > 20
    # the check of the check: no falloff, or a hard cone, changes the picture beyond a level
    for other in (ng.Light(light.position, 4.0, 0.0, 0.0, light.direction, 0.45, 0.5, light.color),
                  ng.Light(light.position, 4.0, 0.0, 1.2, light.direction, 0.45, 0.0, light.color)):
        assert np.abs(got - srgb_float(floor_direct(eye, d, -1.0, other, mat, TINT))).max() > 1
