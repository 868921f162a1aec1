"""The oracle's normals and its sphere and floor hits, pinned against what fixes them independently of the code
under test: the reference's own known answers (tests/test_renderer.cpp:722-768), the normal of a flat face under an
affine map (the inverse transpose of the map applied to the face's normal), and, for vertex normals, the sphere:
with unit vertex normals equal to the vertices of a unit icosphere, the normal interpolated at a point of a flat
triangle is that point itself, so the smooth normal is the transformed direction of the hit point, found from the
ray and its distance alone (src/mesh_dev.h:238-262). A rows-for-columns slip, a missing inverse transpose or
swapped barycentrics each fail one of these."""
import numpy as np
import pytest

from neraulgraphics_amd.geometry import Instance, Mesh, Transform, make_box, make_icosphere
from oracle.oracle import BLAS_NODE, TLAS_NODE, forward

W, H = 320, 180                      # tests/test_renderer.cpp:54
FOV = 1.5708                         # camera_at_origin (tests/test_renderer.cpp:78): at the origin looking down -z
IDENTITY_CAMERA = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)


def _no_mesh_render(oracle, spheres=None, floor=None):
    """One world with no mesh instances: spheres (n, 4) and a floor (y, half extent x, z min, z max)."""
    return oracle.render(np.zeros(1, TLAS_NODE), np.array([-1], np.int32), np.zeros((1, 12), np.float32), np.zeros(1, np.int32),
                         np.zeros(1, BLAS_NODE), np.zeros((1, 12), np.float32), None, np.array([0, 0], np.int32), IDENTITY_CAMERA,
                         FOV, W, H, spheres=spheres, floor=floor)


def _scene(meshes, matrices, cameras=IDENTITY_CAMERA, width=W, height=H, fov=FOV):
    matrices = np.asarray(matrices, np.float32).reshape(1, -1, 12)
    insts = [Instance(m) for m in range(len(meshes))] if len(meshes) == matrices.shape[1] else \
        [Instance(0) for _ in range(matrices.shape[1])]
    return dict(width=width, height=height, worlds=1, cameras_per_world=1, meshes=meshes, instances=insts, counts=None,
                mode="matrices", matrices=matrices, cameras=np.asarray(cameras, np.float32), fov=fov)


def _quad(half):
    """tests/test_renderer.cpp:194-200: a square facing the camera at z = -5."""
    v = np.array([[-half, -half, -5], [half, -half, -5], [half, half, -5], [-half, half, -5]], np.float32)
    return Mesh(v, np.array([0, 1, 2, 0, 2, 3], np.uint32))


def test_a_spheres_near_point_depth_normal_and_id(oracle):
    # tests/test_renderer.cpp:739-749: a sphere 8 away at its nearest point
    f = _no_mesh_render(oracle, spheres=np.array([[0, 0, -10, 2]], np.float32))
    d, n, i = f["depth"][0], f["normal"][0], f["object_id"][0]
    assert abs(d[H // 2, W // 2] - 8.0) < 0.02
    assert abs(n[H // 2, W // 2, 0]) < 0.05 and abs(n[H // 2, W // 2, 1]) < 0.05 and n[H // 2, W // 2, 2] > 0.99
    assert i[H // 2, W // 2] == 1  # OBJECT_SPHERE
    # where nothing is hit: depth +infinity, normal 0, ID OBJECT_NONE
    assert np.isposinf(d[3, 3]) and i[3, 3] == 0 and n[3, 3, 0] == 0 and n[3, 3, 2] == 0
    # every sphere pixel's normal is the unit vector from the centre to the hit point, facing the camera
    dirs = _pixel_rays()
    hit = np.isfinite(d)
    t = d[hit] / np.abs(dirs[hit][:, 2])          # depth is t times the ray's cosine to the view axis (-z)
    p = dirs[hit] * t[:, None]
    want = (p - [0, 0, -10]) / 2.0
    assert np.max(np.abs(n[hit] - want)) < 2e-5
    assert np.all(np.sum(n[hit] * dirs[hit], axis=1) < 0)


def test_a_wall_is_five_deep_everywhere_and_faces_the_camera(oracle):
    # tests/test_renderer.cpp:751-758
    f = forward(_scene([_quad(3.0)], [np.eye(3, 4)]), oracle)
    d, n, i = f["depth"][0], f["normal"][0], f["object_id"][0]
    assert abs(d[H // 2 + 20, W // 2 + 40] - 5.0) < 1e-3 and abs(d[H // 2 - 15, W // 2 - 30] - 5.0) < 1e-3
    assert i[H // 2, W // 2] == 34  # OBJECT_INSTANCE
    assert abs(n[H // 2 + 20, W // 2 + 40, 2] - 1.0) < 1e-4
    hit = np.isfinite(d)
    assert hit.sum() > 1000
    assert np.all(np.abs(d[hit] - 5.0) < 1e-5 * 5.0) and np.all(n[hit] == [0.0, 0.0, 1.0])
    # wound the other way its triangles face away, and the normal is still turned towards the camera
    back = _quad(3.0)
    back.indices = back.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1).copy()
    g = forward(_scene([back], [np.eye(3, 4)]), oracle)
    assert np.array_equal(np.isfinite(g["depth"][0]), hit) and np.all(g["normal"][0][hit] == [0.0, 0.0, 1.0])


def test_the_floor_depth_follows_the_geometry_and_its_normal_is_up(oracle):
    # tests/test_renderer.cpp:760-768: the floor at y = -2 seen from the origin
    f = _no_mesh_render(oracle, floor=np.array([-2.0, 50.0, -100.0, 100.0], np.float32))
    d, n, i = f["depth"][0], f["normal"][0], f["object_id"][0]
    row = H * 3 // 4
    expect = 2.0 / (2.0 * (row + 0.5) / H - 1.0)
    assert i[row, W // 2] == 33 and abs(d[row, W // 2] - expect) < 0.01 * expect  # OBJECT_FLOOR
    assert abs(n[row, W // 2, 1] - 1.0) < 1e-5
    # every floor row: depth 2 / (2 (y + 0.5) / H - 1) (the floor is 2 below the eye), the upper half sees nothing
    rows = np.arange(H // 2 + 1, H)
    want = 2.0 / (2.0 * (rows + 0.5) / H - 1.0)
    centre = d[rows, W // 2]
    seen = np.isfinite(centre)
    assert seen.sum() > 60
    assert np.all(np.abs(centre[seen] - want[seen]) < 1e-5 * want[seen])
    assert np.all(np.isposinf(d[: H // 2]))
    assert np.all(n[np.isfinite(d)] == [0.0, 1.0, 0.0])


def test_the_nearer_of_two_spheres_is_reported(oracle):
    # tests/test_renderer.cpp:770-775
    f = _no_mesh_render(oracle, spheres=np.array([[0, 0, -3, 1], [0, 0, -8, 2]], np.float32))
    assert f["object_id"][0][H // 2, W // 2] == 1 and abs(f["depth"][0][H // 2, W // 2] - 2.0) < 0.02
    # the farther sphere (14.5 degrees across from the eye) hides wholly behind the nearer one (19.5 degrees)
    ids = f["object_id"][0]
    assert (ids == 1).sum() > 1000 and set(np.unique(ids)) == {0, 1}


def _pixel_rays(width=W, height=H, fov=FOV):
    """Unit directions through the pixel centres of the camera at the origin looking down -z (float64, from the
    pinhole geometry: x right, y up, the image plane at distance 1 spanning tan(fov / 2) vertically)."""
    tan = np.tan(fov / 2)
    sx = (2.0 * (np.arange(width) + 0.5) / width - 1.0) * tan * width / height
    sy = -(2.0 * (np.arange(height) + 0.5) / height - 1.0) * tan
    gx, gy = np.meshgrid(sx, sy)
    d = np.stack([gx, gy, -np.ones_like(gx)], axis=-1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)], axis=1)


def _random_affine(rng, n, scale_lo=0.4, scale_hi=1.8):
    """n matrices R1 diag(s) R2 | p: random rotations and per-axis scales along axes that are not the mesh's (so
    the map shears the mesh's faces), placed in front of the camera."""
    s = rng.uniform(scale_lo, scale_hi, size=(n, 3))
    A = _rotations(rng, n) @ (s[:, :, None] * _rotations(rng, n))
    p = np.stack([rng.uniform(-3, 3, n), rng.uniform(-1.5, 1.5, n), rng.uniform(-9, -6, n)], axis=1)
    return np.concatenate([A, p[:, :, None]], axis=2).astype(np.float32)


def test_face_normals_of_rotated_unevenly_scaled_boxes(oracle):
    """Each hit pixel's normal is the box face's normal carried to the world by the inverse transpose of the
    instance's map (M^-T n), normalized and turned to face the ray."""
    rng = np.random.default_rng(31)
    mats = _random_affine(rng, 20)
    cfg = _scene([make_box((0.5, 0.5, 0.5))], mats.reshape(1, 20, 12), width=192, height=108, fov=0.7)
    f = forward(cfg, oracle)
    n, inst, tri = f["normal"][0], f["hit_inst"][0], f["hit_tri"][0]
    hit = inst >= 0
    assert hit.sum() > 2000 and len(np.unique(inst[hit])) >= 8
    dirs = _pixel_rays(192, 108, 0.7)
    tris = f["meshes"]["tris"]
    worst = 0.0
    faces = set()
    for (y, x) in zip(*np.nonzero(hit)):
        t = tris[tri[y, x]]
        local = np.cross(t[4:7].astype(np.float64), t[8:11].astype(np.float64))
        axis = int(np.argmax(np.abs(local)))
        face = np.zeros(3)
        face[axis] = np.sign(local[axis])      # the box's faces are axis aligned: exactly one component is not zero
        faces.add((axis, face[axis]))
        M = mats[inst[y, x]].astype(np.float64)
        want = np.linalg.inv(M[:, :3]).T @ face
        want /= np.linalg.norm(want)
        if want @ dirs[y, x] > 0:
            want = -want
        worst = max(worst, float(np.max(np.abs(n[y, x] - want))))
    assert len(faces) == 6
    assert worst < 2e-6, worst  # float32 inverse rows and normalization: a few units in the last place
    # the image's normals all face the camera
    assert np.all(np.sum(n[hit] * dirs[hit], axis=1) <= 0)


def test_the_inverse_transpose_matters_here():
    """A check of the check: for these maps, the plain map (M n) and the inverse (M^-1 n) give other directions than
    the inverse transpose, so the test above would catch either slip."""
    rng = np.random.default_rng(31)
    mats = _random_affine(rng, 20).astype(np.float64)
    for M in mats:
        for face in np.eye(3):
            right = np.linalg.inv(M[:, :3]).T @ face
            right /= np.linalg.norm(right)
            for wrong in (M[:, :3] @ face, np.linalg.inv(M[:, :3]) @ face):
                wrong /= np.linalg.norm(wrong)
                assert np.max(np.abs(np.abs(wrong) - np.abs(right))) > 1e-3


def _unit_icosphere_with_normals(subdivisions=2):
    m = make_icosphere(1.0, subdivisions)
    return Mesh(m.vertices, m.indices, normals=m.vertices.copy())


def test_smooth_normals_of_a_unit_icosphere_are_the_hit_points_direction(oracle):
    """Vertex normals equal to the unit vertices: the normal interpolated at a point x of a flat triangle is x, so
    the smooth normal is M^-T x normalized, x being the hit point in mesh space (from the ray and its distance)."""
    rng = np.random.default_rng(5)
    mats = _random_affine(rng, 1, 0.6, 1.6)
    mats[0, :, 3] = [0.2, -0.1, -6.0]
    mesh = _unit_icosphere_with_normals(2)
    built = oracle.build_meshes([mesh])
    assert built["attrs"] is not None
    # rays from the origin through a grid of points around the instance's centre
    n_side = 48
    gx, gy = np.meshgrid(np.linspace(-0.35, 0.35, n_side), np.linspace(-0.3, 0.3, n_side))
    dirs = np.stack([gx.ravel(), gy.ravel(), -np.ones(gx.size)], axis=1).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    cfg = _scene([mesh], mats)
    st = forward(dict(cfg, cameras=None), oracle, outputs=())
    hits = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], built["nodes"], built["tris"], built["attrs"],
                            st["world_off"], np.zeros((1, len(dirs), 3), np.float32), dirs.reshape(1, -1, 3), 100.0)
    t, n, tri = hits["distance"][0], hits["normal"][0], hits["hit_tri"][0]
    hit = np.isfinite(t)
    assert hit.sum() > 800
    M = mats[0].astype(np.float64)
    A, p = M[:, :3], M[:, 3]
    x_world = dirs[hit].astype(np.float64) * t[hit, None].astype(np.float64)
    x_local = np.linalg.solve(A, (x_world - p).T).T
    want = (np.linalg.inv(A).T @ x_local.T).T
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.max(np.abs(n[hit] - want)) < 1e-5, np.max(np.abs(n[hit] - want))
    # and this is the smooth branch: the flat faces' normals differ from it on most hits
    tris = built["tris"][tri[hit]].astype(np.float64)
    flat = np.cross(tris[:, 4:7], tris[:, 8:11])
    flat = (np.linalg.inv(A).T @ flat.T).T
    flat /= np.linalg.norm(flat, axis=1, keepdims=True)
    flat *= np.sign(np.sum(flat * want, axis=1))[:, None]
    assert np.mean(np.max(np.abs(flat - want), axis=1) > 1e-3) > 0.9


def test_inward_vertex_normals_are_turned_to_the_geometric_side(oracle):
    """Vertex normals pointing into the sphere: the smooth normal is flipped to the side of the geometric normal,
    which faces the ray (mesh_dev.h:255-257), so the image is the same as with outward normals."""
    mesh_out = _unit_icosphere_with_normals(2)
    mesh_in = Mesh(mesh_out.vertices, mesh_out.indices, normals=-mesh_out.vertices)
    mats = np.array([[1.3, 0, 0, 0, 0, 0.8, 0, 0, 0, 0, 1.1, -5]], np.float32)
    a = forward(_scene([mesh_out], mats, width=96, height=54, fov=0.8), oracle)
    b = forward(_scene([mesh_in], mats, width=96, height=54, fov=0.8), oracle)
    hit = np.isfinite(a["depth"][0])
    assert hit.sum() > 500
    assert np.array_equal(a["normal"][0], b["normal"][0])


def test_a_mesh_without_normals_keeps_its_face_normal_beside_one_with_them(oracle):
    """Attributes are per mesh: in a scene where one mesh has vertex normals and one has none, the second keeps
    its flat normal (mesh_dev.h:251 tests the mesh's flag)."""
    smooth = _unit_icosphere_with_normals(1)
    flat = make_icosphere(1.0, 1)
    mats = np.array([[1, 0, 0, -1.2, 0, 1, 0, 0, 0, 0, 1, -5], [1, 0, 0, 1.2, 0, 1, 0, 0, 0, 0, 1, -5]], np.float32)
    cfg = _scene([smooth, flat], mats, width=192, height=108, fov=0.8)
    f = forward(cfg, oracle)
    n, ids, tri = f["normal"][0], f["object_id"][0], f["hit_tri"][0]
    tris = f["meshes"]["tris"]
    for obj, smooth_expected in ((34, True), (35, False)):
        sel = ids == obj
        assert sel.sum() > 100
        g = np.cross(tris[tri[sel]][:, 4:7].astype(np.float64), tris[tri[sel]][:, 8:11].astype(np.float64))
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        g *= np.sign(np.sum(g * n[sel], axis=1))[:, None]
        same = np.max(np.abs(g - n[sel]), axis=1) < 1e-6
        assert (same.mean() < 0.1) if smooth_expected else same.all()
