"""The oracle on the reference's known-answer scenes, body-pose composition against independent
constructions, and the FP64 pose gradient against central differences."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import workloads
from oracle.oracle import forward


def _world_boxes(oracle, positions, half=1.0, counts=None, cams=None, size=(96, 54), fov=1.5708):
    W, H = size
    worlds = len(positions)
    per = max(len(p) for p in positions)
    mats = np.zeros((worlds, per, 12), np.float32)
    for w, ps in enumerate(positions):
        for i, p in enumerate(ps):
            mats[w, i] = ng.make_transform(p).matrix.reshape(12)
    cfg = dict(worlds=worlds, instances=[ng.Instance(0)] * per, counts=counts, meshes=[ng.make_box((half,) * 3)],
               mode="matrices", matrices=mats, width=W, height=H, fov=fov, cameras_per_world=1,
               cameras=np.tile(ng.look_at((0, 0, 0), (0, 0, -1)).reshape(12), (worlds, 1)) if cams is None else cams)
    return cfg, forward(cfg, oracle)


def test_worlds_do_not_see_one_another(oracle):
    """tests/test_batch.cpp:185-214: world 0's box at z=-8 (front face 7), world 1's at z=-3 (front face 2)."""
    cfg, out = _world_boxes(oracle, [[(0, 0, -8)], [(0, 0, -3)]])
    assert np.allclose(out["depth"][:, 27, 48], [7.0, 2.0], atol=0.02)
    st = out
    hits = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                            None, st["world_off"], np.zeros((2, 1, 3)), np.array([[[0, 0, -1]], [[0, 0, -1]]]), 100.0)
    assert np.allclose(hits["distance"][:, 0], [7.0, 2.0], atol=1e-3) and np.all(hits["object_id"] == ng.OBJECT_INSTANCE)


def test_single_instance_worlds_and_a_pile(oracle):
    """tests/test_batch.cpp:311-331: one-leaf trees at 3 + w; 1,000 coincident boxes whose front face is at 4.8."""
    _, out = _world_boxes(oracle, [[(0, 0, -4.0 - w)] for w in range(4)])
    assert np.allclose(out["depth"][:, 27, 48], 3.0 + np.arange(4), atol=0.02)
    _, out = _world_boxes(oracle, [[(0, 0, -5.0)] * 1000], half=0.2)
    assert abs(out["depth"][0, 27, 48] - 4.8) < 0.02


def test_ray_distance_and_max_distance(oracle):
    """tests/test_renderer.cpp:804-812: the distance does not depend on the direction's length; a hit beyond
    max_distance is a miss."""
    cfg, st = _world_boxes(oracle, [[(0, 0, -9.0)]])
    args = (st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"], None, st["world_off"])
    for scale in (1e-3, 1.0, 250.0):
        h = oracle.cast_rays(*args, np.zeros((1, 1, 3)), np.array([[[0, 0, -scale]]]), 100.0)
        assert abs(h["distance"][0, 0] - 8.0) < 1e-5
    assert np.isinf(oracle.cast_rays(*args, np.zeros((1, 1, 3)), np.array([[[0, 0, -1]]]), 7.9)["distance"][0, 0])
    assert abs(oracle.cast_rays(*args, np.zeros((1, 1, 3)), np.array([[[0, 0, -1]]]), 8.1)["distance"][0, 0] - 8.0) < 1e-5
    for bad in ([0, 0, 0], [np.nan, 0, -1], [np.inf, 0, 0]):
        h = oracle.cast_rays(*args, np.zeros((1, 1, 3)), np.array([[bad]], np.float32), 100.0)
        assert np.isinf(h["distance"][0, 0]) and h["object_id"][0, 0] == ng.OBJECT_NONE


def test_depth_is_ray_distance_along_the_view_axis(oracle):
    """tests/test_renderer.cpp:822-834: depth equals the distance of the pixel's ray times cos of its angle."""
    cfg = workloads.make("instanced_worlds", worlds=2)
    out = forward(cfg, oracle)
    W, H = cfg["width"], cfg["height"]
    o = np.zeros((2, H * W, 3), np.float32)
    d = np.zeros((2, H * W, 3), np.float32)
    cos = np.zeros((2, H * W))
    for v in range(2):
        for y in range(H):
            for x in range(W):
                oo, dd, ff = oracle.camera_ray(cfg["cameras"][v], cfg["fov"], W, H, x, y)
                o[v, y * W + x], d[v, y * W + x] = oo, dd
                cos[v, y * W + x] = float(np.dot(dd, ff))
    st = out
    hits = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"], None,
                            st["world_off"], o, d, 1e9)
    dep = out["depth"].reshape(2, -1)
    fin = np.isfinite(dep)
    assert np.array_equal(fin, np.isfinite(hits["distance"]))
    assert np.allclose(dep[fin], (hits["distance"] * cos)[fin], rtol=2e-6)


def _q2m(q):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_body_poses_compose_as_the_reference_test_does(oracle):
    """tests/test_batch.cpp:216-258: two bodies per world carry an instance each with an offset and a scale, a
    third is fixed; the matrix is the body's pose, then the offset, then the scale."""
    worlds = 4
    boi = np.array([0, 1, -1], np.int32)
    local = np.array([[0.5, 0, 0, 0, 0, 0, 1], [0, 0.4, 0, 0, 0.3826834, 0, 0.9238795], [2, 0, -7, 0, 0, 0, 1]], np.float32)
    scale = np.array([[1, 1, 1], [0.5, 0.5, 1.0], [1, 1, 1]], np.float32)
    rng = np.random.default_rng(777)
    bodies = np.zeros((worlds, 2, 7), np.float32)
    for w in range(worlds):
        for k in range(2):
            bodies[w, k] = [-1.5 + 3.0 * k + 0.2 * w, 0.0, -6.0 - rng.random(), 0.0, 0.2588190 * (1 if k else -1), 0.0, 0.9659258]
    off = np.arange(worlds + 1, dtype=np.int32) * 3
    got = oracle.body_poses_to_matrices(off, 2, boi, local, scale, bodies).reshape(worlds, 3, 3, 4)
    for w in range(worlds):
        for i in range(3):
            rl = _q2m(local[i, 3:])
            if boi[i] < 0:
                r, p = rl, local[i, :3].astype(np.float64)
            else:
                bp = bodies[w, boi[i]].astype(np.float64)
                rb = _q2m(bp[3:])
                r, p = rb @ rl, bp[:3] + rb @ local[i, :3]
            want = np.concatenate([r * scale[i].astype(np.float64)[None, :], p[:, None]], axis=1)
            # FP64 composition rounded once: within one float32 rounding of the exact product
            assert np.allclose(got[w, i], want, rtol=0, atol=2e-7 * np.abs(want).max() + 1e-12)


def test_body_poses_against_a_library_rotation(oracle):
    from scipy.spatial.transform import Rotation
    cfg = workloads.make("body_worlds", worlds=3)
    off = np.concatenate([[0], np.cumsum(cfg["counts"])]).astype(np.int32)
    got = oracle.body_poses_to_matrices(off, cfg["bodies"], cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["body_poses"])
    for w in range(3):
        for i in range(cfg["counts"][w]):
            b = cfg["body_of_instance"][i]
            bp = cfg["body_poses"][w, b].astype(np.float64)
            lp = cfg["local_pose"][i].astype(np.float64)
            rb, rl = Rotation.from_quat(bp[3:]).as_matrix(), Rotation.from_quat(lp[3:]).as_matrix()
            want = np.concatenate([(rb @ rl) * cfg["scale"][i].astype(np.float64)[None, :], (bp[:3] + rb @ lp[:3])[:, None]], axis=1)
            assert np.allclose(got[off[w] + i].reshape(3, 4), want, rtol=0, atol=3e-6)


def test_zero_quaternion_is_the_identity(oracle):
    m = oracle.body_poses_to_matrices(np.array([0, 1], np.int32), 1, np.array([0], np.int32),
                                      np.array([[1, 2, 3, 0, 0, 0, 0]], np.float32), np.ones((1, 3), np.float32),
                                      np.array([[[0, 0, 0, 0, 0, 0, 0]]], np.float32))
    assert np.array_equal(m.reshape(3, 4), np.array([[1, 0, 0, 1], [0, 1, 0, 2], [0, 0, 1, 3]], np.float32))


def test_fp64_pose_gradient_equals_central_differences(oracle):
    """The oracle's closed-form dt/d(pose) through a fixed triangle against FP64 central differences: a wrong
    sign, a dropped normalization term or a transposed rotation is off by O(1)."""
    cfg = workloads.make("posed_bodies", worlds=3)
    out = forward(cfg, oracle)
    V, H, W = out["depth"].shape
    rng = np.random.default_rng(1)
    hit = np.flatnonzero(out["hit_inst"].reshape(-1) >= 0)
    pick = rng.choice(hit, size=min(400, len(hit)), replace=False)
    origins, dirs, cosv = np.zeros((len(pick), 3)), np.zeros((len(pick), 3)), np.zeros(len(pick))
    for k, p in enumerate(pick):
        v, y, x = np.unravel_index(p, (V, H, W))
        o, d, f = oracle.camera_ray(cfg["cameras"][v], cfg["fov"], W, H, int(x), int(y))
        origins[k], dirs[k], cosv[k] = o, d, np.dot(d.astype(np.float64), f.astype(np.float64))
    weight = rng.normal(size=len(pick))
    ray_world = (pick // (H * W)).astype(np.int32)
    args = (origins, dirs, weight, cosv, out["hit_inst"].reshape(-1)[pick], out["hit_tri"].reshape(-1)[pick], ray_world,
            out["world_off"], cfg["bodies"], cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["body_poses"],
            out["meshes"]["tris"])
    cf = oracle.pose_grad(0, 0.0, *args)
    fd = oracle.pose_grad(1, 1e-6, *args)
    assert np.linalg.norm(cf - fd) <= 1e-7 * np.linalg.norm(fd)
    assert np.count_nonzero(np.abs(cf) > 0) > 20


def test_a_sample_of_worlds_renders_as_in_the_full_batch(oracle):
    """Checking a full-size GPU batch on a sample of its worlds rests on this: the oracle run on some worlds alone
    gives those worlds' pixels as the run on the whole batch does (equal distances; IDs equal but for ties)."""
    from tests.helpers import compare_images, world_subset
    cfg = workloads.make("body_worlds", worlds=10)
    full = forward(cfg, oracle)
    sample = np.array([0, 4, 9])
    part = forward(world_subset(cfg, sample), oracle)
    got = {k: full[k][sample] for k in ("depth", "object_id", "normal")}
    rep = compare_images(got, part, "sample")
    assert rep["depth_bitexact"] == 1.0 and rep["hits"] > 0
