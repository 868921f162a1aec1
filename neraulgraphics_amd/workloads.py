"""Seeded synthetic workloads of this backend, at full size or scaled down, by name (`make(name, ...)`).

  one_world_icospheres  one world, 64x64 depth+ID, 16 icospheres (320 triangles) at U[-5,5]^3 + (0,0,-10),
      random unit quaternion, uniform scale U[0.5,1.5]; camera at the origin looking down -z; fov 1.0. Seed 1000.
  instanced_worlds  4096 worlds, 64x64 depth+ID; 32 instances per world over three shared meshes (icosphere
      320, box 12, displaced grid 2000), placed as above per world; one camera per world at U[-2,2]^3 looking
      at the world's instance centroid; new instance matrices every step. Seed 1001.
  body_worlds  4096 worlds, 128x128 depth+ID; a 256-instance template over the three meshes, U{8..256}
      instances per world; instances carried by 64 bodies (body i mod 64, local offset U[-0.5,0.5]^3 and
      random rotation, per-axis scale U[0.5,1.5]); body poses at U[-5,5]^3 + (0,0,-10) with random rotations;
      new body poses every step. Seed 1002.
  lidar_sweeps  1024 worlds of 65,536 rays (64 beams over elevations [-15, 15] degrees x 1024 azimuths over
      [0, 360)) from an origin U[-1,1]^3; 100,000 instances over all worlds (672 worlds of 98, 352 of 97) of
      one 138,000-triangle displaced grid at U[-10,10]^3; max distance 100, miss = inf. Seed 1003.
  posed_bodies  differentiable depth: 1024 worlds, 64x64 depth, 32 instances as in instanced_worlds, each on
      its own body; the target is rendered from poses perturbed by N(0, 0.01) in translation and 1 degree
      about a random axis. Seed 1004.

Quaternions are Gaussian four-vectors normalized. Everything is float32.

falling_bodies(nbody, ...) is the scene of the reference's own batch-renderer benchmark (bench/scene.py,
bench/run_ours.py): per world `nbody` free bodies in rows of 8, even ones a ball (icosphere of 2 subdivisions,
radius 0.12, 320 triangles), odd ones a box (half-size 0.1, 12 triangles), no ground; one camera framing the rows;
one point light at (0, 0, 4) with no radius; RGB, one sample a pixel. The reference replays poses recorded from
MuJoCo Warp physics; MuJoCo is absent here, so the bodies fall under gravity from the same starting layout (nudged
and spun as bench/gen_poses.py:33-40 does) to rest on the ground, turning as they go, over the same 24 frames
20 ms apart."""
import numpy as np

from .geometry import Instance, Transform, look_at, make_box, make_grid, make_icosphere, make_transforms

FOV = 1.0


def _quat(rng, n):
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def three_meshes(seed=7):
    return [make_icosphere(1.0, 2), make_box((0.5, 0.5, 0.5)), make_grid(40, 25, 2.0, 1.25, 0.05, seed=seed)]


def _placed_worlds(rng, worlds, per_world, offset=(0.0, 0.0, -10.0), spread=5.0):
    """Per world: positions, quaternions, uniform scales and the float32 matrices they make."""
    pos = (rng.uniform(-spread, spread, size=(worlds, per_world, 3)) + np.asarray(offset)).astype(np.float32)
    quat = _quat(rng, worlds * per_world).reshape(worlds, per_world, 4)
    scale = rng.uniform(0.5, 1.5, size=(worlds, per_world)).astype(np.float32)
    return pos, quat, scale, make_transforms(pos, quat, scale)


def _cameras_at(rng, targets):
    cams = np.empty((len(targets), 12), np.float32)
    for v, t in enumerate(targets):
        eye = rng.uniform(-2.0, 2.0, size=3)
        cams[v] = look_at(eye, t).reshape(12)
    return cams


NAMES = ("one_world_icospheres", "instanced_worlds", "body_worlds", "lidar_sweeps", "posed_bodies")


def make(name, worlds=None, rays_per_world=None, seed=None):
    """The named workload as a dict (see the module docstring); `worlds` and `rays_per_world` scale it down."""
    if name not in NAMES:
        raise ValueError(f"no workload {name!r}; the workloads are {', '.join(NAMES)}")
    rng = np.random.default_rng(1000 + NAMES.index(name) if seed is None else seed)
    if name == "one_world_icospheres":
        mesh = make_icosphere(1.0, 2)
        pos, quat, scale, mats = _placed_worlds(rng, 1, 16)
        cams = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)
        return dict(width=64, height=64, worlds=1, cameras_per_world=1, meshes=[mesh],
                    instances=[Instance(0, Transform(mats[0, i].reshape(3, 4))) for i in range(16)], counts=None,
                    mode="matrices", matrices=mats, cameras=cams, fov=FOV)
    if name in ("instanced_worlds", "posed_bodies"):
        W = worlds or (4096 if name == "instanced_worlds" else 1024)
        P = 32
        meshes = three_meshes()
        pos, quat, scale, mats = _placed_worlds(rng, W, P)
        cams = _cameras_at(rng, pos.mean(axis=1))
        insts = [Instance(i % 3) for i in range(P)]
        out = dict(width=64, height=64, worlds=W, cameras_per_world=1, meshes=meshes, instances=insts, counts=None,
                   cameras=cams, fov=FOV)
        if name == "instanced_worlds":
            out.update(mode="matrices", matrices=mats, positions=pos, quaternions=quat, scales=scale)
            return out
        # each instance on its own body: the body pose is the instance's position and rotation
        poses = np.concatenate([pos, quat], axis=-1).astype(np.float32)
        local = np.tile(np.array([0, 0, 0, 0, 0, 0, 1], np.float32), (P, 1))
        # scales are per template instance in a binding: use world 0's
        sc = np.repeat(scale[0][:, None], 3, axis=1).astype(np.float32)
        axis = rng.normal(size=(W, P, 3))
        axis /= np.linalg.norm(axis, axis=-1, keepdims=True)
        half = np.deg2rad(1.0) / 2
        dq = np.concatenate([axis * np.sin(half), np.full((W, P, 1), np.cos(half))], axis=-1)
        target = poses.copy()
        target[..., :3] += rng.normal(0.0, 0.01, size=(W, P, 3))
        target[..., 3:] = _qmul(dq, poses[..., 3:].astype(np.float64))
        out.update(mode="bodies", bodies=P, body_of_instance=np.arange(P, dtype=np.int32), local_pose=local, scale=sc,
                   body_poses=poses, target_poses=target.astype(np.float32))
        return out
    if name == "body_worlds":
        W = worlds or 4096
        P, B = 256, 64
        meshes = three_meshes()
        counts = rng.integers(8, 257, size=W).astype(np.int32)
        boi = (np.arange(P) % B).astype(np.int32)
        local = np.concatenate([rng.uniform(-0.5, 0.5, size=(P, 3)), _quat(rng, P)], axis=1).astype(np.float32)
        sc = rng.uniform(0.5, 1.5, size=(P, 3)).astype(np.float32)
        bpos = (rng.uniform(-5, 5, size=(W, B, 3)) + np.array([0, 0, -10.0])).astype(np.float32)
        poses = np.concatenate([bpos, _quat(rng, W * B).reshape(W, B, 4)], axis=-1).astype(np.float32)
        cams = _cameras_at(rng, bpos.mean(axis=1))
        insts = [Instance(i % 3) for i in range(P)]
        return dict(width=128, height=128, worlds=W, cameras_per_world=1, meshes=meshes, instances=insts, counts=counts,
                    mode="bodies", bodies=B, body_of_instance=boi, local_pose=local, scale=sc, body_poses=poses,
                    cameras=cams, fov=FOV)
    if name == "lidar_sweeps":
        W = worlds or 1024
        R = rays_per_world or 65536
        mesh = make_grid(460, 150, 4.0, 4.0 * 150 / 460, 0.05, seed=11)
        full = 100000
        if W == 1024:
            counts = np.array([98] * 672 + [97] * 352, np.int32)
        else:
            counts = np.full(W, min(98, max(1, full // 1024)), np.int32)
        P = int(counts.max())
        pos, quat, scale, mats = _placed_worlds(rng, W, P, offset=(0.0, 0.0, 0.0), spread=10.0)
        origin = rng.uniform(-1, 1, size=(W, 3)).astype(np.float32)
        beams = 64 if R >= 64 else R
        az_n = R // beams
        el = np.deg2rad(np.linspace(-15.0, 15.0, beams))
        az = 2 * np.pi * np.arange(az_n) / az_n
        e, a = np.meshgrid(el, az, indexing="ij")
        d = np.stack([np.cos(e) * np.sin(a), np.sin(e), -np.cos(e) * np.cos(a)], axis=-1).reshape(-1, 3).astype(np.float32)
        dirs = np.broadcast_to(d, (W, beams * az_n, 3)).copy()
        origins = np.broadcast_to(origin[:, None, :], (W, beams * az_n, 3)).copy()
        insts = [Instance(0) for _ in range(P)]
        return dict(width=64, height=64, worlds=W, cameras_per_world=1, meshes=[mesh], instances=insts, counts=counts,
                    mode="matrices", matrices=mats, origins=origins, directions=dirs, max_distance=100.0, positions=pos,
                    quaternions=quat, scales=scale)
    raise AssertionError(name)


class _Lcg:
    """A 32-bit linear congruential generator (multiplier 1664525, increment 1013904223, start 777) giving
    floats in [0, 1) from its top 24 bits: the stream tests/test_batch.cpp:50-56 of the reference draws from."""

    def __init__(self, state=777):
        self.state = state

    def next(self):
        self.state = (self.state * 1664525 + 1013904223) & 0xFFFFFFFF
        return np.float32(self.state >> 8) / np.float32(1 << 24)


def scattered(worlds, per_world, counts=None, width=96, height=54, spread=None):
    """Boxes (half size 0.4) and icospheres (radius 0.5, one subdivision), instance i using mesh i mod 2, scattered
    through each world as in the reference's large-scene test (tests/test_batch.cpp:338-349): position
    (s (r-0.5), s/2 (r-0.5), -10 - s r) with s = 60 for one world and 3 otherwise, rotation (r-0.5, r-0.5, r-0.5, 1)
    normalized, scale 0.5 + r; every world seen from (0, 0, 5) looking at (0, 0, -30) with a 90 degree field of
    view. `counts` (optional) gives each world's number of instances out of per_world; `spread` overrides s."""
    meshes = [make_box((0.4, 0.4, 0.4)), make_icosphere(0.5, 1)]
    insts = [Instance(i % 2) for i in range(per_world)]
    rng = _Lcg()
    spread = np.float32(spread if spread is not None else (60.0 if worlds == 1 else 3.0))
    half = np.float32(0.5)
    pos = np.empty((worlds * per_world, 3), np.float32)
    rot = np.empty((worlds * per_world, 4), np.float32)
    sc = np.empty(worlds * per_world, np.float32)
    for g in range(worlds * per_world):
        pos[g] = (spread * (rng.next() - half), spread * half * (rng.next() - half), np.float32(-10.0) - spread * rng.next())
        rot[g] = (rng.next() - half, rng.next() - half, rng.next() - half, 1.0)
        sc[g] = half + rng.next()
    mats = make_transforms(pos, rot, sc).reshape(worlds, per_world, 12)
    cam = look_at((0.0, 0.0, 5.0), (0.0, 0.0, -30.0)).reshape(12)
    return dict(width=width, height=height, worlds=worlds, cameras_per_world=1, meshes=meshes, instances=insts,
                counts=None if counts is None else np.asarray(counts, np.int32), mode="matrices", matrices=mats,
                cameras=np.tile(cam, (worlds, 1)), fov=1.5708)


def _qmul(a, b):
    """Hamilton product of quaternions stored x, y, z, w."""
    ax, ay, az, aw = np.moveaxis(a, -1, 0)
    bx, by, bz, bw = np.moveaxis(b, -1, 0)
    return np.stack([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], axis=-1)


def make_renderer(cfg, **kw):
    """A BatchRenderer for a workload dict, its poses applied (host arrays)."""
    from .batch import BatchRenderer
    r = BatchRenderer(cfg["width"], cfg["height"], cfg["worlds"], cfg["cameras_per_world"], meshes=cfg["meshes"],
                      instances=cfg["instances"], instances_in_world=cfg["counts"], **kw)
    if cfg["mode"] == "bodies":
        r.bind_bodies(cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["bodies"])
        r.set_body_poses(cfg["body_poses"])
    else:
        r.set_instance_transforms(cfg["matrices"])
    return r


def falling_bodies(nbody=8, worlds=4096, frames=24, seed=1010):
    """The reference benchmark's scene (see the module docstring): a MuJoCo-like model object (the MjModel fields
    neraulgraphics_amd.simulators.mujoco_shapes reads), per-frame geom frames geom_xpos (frames, worlds, nbody, 3)
    and geom_xmat (frames, worlds, nbody, 9), the camera (3 x 4, camera-to-world), the vertical fov in radians and
    the light position."""
    from types import SimpleNamespace
    from .shapes import look_at, quat_to_matrix
    rng = np.random.default_rng(seed)
    columns, spacing = 8, 0.3
    ball, box = make_icosphere(0.12, 2), make_box((0.1, 0.1, 0.1))
    model = SimpleNamespace(
        ngeom=nbody, geom_type=np.full(nbody, 7), geom_size=np.zeros((nbody, 3)),
        geom_rgba=np.array([[0.8, 0.2, 0.2, 1.0] if i % 2 == 0 else [0.2, 0.4, 0.8, 1.0] for i in range(nbody)], np.float32),
        geom_group=np.zeros(nbody, np.int32), geom_matid=np.full(nbody, -1), mat_rgba=np.zeros((0, 4), np.float32),
        geom_dataid=np.arange(nbody) % 2,
        mesh_vertadr=np.array([0, len(ball.vertices)]), mesh_vertnum=np.array([len(ball.vertices), len(box.vertices)]),
        mesh_faceadr=np.array([0, len(ball.indices) // 3]), mesh_facenum=np.array([len(ball.indices) // 3, len(box.indices) // 3]),
        mesh_vert=np.concatenate([ball.vertices, box.vertices]),
        mesh_face=np.concatenate([ball.indices.reshape(-1, 3), box.indices.reshape(-1, 3)]))
    i = np.arange(nbody)
    start = np.stack([(i % columns - (columns - 1) / 2) * spacing, (i // columns) * spacing - 0.2, 0.3 + 0.02 * i], axis=1)
    p0 = np.broadcast_to(start, (worlds, nbody, 3)).copy()
    p0[..., :2] += rng.uniform(-0.04, 0.04, (worlds, nbody, 2))
    p0[..., 2] += rng.uniform(0.0, 0.08, (worlds, nbody))
    q0 = rng.normal(size=(worlds, nbody, 4))
    q0 /= np.linalg.norm(q0, axis=-1, keepdims=True)
    axis = rng.normal(size=(worlds, nbody, 3))
    axis /= np.linalg.norm(axis, axis=-1, keepdims=True)
    spin = rng.uniform(0.0, 3.0, (worlds, nbody))
    rest = np.where(i % 2 == 0, 0.12, 0.1)
    xpos = np.empty((frames, worlds, nbody, 3), np.float32)
    xmat = np.empty((frames, worlds, nbody, 9), np.float32)
    for f in range(frames):
        t = 0.02 * (f + 1)
        p = p0.copy()
        p[..., 2] = np.maximum(rest, p0[..., 2] - 4.905 * t * t)
        half = 0.5 * spin * t
        dq = np.concatenate([axis * np.sin(half)[..., None], np.cos(half)[..., None]], axis=-1)
        xpos[f] = p
        xmat[f] = quat_to_matrix(_qmul(dq, q0).reshape(-1, 4)).reshape(worlds, nbody, 9)
    rows = (nbody + columns - 1) // columns
    camera = look_at([0.0, -1.9, 1.1], [0.0, (rows - 1) * spacing / 2 - 0.2, 0.15])[0]
    return dict(model=model, geom_xpos=xpos, geom_xmat=xmat, camera=camera, fov=float(np.radians(45.0)),
                light=(0.0, 0.0, 4.0), worlds=worlds, nbody=nbody)
