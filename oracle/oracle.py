"""TEST INFRASTRUCTURE, NOT PRODUCT CODE: ctypes access to the checker.

  Oracle  libngx_oracle.so, the plain-C restatement of the reference's path (ngx_oracle.c; its header
          lists what each function restates and what pins it).
  Ref     oracle/_ref/libngx_ref.so, the reference's own bvh.cpp, mesh.cpp and header-only traversal core
          compiled from its sources (oracle/Makefile); present only where it was built.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, and only as the
checker or the timed CPU baseline. Nothing in neraulgraphics_amd/ imports it."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NGX_ORACLE_LIB may name a sanitizer build of the same source (oracle/Makefile's `sanitize` target)
ORACLE_LIB = os.environ.get("NGX_ORACLE_LIB", os.path.join(HERE, "libngx_oracle.so"))
REF_LIB = os.path.join(HERE, "_ref", "libngx_ref.so")

TLAS_NODE = np.dtype([("lo", "<f4", (3,)), ("left", "<i4"), ("hi", "<f4", (3,)), ("right", "<i4")])
BLAS_NODE = np.dtype([("lo", "<f4", (3,)), ("a", "<i4"), ("hi", "<f4", (3,)), ("count", "<i4")])
VP = ctypes.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(VP)


def build(reference="/root/reference"):
    """Compiles the oracle, and the reference pieces when the reference tree is present."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir(os.path.join(reference, "src")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={reference}"], check=True)


def ref_available():
    return os.path.exists(REF_LIB)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_LIB):
            build()
        self.lib = ctypes.CDLL(ORACLE_LIB)
        L = self.lib
        L.orc_sah_build.restype = ctypes.c_int
        L.orc_sah_cost.restype = ctypes.c_float
        L.orc_build_meshes.restype = ctypes.c_int
        L.orc_spread_bits.restype = ctypes.c_uint32
        L.orc_spread_bits.argtypes = [ctypes.c_uint32]
        L.orc_transform_ok.restype = ctypes.c_int
        L.orc_sah_tlas_create.restype = VP
        L.orc_sah_tlas_nodes.restype = VP
        L.orc_sah_tlas_roots.restype = VP
        L.orc_sah_tlas_node_count.restype = ctypes.c_int
        for name in ("orc_sah_tlas_destroy", "orc_sah_tlas_node_count", "orc_sah_tlas_nodes", "orc_sah_tlas_roots"):
            getattr(L, name).argtypes = [VP]
        L.orc_sah_tlas_update.argtypes = [VP, VP, VP, VP]
        L.orc_render_batch.argtypes = [VP] * 7 + [ctypes.c_int, VP, ctypes.c_int, VP, VP, VP, ctypes.c_float, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int] + [VP] * 5 + [ctypes.c_int]
        L.orc_cast_rays_batch.argtypes = [VP] * 7 + [ctypes.c_int, VP, ctypes.c_int, VP, VP, VP, VP, ctypes.c_int,
                                                     ctypes.c_float] + [VP] * 5 + [ctypes.c_int]
        L.orc_pose_grad.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int] + [VP] * 8 + [ctypes.c_int] + [VP] * 6
        L.orc_render_rgb.argtypes = [VP] * 7 + [ctypes.c_int, VP, ctypes.c_int, VP, VP, VP, VP, VP, ctypes.c_float, ctypes.c_int, VP,
                                                ctypes.c_float, VP, VP, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, VP, ctypes.c_int]

    # ---- bottom level
    def sah_build(self, bounds6, max_leaf):
        bounds6 = np.ascontiguousarray(bounds6, np.float32).reshape(-1, 6)
        n = len(bounds6)
        nodes = np.zeros(max(2 * n, 1), BLAS_NODE)
        order = np.zeros(n, np.uint32)
        k = self.lib.orc_sah_build(_p(bounds6), ctypes.c_int(n), ctypes.c_int(max_leaf), _p(nodes), _p(order))
        return nodes[:k], order

    def sah_cost(self, nodes):
        nodes = np.ascontiguousarray(nodes, BLAS_NODE)
        return float(self.lib.orc_sah_cost(_p(nodes), ctypes.c_int(len(nodes))))

    def sah_refit(self, nodes, order, bounds6):
        nodes = np.array(nodes, BLAS_NODE)
        self.lib.orc_sah_refit(_p(nodes), ctypes.c_int(len(nodes)), _p(np.ascontiguousarray(order, np.uint32)),
                               _p(np.ascontiguousarray(bounds6, np.float32)))
        return nodes

    def build_meshes(self, meshes):
        verts = np.concatenate([m.vertices for m in meshes]).astype(np.float32)
        idx = np.concatenate([m.indices for m in meshes]).astype(np.uint32)
        vo = np.cumsum([0] + [len(m.vertices) for m in meshes]).astype(np.int32)
        io = np.cumsum([0] + [m.indices.size for m in meshes]).astype(np.int32)
        has = np.array([m.normals is not None for m in meshes], np.int32)
        nrm = np.concatenate([m.normals if m.normals is not None else np.zeros_like(m.vertices) for m in meshes]).astype(np.float32) \
            if has.any() else None
        T = int(idx.size // 3)
        nodes = np.zeros(2 * T + len(meshes), BLAS_NODE)
        tris = np.zeros((T, 12), np.float32)
        attrs = np.zeros((T, 16), np.float32) if has.any() else None
        roots = np.zeros(len(meshes), np.int32)
        bounds = np.zeros((len(meshes), 6), np.float32)
        k = self.lib.orc_build_meshes(ctypes.c_int(len(meshes)), _p(verts), _p(vo), _p(idx), _p(io), _p(nrm), _p(has),
                                      _p(nodes), _p(tris), _p(attrs), _p(roots), _p(bounds))
        return dict(nodes=nodes[:k], tris=tris, attrs=attrs, roots=roots, bounds=bounds)

    # ---- poses and the top level
    def invert_affine(self, m12):
        out = np.zeros(12, np.float32)
        self.lib.orc_invert_affine(_p(np.ascontiguousarray(m12, np.float32)), _p(out))
        return out

    def transform_ok(self, m12):
        return bool(self.lib.orc_transform_ok(_p(np.ascontiguousarray(m12, np.float32))))

    def spread_bits(self, v):
        return int(self.lib.orc_spread_bits(ctypes.c_uint32(v)))

    def body_poses_to_matrices(self, world_off, bodies, body_of_instance, local_pose, scale, body_poses):
        world_off = np.ascontiguousarray(world_off, np.int32)
        W = len(world_off) - 1
        out = np.zeros((int(world_off[-1]), 12), np.float32)
        self.lib.orc_body_poses_to_matrices(ctypes.c_int(W), _p(world_off), ctypes.c_int(bodies),
                                            _p(np.ascontiguousarray(body_of_instance, np.int32)),
                                            _p(np.ascontiguousarray(local_pose, np.float32)),
                                            _p(np.ascontiguousarray(scale, np.float32)),
                                            _p(np.ascontiguousarray(body_poses, np.float32)), _p(out))
        return out

    def lbvh_build(self, world_off, matrices, local_bounds):
        """The reference's GPU tree build (cuda_tlas.cuh) over all worlds: matrices [N][12] compact,
        local_bounds [P][6] per template instance."""
        world_off = np.ascontiguousarray(world_off, np.int32)
        W = len(world_off) - 1
        N = int(world_off[-1])
        out = dict(inv=np.zeros((N, 12), np.float32), aabb=np.zeros((N, 6), np.float32),
                   world_bounds=np.zeros((W, 6), np.float32), keys=np.zeros(N, np.uint64), sorted_keys=np.zeros(N, np.uint64),
                   sorted_ids=np.zeros(N, np.int32), nodes=np.zeros(max(2 * N - 1, 1), TLAS_NODE),
                   parent=np.zeros(max(2 * N - 1, 1), np.int32), range=np.zeros((max(N - 1, 1), 2), np.int32),
                   roots=np.zeros(W, np.int32))
        self.lib.orc_lbvh_build(ctypes.c_int(W), _p(world_off), _p(np.ascontiguousarray(matrices, np.float32)),
                                _p(np.ascontiguousarray(local_bounds, np.float32)), _p(out["inv"]), _p(out["aabb"]),
                                _p(out["world_bounds"]), _p(out["keys"]), _p(out["sorted_keys"]), _p(out["sorted_ids"]),
                                _p(out["nodes"]), _p(out["parent"]), _p(out["range"]), _p(out["roots"]))
        if N == 0:
            out["nodes"] = out["nodes"][:0]
        return out

    # ---- tracing
    def render(self, tlas, roots, inv, blas_root, blas, tris, attrs, world_off, cams, fov, width, height, cameras_per_world=1,
               spheres=None, floor=None, threads=None, outputs=("depth", "normal", "id")):
        world_off = np.ascontiguousarray(world_off, np.int32)
        W = len(world_off) - 1
        V = W * cameras_per_world
        px = V * width * height
        depth = np.zeros(px, np.float32) if "depth" in outputs else None
        normal = np.zeros(3 * px, np.float32) if "normal" in outputs else None
        oid = np.zeros(px, np.int32) if "id" in outputs else None
        hi = np.zeros(px, np.int32)
        ht = np.zeros(px, np.int32)
        sph = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4) if spheres is not None and len(spheres) else None
        fl = np.ascontiguousarray(floor, np.float32) if floor is not None else None
        self.lib.orc_render_batch(_p(np.ascontiguousarray(tlas, TLAS_NODE)), _p(np.ascontiguousarray(roots, np.int32)),
                                  _p(np.ascontiguousarray(inv, np.float32)), _p(np.ascontiguousarray(blas_root, np.int32)),
                                  _p(np.ascontiguousarray(blas, BLAS_NODE)), _p(np.ascontiguousarray(tris, np.float32)),
                                  _p(attrs), ctypes.c_int(W), _p(world_off),
                                  ctypes.c_int(0 if sph is None else len(sph)), _p(sph), _p(fl),
                                  _p(np.ascontiguousarray(cams, np.float32)), ctypes.c_float(_tanf_half(fov)),
                                  ctypes.c_int(width), ctypes.c_int(height), ctypes.c_int(cameras_per_world),
                                  _p(depth), _p(normal), _p(oid), _p(hi), _p(ht), ctypes.c_int(threads or os.cpu_count()))
        shape = (V, height, width)
        return dict(depth=None if depth is None else depth.reshape(shape),
                    normal=None if normal is None else normal.reshape(shape + (3,)),
                    object_id=None if oid is None else oid.reshape(shape), hit_inst=hi.reshape(shape), hit_tri=ht.reshape(shape))

    def render_rgb(self, tlas, roots, inv, blas_root, blas, tris, attrs, world_off, cams, fov, width, height, shading,
                   cameras_per_world=1, spp=1, frame=0, spheres=None, floor=None, threads=None):
        """The classic RGB picture (render_core.h:146-283) as (views, H, W, 3) bytes; shading as from
        neraulgraphics_amd.geometry.shading_arrays."""
        world_off = np.ascontiguousarray(world_off, np.int32)
        W = len(world_off) - 1
        V = W * cameras_per_world
        rgb = np.zeros(V * width * height * 3, np.uint8)
        sph = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4) if spheres is not None and len(spheres) else None
        fl = np.ascontiguousarray(floor, np.float32) if floor is not None else None
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        im, sm, fm = f32(shading["instance_materials"]), f32(shading["sphere_materials"]), f32(shading["floor_material"])
        lights, sky = f32(shading["lights"]), f32(shading["sky"])
        self.lib.orc_render_rgb(_p(np.ascontiguousarray(tlas, TLAS_NODE)), _p(np.ascontiguousarray(roots, np.int32)),
                                _p(np.ascontiguousarray(inv, np.float32)), _p(np.ascontiguousarray(blas_root, np.int32)),
                                _p(np.ascontiguousarray(blas, BLAS_NODE)), _p(np.ascontiguousarray(tris, np.float32)),
                                _p(attrs), ctypes.c_int(W), _p(world_off),
                                ctypes.c_int(0 if sph is None else len(sph)), _p(sph), _p(fl),
                                _p(im), _p(sm), _p(fm), ctypes.c_float(shading["floor_checker_scale"]),
                                ctypes.c_int(len(lights)), _p(lights), ctypes.c_float(shading["ambient"]), _p(sky),
                                _p(np.ascontiguousarray(cams, np.float32)), ctypes.c_float(_tanf_half(fov)),
                                ctypes.c_int(width), ctypes.c_int(height), ctypes.c_int(cameras_per_world), ctypes.c_int(spp),
                                ctypes.c_int(frame), _p(rgb), ctypes.c_int(threads or os.cpu_count()))
        return rgb.reshape(V, height, width, 3)

    def cast_rays(self, tlas, roots, inv, blas_root, blas, tris, attrs, world_off, origins, directions, max_distance,
                  spheres=None, floor=None, threads=None):
        world_off = np.ascontiguousarray(world_off, np.int32)
        W = len(world_off) - 1
        origins = np.ascontiguousarray(origins, np.float32).reshape(W, -1, 3)
        R = origins.shape[1]
        n = W * R
        dist = np.zeros(n, np.float32)
        normal = np.zeros(3 * n, np.float32)
        oid = np.zeros(n, np.int32)
        hi = np.zeros(n, np.int32)
        ht = np.zeros(n, np.int32)
        sph = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4) if spheres is not None and len(spheres) else None
        fl = np.ascontiguousarray(floor, np.float32) if floor is not None else None
        self.lib.orc_cast_rays_batch(_p(np.ascontiguousarray(tlas, TLAS_NODE)), _p(np.ascontiguousarray(roots, np.int32)),
                                     _p(np.ascontiguousarray(inv, np.float32)), _p(np.ascontiguousarray(blas_root, np.int32)),
                                     _p(np.ascontiguousarray(blas, BLAS_NODE)), _p(np.ascontiguousarray(tris, np.float32)),
                                     _p(attrs), ctypes.c_int(W), _p(world_off), ctypes.c_int(0 if sph is None else len(sph)),
                                     _p(sph), _p(fl), _p(origins), _p(np.ascontiguousarray(directions, np.float32)),
                                     ctypes.c_int(R), ctypes.c_float(max_distance), _p(dist), _p(normal), _p(oid), _p(hi), _p(ht),
                                     ctypes.c_int(threads or os.cpu_count()))
        return dict(distance=dist.reshape(W, R), normal=normal.reshape(W, R, 3), object_id=oid.reshape(W, R),
                    hit_inst=hi.reshape(W, R), hit_tri=ht.reshape(W, R))

    def camera_ray(self, cam12, fov, width, height, x, y):
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        f = np.zeros(3, np.float32)
        self.lib.orc_camera_ray(_p(np.ascontiguousarray(cam12, np.float32)), ctypes.c_float(_tanf_half(fov)), ctypes.c_int(width),
                                ctypes.c_int(height), ctypes.c_int(x), ctypes.c_int(y), _p(o), _p(d), _p(f))
        return o, d, f

    def pose_grad(self, mode, h, origins, dirs, weight, out_scale, hit_inst, hit_tri, ray_world, world_off, bodies,
                  body_of_instance, local_pose, scale, body_poses, tris):
        W = len(world_off) - 1
        grad = np.zeros((W, bodies, 7), np.float64)
        n = len(weight)
        self.lib.orc_pose_grad(ctypes.c_int(mode), ctypes.c_double(h), ctypes.c_int(n),
                               _p(np.ascontiguousarray(origins, np.float64)), _p(np.ascontiguousarray(dirs, np.float64)),
                               _p(np.ascontiguousarray(weight, np.float64)), _p(np.ascontiguousarray(out_scale, np.float64)),
                               _p(np.ascontiguousarray(hit_inst, np.int32)), _p(np.ascontiguousarray(hit_tri, np.int32)),
                               _p(np.ascontiguousarray(ray_world, np.int32)), _p(np.ascontiguousarray(world_off, np.int32)),
                               ctypes.c_int(bodies), _p(np.ascontiguousarray(body_of_instance, np.int32)),
                               _p(np.ascontiguousarray(local_pose, np.float32)), _p(np.ascontiguousarray(scale, np.float32)),
                               _p(np.ascontiguousarray(body_poses, np.float32)), _p(np.ascontiguousarray(tris, np.float32)),
                               _p(grad))
        return grad


class SahTlas:
    """The reference CPU backend's per-world SAH top level (instances.cpp:31-121), kept across updates."""

    def __init__(self, oracle, world_off):
        self.o = oracle
        self.world_off = np.ascontiguousarray(world_off, np.int32)
        self.h = VP(oracle.lib.orc_sah_tlas_create(ctypes.c_int(len(world_off) - 1), _p(self.world_off)))

    def update(self, matrices, local_bounds):
        N = int(self.world_off[-1])
        inv = np.zeros((N, 12), np.float32)
        self.o.lib.orc_sah_tlas_update(self.h, _p(np.ascontiguousarray(matrices, np.float32)),
                                       _p(np.ascontiguousarray(local_bounds, np.float32)), _p(inv))
        n = self.o.lib.orc_sah_tlas_node_count(self.h)
        nodes = np.ctypeslib.as_array(ctypes.cast(self.o.lib.orc_sah_tlas_nodes(self.h), ctypes.POINTER(ctypes.c_uint8)),
                                      shape=(n * 32,)).view(TLAS_NODE).copy() if n else np.zeros(0, TLAS_NODE)
        W = len(self.world_off) - 1
        roots = np.ctypeslib.as_array(ctypes.cast(self.o.lib.orc_sah_tlas_roots(self.h), ctypes.POINTER(ctypes.c_int32)),
                                      shape=(W,)).copy()
        return nodes, roots, inv

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.orc_sah_tlas_destroy(self.h)
            self.h = None


def _tanf_half(fov):
    """tanf(fov * 0.5f) as the reference's host computes it (glibc tanf), via the C library."""
    libm = ctypes.CDLL("libm.so.6")
    libm.tanf.restype = ctypes.c_float
    libm.tanf.argtypes = [ctypes.c_float]
    return float(libm.tanf(np.float32(np.float32(fov) * np.float32(0.5))))


class Ref:
    """The reference's own code, compiled from its sources into oracle/_ref."""

    def __init__(self):
        self.lib = ctypes.CDLL(REF_LIB)
        self.lib.ref_sah_build.restype = ctypes.c_int
        self.lib.ref_sah_refit_cost.restype = ctypes.c_float

    def sah_build(self, bounds6, max_leaf):
        bounds6 = np.ascontiguousarray(bounds6, np.float32).reshape(-1, 6)
        n = len(bounds6)
        nodes = np.zeros(max(2 * n, 1), BLAS_NODE)
        order = np.zeros(n, np.uint32)
        k = self.lib.ref_sah_build(_p(bounds6), ctypes.c_int(n), ctypes.c_int(max_leaf), _p(nodes), _p(order))
        return nodes[:k], order

    def sah_refit_cost(self, nodes, order, bounds6):
        nodes = np.array(nodes, BLAS_NODE)
        order = np.ascontiguousarray(order, np.uint32)
        c = self.lib.ref_sah_refit_cost(_p(nodes), ctypes.c_int(len(nodes)), _p(order), ctypes.c_int(len(order)),
                                        _p(np.ascontiguousarray(bounds6, np.float32)))
        return nodes, float(c)

    def _mesh(self, fn, *args):
        nv, ni = ctypes.c_int(), ctypes.c_int()
        fn(*args, None, None, ctypes.byref(nv), ctypes.byref(ni))
        v = np.zeros((nv.value, 3), np.float32)
        i = np.zeros(ni.value, np.uint32)
        fn(*args, _p(v), _p(i), ctypes.byref(nv), ctypes.byref(ni))
        return v, i

    def make_icosphere(self, radius, subdivisions):
        return self._mesh(self.lib.ref_make_icosphere, ctypes.c_float(radius), ctypes.c_int(subdivisions))

    def make_box(self, hx, hy, hz):
        return self._mesh(self.lib.ref_make_box, ctypes.c_float(hx), ctypes.c_float(hy), ctypes.c_float(hz))

    def make_transform(self, position, quat, scale):
        out = np.zeros(12, np.float32)
        self.lib.ref_make_transform(_p(np.ascontiguousarray(position, np.float32)), _p(np.ascontiguousarray(quat, np.float32)),
                                    ctypes.c_float(scale), _p(out))
        return out

    def camera_rays(self, cams, fov, width, height):
        cams = np.ascontiguousarray(cams, np.float32).reshape(-1, 12)
        V = len(cams)
        o = np.zeros((V, height, width, 3), np.float32)
        d = np.zeros((V, height, width, 3), np.float32)
        f = np.zeros((V, 3), np.float32)
        self.lib.ref_camera_rays(_p(cams), ctypes.c_int(V), ctypes.c_float(fov), ctypes.c_int(width), ctypes.c_int(height),
                                 _p(o), _p(d), _p(f))
        return o, d, f

    def mesh_intersect(self, tlas, roots, inv, blas_root, blas, tris, ray_world, origins, dirs, tmax=None):
        n = len(ray_world)
        tmax = np.full(n, 1e30, np.float32) if tmax is None else np.ascontiguousarray(tmax, np.float32)
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        self.lib.ref_mesh_intersect(_p(np.ascontiguousarray(tlas, TLAS_NODE)), _p(np.ascontiguousarray(roots, np.int32)),
                                    _p(np.ascontiguousarray(inv, np.float32)), _p(np.ascontiguousarray(blas_root, np.int32)),
                                    ctypes.c_int(len(blas_root)), _p(np.ascontiguousarray(blas, BLAS_NODE)),
                                    _p(np.ascontiguousarray(tris, np.float32)), ctypes.c_int(n),
                                    _p(np.ascontiguousarray(ray_world, np.int32)), _p(np.ascontiguousarray(origins, np.float32)),
                                    _p(np.ascontiguousarray(dirs, np.float32)), _p(tmax), _p(t), _p(inst), _p(tri), _p(u), _p(v))
        return dict(t=t, inst=inst, tri=tri, u=u, v=v)


def forward(cfg, oracle=None, tlas="lbvh", threads=None, body_poses=None, matrices=None, outputs=("depth", "normal", "id")):
    """The whole path on the CPU for a workload dict (neraulgraphics_amd.workloads): meshes, instance matrices
    (FP64 body composition or the given matrices), the top level (the reference GPU algorithm 'lbvh' or the
    reference CPU backend's 'sah'), then the render. Returns the render dict plus the intermediate state."""
    o = oracle or Oracle()
    W, P = cfg["worlds"], len(cfg["instances"])
    counts = np.full(W, P, np.int32) if cfg["counts"] is None else np.asarray(cfg["counts"], np.int32)
    world_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    meshes = o.build_meshes(cfg["meshes"])
    tmpl_mesh = np.array([i.mesh for i in cfg["instances"]], np.int32)
    local_bounds = meshes["bounds"][tmpl_mesh]
    if cfg["mode"] == "bodies":
        poses = cfg["body_poses"] if body_poses is None else body_poses
        mats = o.body_poses_to_matrices(world_off, cfg["bodies"], cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], poses)
    else:
        full = cfg["matrices"] if matrices is None else matrices
        full = np.asarray(full, np.float32).reshape(W, P, 12)
        mats = np.concatenate([full[w, :counts[w]] for w in range(W)]) if counts.sum() else np.zeros((0, 12), np.float32)
    tmpl_of = np.concatenate([np.arange(c) for c in counts]).astype(np.int32) if counts.sum() else np.zeros(0, np.int32)
    blas_root = meshes["roots"][tmpl_mesh][tmpl_of]
    if tlas == "lbvh":
        st = o.lbvh_build(world_off, mats, local_bounds)
        nodes, roots, inv = st["nodes"], st["roots"], st["inv"]
    else:
        st = {}
        nodes, roots, inv = SahTlas(o, world_off).update(mats, local_bounds)
    out = {}
    spheres, floor = cfg.get("spheres"), cfg.get("floor")
    aov = tuple(k for k in outputs if k != "rgb")
    if "cameras" in cfg and aov:
        out = o.render(nodes, roots, inv, blas_root, meshes["nodes"], meshes["tris"], meshes["attrs"], world_off, cfg["cameras"],
                       cfg["fov"], cfg["width"], cfg["height"], cfg["cameras_per_world"], spheres=spheres, floor=floor,
                       threads=threads, outputs=aov)
    if "cameras" in cfg and "rgb" in outputs:
        out["rgb"] = o.render_rgb(nodes, roots, inv, blas_root, meshes["nodes"], meshes["tris"], meshes["attrs"], world_off,
                                  cfg["cameras"], cfg["fov"], cfg["width"], cfg["height"], cfg["shading"],
                                  cameras_per_world=cfg["cameras_per_world"], spp=cfg.get("spp", 1), frame=cfg.get("frame", 0),
                                  spheres=spheres, floor=floor, threads=threads)
    return dict(out, state=st, nodes=nodes, roots=roots, inv=inv, matrices=mats, world_off=world_off, meshes=meshes,
                blas_root=blas_root, local_bounds=local_bounds)
