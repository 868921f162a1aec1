"""BatchRenderer: the Python edge of the MI355X backend, with the surface of the reference's
ng.BatchRenderer (python/bindings.cpp:667-800): set_instance_transforms, bind_bodies, set_body_poses,
render, cast_rays and the same properties. Inputs may be NumPy arrays (checked, copied to the device) or
torch tensors on the GPU (used where they are). Device work is queued on torch's current stream.

Differences from the reference, all deliberate:
  * the RGB picture is the classic one (materials, sphere lights, mirror reflections, sky gradient); textures and
    environment maps are not supported (NotImplementedError), and rgb defaults to False here;
  * instances_in_world gives every world its own count of the template's first instances;
  * pose_backward_depth / pose_backward_range give gradients with respect to body poses.
Errors: ValueError for bad input (the reference's std::invalid_argument), RuntimeError for device failures.
Threads: one renderer may be shared; its calls are serialized by a per-renderer lock, and the library call itself
runs without the GIL (ctypes releases it), as the reference's edge does (python/bindings.cpp:242-251, 287-289)."""
import ctypes
import functools
import threading

import numpy as np

from . import _lib
from .dlpack import to_torch
from .geometry import Environment, Instance, Mesh, Plane, Sphere, shading_arrays

_T = None


def _torch():
    global _T
    if _T is None:
        import torch
        _T = torch
    return _T


def _is_device_tensor(a):
    t = _torch() if not isinstance(a, np.ndarray) else None
    return t is not None and isinstance(a, t.Tensor) and a.is_cuda


def _stream():
    t = _torch()
    return ctypes.c_void_p(t.cuda.current_stream().cuda_stream)


def _as_input(a, what, count=None, dtype=np.float32):
    """(pointer, memory, keepalive) for a NumPy array or a device tensor of float32 values."""
    if _is_device_tensor(a):
        t = _torch()
        want = t.float32 if dtype == np.float32 else t.int32
        if a.dtype != want:
            raise ValueError(f"{what} must be {np.dtype(dtype).name}, got {a.dtype}")
        a = a.contiguous()
        if count is not None and a.numel() != count:
            raise ValueError(f"{what} has {a.numel()} values, but {count} are needed")
        return ctypes.c_void_p(a.data_ptr()), _lib.NGX_MEM_DEVICE, a
    arr = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    if count is not None and arr.size != count:
        raise ValueError(f"{what} has {arr.size} values, but {count} are needed")
    return arr.ctypes.data_as(ctypes.c_void_p), _lib.NGX_MEM_HOST, arr


def _device_operand(a, what, count, device_id):
    """`a` as a contiguous float32 torch tensor of `count` values on GPU `device_id`, for the entry points that take
    device pointers only (the pose backward); anything else is refused with ValueError before any device work."""
    t = _torch()
    if not isinstance(a, t.Tensor):
        raise ValueError(f"{what} must be a torch tensor on the GPU, got {type(a).__name__}")
    if not a.is_cuda or a.device.index != device_id:
        raise ValueError(f"{what} must be on the renderer's GPU (cuda:{device_id}), got {a.device}")
    if a.dtype != t.float32:
        raise ValueError(f"{what} must be float32, got {a.dtype}")
    if a.numel() != count:
        raise ValueError(f"{what} has {a.numel()} values, but {count} are needed")
    return a.contiguous()


def _serialized(method):
    """Runs the method under the renderer's lock: the backend context is not thread-safe."""
    @functools.wraps(method)
    def run(self, *args, **kwargs):
        with self._mutex:
            return method(self, *args, **kwargs)
    return run


class BatchRenderer:
    """Many worlds rendered at once: thousands of copies of one environment, each seen by its own camera(s).

    All worlds share the meshes, spheres and floor. Each holds the template's instances (its first
    instances_in_world[w] of them), placed differently, and a camera sees only its own world's. Views are
    numbered world-major: view v looks at world v // cameras_per_world."""

    def __init__(self, width, height, world_count, cameras_per_world=1, spheres=(), lights=(), floor=None, meshes=(),
                 instances=(), textures=(), device="auto", instances_in_world=None, device_id=None):
        self._mutex = threading.RLock()
        lib = _lib.load()
        if device not in ("auto", "cuda", "gpu"):
            raise ValueError(f"this backend runs on the GPU; device {device!r} is not one it serves")
        if textures:
            raise NotImplementedError("textured meshes are not supported by this backend's RGB picture")
        floor = floor if floor is not None else Plane()
        spheres = [s if isinstance(s, Sphere) else Sphere(*s) for s in spheres]
        meshes = [m if isinstance(m, Mesh) else Mesh(*m) for m in meshes]
        instances = [i if isinstance(i, Instance) else Instance(*i) for i in instances]
        if device_id is None:
            try:
                device_id = _torch().cuda.current_device()
            except Exception:
                device_id = 0
        self._keep = []
        vert = np.concatenate([m.vertices for m in meshes]).astype(np.float32) if meshes else np.zeros((1, 3), np.float32)
        vo = np.zeros(len(meshes) + 1, np.int32)
        io = np.zeros(len(meshes) + 1, np.int32)
        for k, m in enumerate(meshes):
            vo[k + 1] = vo[k] + len(m.vertices)
            io[k + 1] = io[k] + m.indices.size
        idx = np.concatenate([m.indices for m in meshes]).astype(np.uint32) if meshes else np.zeros(1, np.uint32)
        has_n = np.array([m.normals is not None for m in meshes] or [0], np.uint8)
        nrm = np.concatenate([m.normals if m.normals is not None else np.zeros_like(m.vertices) for m in meshes]) \
            if meshes and has_n.any() else None
        inst_mesh = np.array([i.mesh for i in instances] or [0], np.int32)
        inst_tf = np.concatenate([i.transform.matrix.reshape(1, 12) for i in instances]).astype(np.float32) \
            if instances else np.zeros((1, 12), np.float32)
        sph = np.array([[*s.center, s.radius] for s in spheres], np.float32).reshape(-1, 4) if spheres else np.zeros((1, 4), np.float32)
        counts = None
        if instances_in_world is not None:
            counts = np.ascontiguousarray(np.asarray(instances_in_world, dtype=np.int32))
            if counts.shape != (world_count,):
                raise ValueError(f"instances_in_world must hold one count for each of the {world_count} worlds")
        self._keep += [vert, vo, io, idx, has_n, nrm, inst_mesh, inst_tf, sph, counts]

        def ptr(a, t):
            return a.ctypes.data_as(t) if a is not None else ctypes.cast(None, t)

        desc = _lib.SceneDesc(
            width, height, world_count, cameras_per_world,
            len(meshes), ptr(vert, _lib.c_float_p), ptr(vo, _lib.c_int32_p), ptr(idx, _lib.c_uint32_p),
            ptr(io, _lib.c_int32_p), ptr(nrm, _lib.c_float_p), ptr(has_n, _lib.c_uint8_p),
            len(instances), ptr(inst_mesh, _lib.c_int32_p), ptr(inst_tf, _lib.c_float_p), ptr(counts, _lib.c_int32_p),
            len(spheres), ptr(sph, _lib.c_float_p),
            float(floor.y), float(floor.half_extent_x), float(floor.z_min), float(floor.z_max),
            int(device_id))
        handle = ctypes.c_void_p()
        _lib.check(lib.ngx_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._h = handle
        self._lib = lib
        self._width, self._height = width, height
        self._worlds, self._cpw = world_count, cameras_per_world
        self._per_world = len(instances)
        self._mesh_vertex_counts = [len(m.vertices) for m in meshes]
        self._mesh_signature = ",".join(str(len(m.indices) // 3 if hasattr(m, "indices") else 0) for m in meshes)
        self._bodies = None
        self._geoms = None
        self._device_id = int(device_id)
        self._keep = []
        self._shading = shading_arrays(instances, spheres, floor, list(lights))
        self._apply_shading()

    def _apply_shading(self):
        """Hands the materials, lights, ambient level and sky to the backend (ngx_set_shading)."""
        sh = self._shading
        im = np.ascontiguousarray(sh["instance_materials"], np.float32)
        sm = np.ascontiguousarray(sh["sphere_materials"], np.float32)
        lights = np.ascontiguousarray(sh["lights"], np.float32)

        def material(row):
            return _lib.Material((ctypes.c_float * 3)(*row[:3]), *[float(v) for v in row[3:7]])

        # the C structs are laid out as the flat arrays without their padding float
        im7 = np.ascontiguousarray(im[:, :7]) if len(im) else np.zeros((1, 7), np.float32)
        sm7 = np.ascontiguousarray(sm[:, :7]) if len(sm) else np.zeros((1, 7), np.float32)
        l14 = lights if len(lights) else np.zeros((1, 14), np.float32)
        sky = sh["sky"]
        desc = _lib.Shading(im7.ctypes.data, sm7.ctypes.data, material(sh["floor_material"]), float(sh["floor_checker_scale"]),
                            len(lights), l14.ctypes.data, float(sh["ambient"]), (ctypes.c_float * 3)(*sky[0:3]),
                            (ctypes.c_float * 3)(*sky[3:6]), (ctypes.c_float * 3)(*sky[6:9]), float(sky[9]))
        _lib.check(self._lib.ngx_set_shading(self._h, ctypes.byref(desc)))

    # ---- the RGB picture's lighting (include/neraulgraphics/batch.h:206-212)
    @_serialized
    def set_ambient(self, ambient):
        old = self._shading["ambient"]
        self._shading["ambient"] = float(ambient)
        try:
            self._apply_shading()
        except Exception:
            self._shading["ambient"] = old
            raise

    @property
    def ambient(self):
        return self._shading["ambient"]

    @_serialized
    def set_environment(self, environment):
        """The sky gradient (an Environment); maps are not supported."""
        if getattr(environment, "map", None):
            raise NotImplementedError("environment maps are not supported by this backend's RGB picture")
        old = self._shading["sky"]
        self._shading["sky"] = (environment if isinstance(environment, Environment) else Environment(**environment)).as_array()
        try:
            self._apply_shading()
        except Exception:
            self._shading["sky"] = old
            raise

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.ngx_destroy(h)
            self._h = None

    # ---- properties (python/bindings.cpp:787-794)
    width = property(lambda self: self._width)
    height = property(lambda self: self._height)
    world_count = property(lambda self: self._worlds)
    cameras_per_world = property(lambda self: self._cpw)
    view_count = property(lambda self: self._worlds * self._cpw)
    instances_per_world = property(lambda self: self._per_world)
    device = property(lambda self: "cuda")
    device_id = property(lambda self: self._device_id)
    bodies_per_world = property(lambda self: self._bodies)

    @property
    def instance_count(self):
        return int(self._lib.ngx_instance_count(self._h))

    @property
    def last_kernel_ms(self):
        return float(self._lib.ngx_last_kernel_ms(self._h))

    @_serialized
    def synchronize(self):
        _lib.check(self._lib.ngx_synchronize(self._h))

    @_serialized
    def set_trace_config(self, tile_w=8, xcd_remap=True, packet=True, pair_limit=64, persistent=True, lane_min_triangles=300,
                         sort_rays=True, render_blocks=0, rgb_blocks=0, ray_blocks=0, backward_blocks=0, rgb_two_pass=True):
        """Tile shape of a wavefront (tile_w x 64/tile_w pixels), XCD-contiguous tile order, packet (True, the
        default) or per-lane (False) walks, (ray, triangle) pairs batched before testing (1..64), the
        persistent fixed-grid launch, the triangle count from which a mesh's bottom level is walked lane by lane
        inside the packet walk, whether ray queries trace each world's rays in an order of their directions, the
        kernel variants by workgroups per compute unit (0: the walk's default), and whether the RGB picture of a
        scene without mirror bounces is drawn in two passes (see ngx_trace_config)."""
        c = _lib.TraceConfig(int(tile_w), int(bool(xcd_remap)), int(bool(packet)), int(pair_limit), int(bool(persistent)),
                             int(lane_min_triangles), int(bool(sort_rays)), int(render_blocks), int(rgb_blocks),
                             int(ray_blocks), int(backward_blocks), int(bool(rgb_two_pass)))
        _lib.check(self._lib.ngx_set_trace_config(self._h, ctypes.byref(c)))

    @property
    def trace_config(self):
        c = _lib.TraceConfig()
        _lib.check(self._lib.ngx_get_trace_config(self._h, ctypes.byref(c)))
        return c.as_dict()

    @_serialized
    def autotune(self, cameras, fov=1.0, *, depth=True, normals=False, object_ids=True, rgb=False, reps=3, cache=True):
        """Times every trace configuration worth trying on these cameras and keeps the fastest (see
        ngx_autotune). With cache, the choice is remembered in a JSON file (NGX_TUNE_CACHE, default
        ~/.cache/neraulgraphics_amd/autotune.json) under a key made of the device, the frame, the scene's sizes
        and the outputs, and reused without timing. Returns the configuration with its time per render (ms;
        None when it came from the cache)."""
        import json
        import os
        outputs = (_lib.OUTPUT_DEPTH if depth else 0) | (_lib.OUTPUT_NORMAL if normals else 0) | \
                  (_lib.OUTPUT_OBJECT_ID if object_ids else 0) | (_lib.OUTPUT_RGB if rgb else 0)
        path = os.environ.get("NGX_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "neraulgraphics_amd", "autotune.json"))
        key = None
        if cache:
            try:
                device = _torch().cuda.get_device_name(self._device_id)
            except Exception:
                device = "device"
            key = "|".join(str(x) for x in (device, self._width, self._height, self._worlds, self._cpw, self._per_world,
                                            self.instance_count, self._mesh_signature, outputs))
            try:
                with open(path) as fh:
                    known = json.load(fh)
                if key in known:
                    self.set_trace_config(**known[key])
                    return dict(known[key], ms=None)
            except (OSError, ValueError, TypeError):
                pass
        views = self.view_count
        p, mem, keep = _as_input(cameras, f"cameras ({views} cameras of 3 x 4 numbers, camera-to-world)", views * 12)
        best, ms = _lib.TraceConfig(), ctypes.c_float()
        _lib.check(self._lib.ngx_autotune(self._h, p, mem, float(fov), outputs, int(reps), self._stream_for(mem),
                                          ctypes.byref(best), ctypes.byref(ms)))
        chosen = best.as_dict()
        if cache:
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                try:
                    with open(path) as fh:
                        known = json.load(fh)
                except (OSError, ValueError):
                    known = {}
                known[key] = chosen
                with open(path + ".tmp", "w") as fh:
                    json.dump(known, fh, indent=1, sort_keys=True)
                os.replace(path + ".tmp", path)
            except OSError:
                pass
        return dict(chosen, ms=float(ms.value))

    @_serialized
    def autotune_rays(self, origins, directions, max_distance=100.0, reps=3):
        """Times ray queries on these rays with the direction order on and off and every variant of the ray kernel
        (ngx_autotune_rays), keeps the fastest and returns the configuration with its time per query (ms)."""
        w = self._worlds
        n = origins.numel() if _is_device_tensor(origins) else np.asarray(origins).size
        if n % (3 * w) != 0:
            raise ValueError(f"origins must hold 3 floats for each ray of each of the {w} worlds")
        r = n // (3 * w)
        po, mo, ko = _as_input(origins, "origins (world_count, rays_per_world, 3)", 3 * w * r)
        pd, md, kd = _as_input(directions, "directions (world_count, rays_per_world, 3)", 3 * w * r)
        if mo != md:
            raise ValueError("origins and directions must both be on the host or both on the GPU")
        best, ms = _lib.TraceConfig(), ctypes.c_float()
        _lib.check(self._lib.ngx_autotune_rays(self._h, po, pd, int(r), float(max_distance), mo, int(reps), self._stream_for(mo),
                                               ctypes.byref(best), ctypes.byref(ms)))
        return dict(best.as_dict(), ms=float(ms.value))

    @_serialized
    def autotune_backward(self, body_poses, cameras, fov, depth, grad_depth, reps=3):
        """Times the depth backward on these arguments (as pose_backward_depth takes them) with every variant of
        its kernel (ngx_autotune_backward), keeps the fastest and returns the configuration with its time (ms)."""
        t = _torch()
        if self._bodies is None:
            raise ValueError("the pose backward needs bind_bodies first")
        views, dev = self.view_count, self._device_id
        pixels = views * self._height * self._width
        poses = _device_operand(body_poses, "body_poses", self._worlds * self._bodies * 7, dev)
        cams = _device_operand(cameras, "cameras", views * 12, dev)
        d = _device_operand(depth, "depth", pixels, dev)
        gd = _device_operand(grad_depth, "grad_depth", pixels, dev)
        grad = t.empty((self._worlds, self._bodies, 7), dtype=t.float32, device=poses.device)
        best, ms = _lib.TraceConfig(), ctypes.c_float()
        _lib.check(self._lib.ngx_autotune_backward(self._h, poses.data_ptr(), cams.data_ptr(), float(fov), d.data_ptr(),
                                                   gd.data_ptr(), grad.data_ptr(), int(reps), _stream(), ctypes.byref(best),
                                                   ctypes.byref(ms)))
        return dict(best.as_dict(), ms=float(ms.value))

    # ---- poses
    @_serialized
    def set_instance_transforms(self, matrices):
        """(world_count, instances_per_world, 3, 4) float32: the top three rows of each instance's 4x4 transform."""
        n = self._worlds * self._per_world * 12
        p, mem, keep = _as_input(matrices, "matrices (world_count, instances_per_world, 3, 4)", n)
        _lib.check(self._lib.ngx_set_instance_matrices(self._h, p, mem, self._stream_for(mem)))
        self._hold(keep, mem)

    @_serialized
    def bind_bodies(self, body_of_instance, local_pose, scale, bodies_per_world):
        boi = np.ascontiguousarray(np.asarray(body_of_instance, dtype=np.int32).reshape(-1))
        lp = np.ascontiguousarray(np.asarray(local_pose, dtype=np.float32).reshape(-1))
        sc = np.ascontiguousarray(np.asarray(scale, dtype=np.float32).reshape(-1))
        n = self._per_world
        if boi.size != n or lp.size != 7 * n or sc.size != 3 * n:
            raise ValueError(f"a body binding needs one body index, 7 pose floats and 3 scale floats for each of the {n} "
                             "instances of a world")
        _lib.check(self._lib.ngx_bind_bodies(self._h, int(bodies_per_world), boi.ctypes.data_as(ctypes.c_void_p),
                                             lp.ctypes.data_as(ctypes.c_void_p), sc.ctypes.data_as(ctypes.c_void_p)))
        self._bodies = int(bodies_per_world)

    @_serialized
    def set_body_poses(self, body_poses):
        """(world_count, bodies_per_world, 7) float32: x, y, z, then a quaternion x, y, z, w."""
        if self._bodies is None:
            raise ValueError("set_body_poses needs bind_bodies to be called first")
        n = self._worlds * self._bodies * 7
        p, mem, keep = _as_input(body_poses, "body poses (world_count, bodies_per_world, 7)", n)
        _lib.check(self._lib.ngx_set_body_poses(self._h, p, mem, self._stream_for(mem)))
        self._hold(keep, mem)

    @_serialized
    def bind_geoms(self, geom_of_instance, scale, geoms_per_world):
        """Binds each template instance to a simulator geom of its world (the MuJoCo Warp flow,
        python/neraulgraphics/mujoco.py:177-302): instance i is drawn at geom geom_of_instance[i]'s world frame,
        scaled by scale[i] along the mesh's own axes. Independent of bind_bodies."""
        goi = np.ascontiguousarray(np.asarray(geom_of_instance, dtype=np.int32).reshape(-1))
        sc = np.ascontiguousarray(np.asarray(scale, dtype=np.float32).reshape(-1))
        n = self._per_world
        if goi.size != n or sc.size != 3 * n:
            raise ValueError(f"a geom binding needs one geom index and 3 scale floats for each of the {n} instances of a world")
        _lib.check(self._lib.ngx_bind_geoms(self._h, int(geoms_per_world), goi.ctypes.data_as(ctypes.c_void_p),
                                            sc.ctypes.data_as(ctypes.c_void_p)))
        self._geoms = int(geoms_per_world)

    @_serialized
    def set_geom_frames(self, geom_xpos, geom_xmat):
        """Places every instance from its geom's world frame and rebuilds the top level, one kernel: geom_xpos
        (world_count, geoms_per_world, 3) and geom_xmat (world_count, geoms_per_world, 3, 3) or (..., 9), float32,
        row-major, as MuJoCo Warp's data.geom_xpos and data.geom_xmat hold them. Both NumPy or both GPU arrays."""
        if self._geoms is None:
            raise ValueError("set_geom_frames needs bind_geoms to be called first")
        n = self._worlds * self._geoms
        p, mem, keep = _as_input(geom_xpos, "geom positions (world_count, geoms_per_world, 3)", 3 * n)
        r, mem_r, keep_r = _as_input(geom_xmat, "geom rotations (world_count, geoms_per_world, 3, 3)", 9 * n)
        if mem != mem_r:
            raise ValueError("geom positions and rotations must both be on the host or both on the GPU")
        _lib.check(self._lib.ngx_set_geom_frames(self._h, p, r, mem, self._stream_for(mem)))
        self._hold((keep, keep_r), mem)

    @_serialized
    def set_mesh_vertices(self, mesh, vertices):
        """Moves mesh `mesh`'s vertices ((vertex_count, 3) float32, NumPy or a GPU tensor) and refits its bottom level on
        the GPU, keeping its tree (no counterpart in the reference, whose meshes are fixed). Set the instance
        transforms or body poses again before the next render: the top level still holds the mesh's old box."""
        if not 0 <= int(mesh) < len(self._mesh_vertex_counts):
            raise ValueError(f"mesh {mesh} is not one of the scene's {len(self._mesh_vertex_counts)} meshes")
        n = 3 * self._mesh_vertex_counts[int(mesh)]
        p, mem, keep = _as_input(vertices, f"vertices of mesh {mesh} ({n // 3}, 3)", n)
        _lib.check(self._lib.ngx_set_mesh_vertices(self._h, int(mesh), p, mem, self._stream_for(mem)))
        self._hold(keep, mem)

    # ---- images and rays
    @_serialized
    def render(self, cameras, fov=1.0, samples_per_pixel=1, *, rgb=False, depth=False, normals=False, object_ids=False,
               frame=0, zero_copy=False):
        """The images asked for, as a dict: 'depth' (views, H, W) float32, 'normals' (views, H, W, 3) float32 and
        'object_ids' (views, H, W) int32. With zero_copy the values are torch tensors viewing the renderer's
        buffers (overwritten by the next render); otherwise NumPy copies."""
        outputs = (_lib.OUTPUT_RGB if rgb else 0) | (_lib.OUTPUT_DEPTH if depth else 0) | \
                  (_lib.OUTPUT_NORMAL if normals else 0) | (_lib.OUTPUT_OBJECT_ID if object_ids else 0)
        views = self.view_count
        p, mem, keep = _as_input(cameras, f"cameras ({views} cameras of 3 x 4 numbers, camera-to-world)", views * 12)
        if mem == _lib.NGX_MEM_DEVICE and not zero_copy:
            raise ValueError("cameras on the GPU need zero_copy=True: the results then stay on the GPU too")
        f = _lib.BatchFrame()
        _lib.check(self._lib.ngx_render_batch(self._h, p, mem, float(fov), int(samples_per_pixel), outputs, int(frame),
                                              self._stream_for(mem), ctypes.byref(f)))
        self._hold(keep, mem)
        v, h, w = views, self._height, self._width
        shapes = {"rgb": ((v, h, w, 3), np.uint8, f.rgb), "depth": ((v, h, w), np.float32, f.depth),
                  "normals": ((v, h, w, 3), np.float32, f.normal), "object_ids": ((v, h, w), np.int32, f.object_id)}
        return self._outputs(shapes, zero_copy)

    @_serialized
    def cast_rays(self, origins, directions, max_distance=float("inf"), zero_copy=None):
        """Rays in every world: origins and directions (world_count, rays_per_world, 3). Returns 'distance'
        (world_count, rays_per_world), 'normal' (..., 3) and 'object_id'. Device tensors in give device tensors
        out, with no copy."""
        w = self._worlds
        po, mo, ko = _as_input(origins, "origins")
        pd, md, kd = _as_input(directions, "directions")
        n_o = ko.numel() if mo == _lib.NGX_MEM_DEVICE else ko.size
        n_d = kd.numel() if md == _lib.NGX_MEM_DEVICE else kd.size
        if n_o % (3 * w) != 0:
            raise ValueError(f"origins must hold 3 floats for each ray of each of the {w} worlds: (world_count, rays_per_world, 3)")
        if n_d != n_o:
            raise ValueError("directions must have the same shape as origins")
        if mo != md:
            raise ValueError("origins and directions must be on the same device")
        r = n_o // (3 * w)
        hits = _lib.RayHits()
        _lib.check(self._lib.ngx_cast_rays_batch(self._h, po, pd, int(r), float(max_distance), mo, self._stream_for(mo),
                                                 ctypes.byref(hits)))
        self._hold((ko, kd), mo)
        if zero_copy is None:
            zero_copy = mo == _lib.NGX_MEM_DEVICE
        shapes = {"distance": ((w, r), np.float32, hits.distance), "normal": ((w, r, 3), np.float32, hits.normal),
                  "object_id": ((w, r), np.int32, hits.object_id)}
        return self._outputs(shapes, zero_copy)

    # ---- gradients (no counterpart in the reference)
    @_serialized
    def pose_backward_depth(self, body_poses, cameras, fov, depth, grad_depth):
        """dL/d(body_poses) given dL/d(depth) of a render from these poses and cameras. Every argument is a float32
        torch tensor on the renderer's GPU: body_poses (world_count, bodies_per_world, 7), cameras (views, 3, 4),
        depth and grad_depth (views, H, W); anything else raises ValueError."""
        t = _torch()
        if self._bodies is None:
            raise ValueError("the pose backward needs bind_bodies first")
        views, dev = self.view_count, self._device_id
        pixels = views * self._height * self._width
        poses = _device_operand(body_poses, "body_poses", self._worlds * self._bodies * 7, dev)
        cams = _device_operand(cameras, "cameras", views * 12, dev)
        d = _device_operand(depth, "depth", pixels, dev)
        gd = _device_operand(grad_depth, "grad_depth", pixels, dev)
        grad = t.empty((self._worlds, self._bodies, 7), dtype=t.float32, device=poses.device)
        _lib.check(self._lib.ngx_pose_backward_depth(self._h, poses.data_ptr(), cams.data_ptr(), float(fov), d.data_ptr(),
                                                     gd.data_ptr(), grad.data_ptr(), _stream()))
        return grad

    @_serialized
    def pose_backward_range(self, body_poses, origins, directions, max_distance, distance, grad_distance):
        """dL/d(body_poses) given dL/d(distance) of a ray query from these poses. Every argument is a float32 torch
        tensor on the renderer's GPU: body_poses (world_count, bodies_per_world, 7), origins and directions
        (world_count, rays_per_world, 3), distance and grad_distance (world_count, rays_per_world)."""
        t = _torch()
        if self._bodies is None:
            raise ValueError("the pose backward needs bind_bodies first")
        w, dev = self._worlds, self._device_id
        n = origins.numel() if isinstance(origins, t.Tensor) else -1
        if n < 0 or n % (3 * w) != 0:
            raise ValueError(f"origins must hold 3 floats for each ray of each of the {w} worlds: (world_count, rays_per_world, 3)")
        r = n // (3 * w)
        poses = _device_operand(body_poses, "body_poses", w * self._bodies * 7, dev)
        o = _device_operand(origins, "origins", 3 * w * r, dev)
        dr = _device_operand(directions, "directions", 3 * w * r, dev)
        dist = _device_operand(distance, "distance", w * r, dev)
        gd = _device_operand(grad_distance, "grad_distance", w * r, dev)
        grad = t.empty((w, self._bodies, 7), dtype=t.float32, device=poses.device)
        _lib.check(self._lib.ngx_pose_backward_range(self._h, poses.data_ptr(), o.data_ptr(), dr.data_ptr(), int(r),
                                                     float(max_distance), dist.data_ptr(), gd.data_ptr(), grad.data_ptr(),
                                                     _stream()))
        return grad

    # ---- build read-back, for checking the top-level build
    @_serialized
    def keep_build_intermediates(self, keep=True):
        _lib.check(self._lib.ngx_debug_keep_build(self._h, int(bool(keep))))

    @_serialized
    def read_build(self, intermediates=False):
        from .layouts import TLAS_NODE
        n = self.instance_count
        out = {"nodes": np.zeros(max(2 * n - 1, 0), TLAS_NODE), "roots": np.zeros(self._worlds, np.int32),
               "inv": np.zeros((n, 12), np.float32), "matrices": np.zeros((n, 12), np.float32)}
        if intermediates:
            out.update(keys=np.zeros(n, np.uint64), sorted_keys=np.zeros(n, np.uint64), sorted_ids=np.zeros(n, np.int32),
                       aabb=np.zeros((n, 6), np.float32), world_bounds=np.zeros((self._worlds, 6), np.float32))
        names = ["nodes", "roots", "inv", "matrices", "keys", "sorted_keys", "sorted_ids", "aabb", "world_bounds"]
        ptrs = [out[k].ctypes.data_as(ctypes.c_void_p) if k in out else None for k in names]
        _lib.check(self._lib.ngx_debug_read_build(self._h, *ptrs))
        return out

    # ---- helpers
    def _stream_for(self, mem):
        if mem == _lib.NGX_MEM_DEVICE:
            return _stream()
        try:
            return _stream()
        except Exception:
            return ctypes.c_void_p(None)

    def _hold(self, keep, mem):
        # host inputs are copied asynchronously from pageable memory: wait before the caller may reuse them
        if mem == _lib.NGX_MEM_HOST:
            self.synchronize()

    def _outputs(self, shapes, zero_copy):
        out = {}
        if zero_copy:
            for k, (shape, dt, ptr) in shapes.items():
                if ptr:
                    out[k] = to_torch(ptr, shape, dt, self._device_id, self)
            return out
        self.synchronize()
        for k, (shape, dt, ptr) in shapes.items():
            if ptr:
                a = np.empty(shape, dt)
                _lib.check(self._lib.ngx_copy_to_host(self._h, ctypes.c_void_p(ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
                out[k] = a
        return out
