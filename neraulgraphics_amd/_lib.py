"""ctypes bindings of libngx.so, the C ABI declared in include/ngx_backend.h.

The library is built in place by neraulgraphics_amd/build.py. There is no fallback: if it is missing or
fails to load, every entry point of this package raises."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# NGX_LIBRARY may name an instrumented build of the same sources (see build.py's extra_flags).
LIB_PATH = os.environ.get("NGX_LIBRARY", os.path.join(_HERE, "libngx.so"))

NGX_OK, NGX_INVALID_ARGUMENT, NGX_DEVICE_ERROR, NGX_OUT_OF_MEMORY, NGX_BAD_STATE, NGX_UNSUPPORTED = range(6)
NGX_MEM_HOST, NGX_MEM_DEVICE = 0, 1
OUTPUT_RGB, OUTPUT_DEPTH, OUTPUT_NORMAL, OUTPUT_OBJECT_ID = 1, 2, 4, 8

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_uint32_p = ctypes.POINTER(ctypes.c_uint32)
c_uint64_p = ctypes.POINTER(ctypes.c_uint64)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("worlds", ctypes.c_int32),
        ("cameras_per_world", ctypes.c_int32),
        ("mesh_count", ctypes.c_int32), ("vertices", c_float_p), ("vertex_offsets", c_int32_p),
        ("indices", c_uint32_p), ("index_offsets", c_int32_p), ("normals", c_float_p), ("mesh_has_normals", c_uint8_p),
        ("instances_per_world", ctypes.c_int32), ("instance_mesh", c_int32_p), ("instance_transforms", c_float_p),
        ("instances_in_world", c_int32_p),
        ("sphere_count", ctypes.c_int32), ("spheres", c_float_p),
        ("floor_y", ctypes.c_float), ("floor_half_extent_x", ctypes.c_float), ("floor_z_min", ctypes.c_float),
        ("floor_z_max", ctypes.c_float),
        ("device_id", ctypes.c_int32),
    ]


class BatchFrame(ctypes.Structure):
    _fields_ = [("memory", ctypes.c_int32), ("device_id", ctypes.c_int32), ("views", ctypes.c_int32),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("rgb", ctypes.c_void_p),
                ("depth", ctypes.c_void_p), ("normal", ctypes.c_void_p), ("object_id", ctypes.c_void_p)]


class Material(ctypes.Structure):
    _fields_ = [("diffuse_color", ctypes.c_float * 3), ("diffuse_albedo", ctypes.c_float), ("specular_albedo", ctypes.c_float),
                ("reflectivity", ctypes.c_float), ("specular_exponent", ctypes.c_float)]


class Light(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("intensity", ctypes.c_float), ("radius", ctypes.c_float),
                ("falloff", ctypes.c_float), ("direction", ctypes.c_float * 3), ("spot_angle", ctypes.c_float),
                ("spot_softness", ctypes.c_float), ("color", ctypes.c_float * 3)]


class Shading(ctypes.Structure):
    _fields_ = [("instance_materials", ctypes.c_void_p), ("sphere_materials", ctypes.c_void_p), ("floor_material", Material),
                ("floor_checker_scale", ctypes.c_float), ("light_count", ctypes.c_int32), ("lights", ctypes.c_void_p),
                ("ambient", ctypes.c_float), ("sky_up", ctypes.c_float * 3), ("sky_color_up", ctypes.c_float * 3),
                ("sky_color_down", ctypes.c_float * 3), ("sky_intensity", ctypes.c_float)]


assert ctypes.sizeof(Material) == 28 and ctypes.sizeof(Light) == 56


class TraceConfig(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_int32), ("xcd_remap", ctypes.c_int32), ("packet", ctypes.c_int32),
                ("pair_limit", ctypes.c_int32), ("persistent", ctypes.c_int32), ("lane_min_triangles", ctypes.c_int32),
                ("sort_rays", ctypes.c_int32), ("render_blocks", ctypes.c_int32), ("rgb_blocks", ctypes.c_int32),
                ("ray_blocks", ctypes.c_int32), ("backward_blocks", ctypes.c_int32), ("rgb_two_pass", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class RayHits(ctypes.Structure):
    _fields_ = [("memory", ctypes.c_int32), ("device_id", ctypes.c_int32), ("count", ctypes.c_int32),
                ("distance", ctypes.c_void_p), ("normal", ctypes.c_void_p), ("object_id", ctypes.c_void_p)]


_VP = ctypes.c_void_p
_SIGNATURES = {
    "ngx_create": (ctypes.c_int32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(_VP)]),
    "ngx_destroy": (None, [_VP]),
    "ngx_set_instance_matrices": (ctypes.c_int32, [_VP, _VP, ctypes.c_int32, _VP]),
    "ngx_bind_bodies": (ctypes.c_int32, [_VP, ctypes.c_int32, _VP, _VP, _VP]),
    "ngx_set_body_poses": (ctypes.c_int32, [_VP, _VP, ctypes.c_int32, _VP]),
    "ngx_bind_geoms": (ctypes.c_int32, [_VP, ctypes.c_int32, _VP, _VP]),
    "ngx_set_geom_frames": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_int32, _VP]),
    "ngx_set_mesh_vertices": (ctypes.c_int32, [_VP, ctypes.c_int32, _VP, ctypes.c_int32, _VP]),
    "ngx_render_batch": (ctypes.c_int32, [_VP, _VP, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_uint32,
                                          ctypes.c_int32, _VP, ctypes.POINTER(BatchFrame)]),
    "ngx_cast_rays_batch": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _VP,
                                             ctypes.POINTER(RayHits)]),
    "ngx_pose_backward_depth": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_float, _VP, _VP, _VP, _VP]),
    "ngx_pose_backward_depth_noalloc": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_float, _VP, _VP, _VP, _VP]),
    "ngx_pose_backward_range": (ctypes.c_int32, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.c_float, _VP, _VP, _VP, _VP]),
    "ngx_copy_to_host": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_size_t]),
    "ngx_synchronize": (ctypes.c_int32, [_VP]),
    "ngx_last_kernel_ms": (ctypes.c_float, [_VP]),
    "ngx_world_count": (ctypes.c_int32, [_VP]),
    "ngx_view_count": (ctypes.c_int32, [_VP]),
    "ngx_instances_per_world": (ctypes.c_int32, [_VP]),
    "ngx_instance_count": (ctypes.c_int32, [_VP]),
    "ngx_device_id": (ctypes.c_int32, [_VP]),
    "ngx_last_error": (ctypes.c_char_p, []),
    "ngx_autotune_rays": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_int32, _VP,
                                           ctypes.POINTER(TraceConfig), ctypes.POINTER(ctypes.c_float)]),
    "ngx_autotune_backward": (ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_float, _VP, _VP, _VP, ctypes.c_int32, _VP,
                                               ctypes.POINTER(TraceConfig), ctypes.POINTER(ctypes.c_float)]),
    "ngx_set_trace_config": (ctypes.c_int32, [_VP, ctypes.POINTER(TraceConfig)]),
    "ngx_get_trace_config": (ctypes.c_int32, [_VP, ctypes.POINTER(TraceConfig)]),
    "ngx_set_shading": (ctypes.c_int32, [_VP, ctypes.POINTER(Shading)]),
    "ngx_autotune": (ctypes.c_int32, [_VP, _VP, ctypes.c_int32, ctypes.c_float, ctypes.c_uint32, ctypes.c_int32, _VP,
                                      ctypes.POINTER(TraceConfig), ctypes.POINTER(ctypes.c_float)]),
    "ngx_debug_keep_build": (ctypes.c_int32, [_VP, ctypes.c_int32]),
    "ngx_debug_read_build": (ctypes.c_int32, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "ngx_build_blas_host": (ctypes.c_int32, [_VP, ctypes.c_int32, _VP, ctypes.c_int32, _VP, c_int32_p, _VP, _VP]),
    "ngx_debug_check_reciprocal": (ctypes.c_int32, [ctypes.c_int32, _VP, _VP, _VP]),
    "ngx_debug_wave_prefix_sum": (ctypes.c_int32, [ctypes.c_int32, _VP, ctypes.c_int32, _VP]),
    "ngx_debug_ray_order": (ctypes.c_int32, [ctypes.c_int32, _VP, ctypes.c_int32, ctypes.c_int32, _VP]),
    "ngx_debug_rgb_passes": (ctypes.c_int32, [_VP, ctypes.POINTER(ctypes.c_int32)]),
    "ngx_debug_set_share_capacity": (ctypes.c_int32, [_VP, ctypes.c_int32]),
    "ngx_build_skip_host": (ctypes.c_int32, [_VP, ctypes.c_int32, _VP, ctypes.c_int32, _VP, c_int32_p]),
}

_lib = None


def load():
    """The loaded library; raises if it is not built or does not load."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run neraulgraphics_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NgxError(RuntimeError):
    pass


def check(status):
    """Raises the Python counterpart of a failed status: ValueError for bad input (the reference's
    std::invalid_argument), NotImplementedError for what this backend does not do, RuntimeError otherwise."""
    if status == NGX_OK:
        return
    msg = load().ngx_last_error().decode()
    if status == NGX_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == NGX_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == NGX_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise NgxError(f"ngx status {status}: {msg}")
