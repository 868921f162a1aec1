"""neraulgraphics_amd: an MI355X-native backend for the batched sensor-rendering path of neraulGraphics.

Per step, for thousands of worlds: new instance or body poses rebuild every world's top-level BVH on the GPU
(one kernel), then depth / normal / object-ID images or LiDAR ray queries are traced by hand-written HIP
kernels for gfx950, and the classic RGB picture (materials, sphere lights with soft shadows, mirror reflections,
sky) when asked; a backward kernel gives depth and range gradients with respect to body poses. The native
library is include/ngx_backend.h's C ABI (libngx.so); this package is its Python edge."""
from .geometry import (OBJECT_FLOOR, OBJECT_INSTANCE, OBJECT_NONE, OBJECT_SPHERE, Environment, Instance, Light, Material, Mesh,
                       Plane, Sphere, Transform, look_at, make_box, make_grid, make_icosphere, make_transform, shading_arrays)
from .batch import BatchRenderer

__all__ = ["BatchRenderer", "Mesh", "Instance", "Transform", "Material", "Sphere", "Plane", "Light", "Environment",
           "shading_arrays", "make_box",
           "make_icosphere", "make_grid", "make_transform", "look_at", "OBJECT_NONE", "OBJECT_SPHERE", "OBJECT_FLOOR",
           "OBJECT_INSTANCE"]
