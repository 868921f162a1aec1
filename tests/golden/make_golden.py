"""Writes tests/golden/reference_vectors.npz: input/output vectors of the reference's own code, compiled from
its sources into oracle/_ref (oracle/Makefile), so the oracle can be checked where the reference is absent.

  meshes      make_icosphere(1, 2) and make_box(0.5) (src/mesh.cpp)
  sah_*       bvh::build on seeded boxes, leaf sizes 1 and 4 (src/bvh.cpp:46-174)
  transform_* make_transform on seeded poses (src/mesh.cpp:20-34)
  rays_*      camera_from_matrix + camera_ray for two cameras (src/render_core.h:264-273, src/device_scene.h:114-119)
  hits_*      mesh_intersect (src/mesh_dev.h:177-205) for those rays over a two-world scene of the three benchmark
              meshes; the top level handed to it is the oracle's LBVH (the reference has no CPU LBVH), and the
              closest hit does not depend on the top level's shape

Run: python tests/golden/make_golden.py (needs oracle/_ref, i.e. the reference tree at build time)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from neraulgraphics_amd import workloads  # noqa: E402
from oracle.oracle import Oracle, Ref, forward  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")


def main():
    ref, orc = Ref(), Oracle()
    v = {}
    v["ico_vertices"], v["ico_indices"] = ref.make_icosphere(1.0, 2)
    v["box_vertices"], v["box_indices"] = ref.make_box(0.5, 0.5, 0.5)
    rng = np.random.default_rng(2024)
    for k, (n, leaf) in enumerate([(37, 1), (300, 4), (64, 4)]):
        lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
        if k == 2:
            lo[:32] = lo[0]  # coincident boxes: the builder must split them by index
        b = np.concatenate([lo, lo + rng.uniform(0, 1, (n, 3)).astype(np.float32)], axis=1)
        nodes, order = ref.sah_build(b, leaf)
        v[f"sah{k}_bounds"], v[f"sah{k}_leaf"], v[f"sah{k}_nodes"], v[f"sah{k}_order"] = b, np.int32(leaf), nodes, order
    pos = rng.uniform(-5, 5, (16, 3)).astype(np.float32)
    quat = rng.normal(size=(16, 4)).astype(np.float32)
    scale = rng.uniform(0.5, 1.5, 16).astype(np.float32)
    v["transform_pos"], v["transform_quat"], v["transform_scale"] = pos, quat, scale
    v["transform_out"] = np.stack([ref.make_transform(pos[i], quat[i], float(scale[i])) for i in range(16)])
    cfg = workloads.make("instanced_worlds", worlds=2)
    W, H = 32, 32
    cams = cfg["cameras"]
    o, d, f = ref.camera_rays(cams, cfg["fov"], W, H)
    v["rays_cams"], v["rays_fov"], v["rays_size"] = cams, np.float32(cfg["fov"]), np.array([W, H], np.int32)
    v["rays_origins"], v["rays_dirs"], v["rays_forward"] = o, d, f
    st = forward(cfg, orc, outputs=())
    world = np.repeat(np.arange(2), W * H).astype(np.int32)
    hits = ref.mesh_intersect(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                              world, o.reshape(-1, 3), d.reshape(-1, 3))
    v["hits_matrices"] = cfg["matrices"]
    for k in ("t", "inst", "tri", "u", "v"):
        v["hits_" + k] = hits[k]
    np.savez_compressed(OUT, **v)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
