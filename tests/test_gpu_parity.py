"""GPU parity: the HIP path through the C ABI against the oracle on the same seeded inputs.

Top-level build: Morton keys, the stable sort, and every world's subtree (node indices, children, boxes,
roots) bit-exact against the restatement of the reference's GPU build; instance inverses and body-pose
matrices bit-exact. Images and rays: see helpers.compare_images."""
import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import OBJECT_INSTANCE, workloads
from oracle.oracle import forward

from .helpers import compare_images, compare_rays, world_subset, world_subtree

pytestmark = pytest.mark.gpu

# The three walks: both levels as a packet; the packet walk with every mesh's bottom level walked lane by lane;
# both levels lane by lane.
WALKS = {"packet": dict(packet=True, lane_min_triangles=1 << 30), "hybrid": dict(packet=True, lane_min_triangles=0),
         "lanes": dict(packet=False)}


def _say(*a):
    print(*a, flush=True)


def _check_build(r, ref, world_off):
    b = r.read_build(intermediates=True)
    st = ref["state"]
    N = int(world_off[-1])
    assert np.array_equal(b["matrices"].view(np.uint32), ref["matrices"].view(np.uint32)), "placed matrices differ"
    assert np.array_equal(b["inv"].view(np.uint32), st["inv"].view(np.uint32)), "inverses differ"
    assert np.array_equal(b["keys"], st["keys"]), "Morton keys differ"
    assert np.array_equal(b["sorted_keys"], st["sorted_keys"]), "sorted keys differ"
    assert np.array_equal(b["sorted_ids"], st["sorted_ids"]), "sort permutation differs"
    assert np.array_equal(b["aabb"], st["aabb"]), "instance boxes differ"
    nonempty = np.diff(world_off) > 0
    assert np.array_equal(b["world_bounds"][nonempty], st["world_bounds"][nonempty]), "world centre bounds differ"
    assert np.array_equal(b["roots"], st["roots"]), "world roots differ"
    nodes_g, nodes_r = b["nodes"], st["nodes"]
    for w in np.flatnonzero(nonempty):
        idx = world_subtree(nodes_r, int(st["roots"][w]), N)
        g, e = nodes_g[idx], nodes_r[idx]
        assert np.array_equal(g["left"], e["left"]) and np.array_equal(g["right"], e["right"]), f"world {w} topology differs"
        assert np.array_equal(g["lo"], e["lo"]) and np.array_equal(g["hi"], e["hi"]), f"world {w} boxes differ"


@pytest.mark.parametrize("name,worlds", [("instanced_worlds", 64), ("body_worlds", 64), ("lidar_sweeps", 16)])
def test_top_level_build_matches_the_reference_algorithm(gpu, oracle, name, worlds):
    _say(f"build parity: {name}, {worlds} worlds")
    cfg = workloads.make(name, worlds=worlds, rays_per_world=64)
    r = workloads.make_renderer(cfg)
    r.keep_build_intermediates(True)
    if cfg["mode"] == "bodies":
        r.set_body_poses(cfg["body_poses"])
    else:
        r.set_instance_transforms(cfg["matrices"])
    ref = forward(cfg, oracle, outputs=())
    _check_build(r, ref, ref["world_off"])


@pytest.mark.parametrize("per_world", [200, 3000])
def test_the_build_sort_is_a_library_stable_sort(gpu, per_world):
    """The top-level build's radix sort (per world in LDS, or device-wide above 1,024 instances) gives the
    permutation a library sort gives: torch.sort(stable=True) on the GPU (rocPRIM underneath) of the same keys."""
    import torch
    _say(f"sort cross-check: 6 worlds of up to {per_world} instances")
    cfg = workloads.scattered(6, per_world, counts=[per_world, per_world // 2, 1, 0, per_world - 1, 17], width=16, height=16)
    r = workloads.make_renderer(cfg)
    r.keep_build_intermediates(True)
    r.set_instance_transforms(cfg["matrices"])
    b = r.read_build(intermediates=True)
    keys = torch.from_numpy(b["keys"].view(np.int64)).cuda()
    _, order = torch.sort(keys, stable=True)
    assert np.array_equal(order.cpu().numpy().astype(np.int32), b["sorted_ids"])
    assert np.array_equal(np.sort(b["keys"]), b["sorted_keys"])


def test_large_worlds_build_and_render_as_the_reference_algorithm(gpu, oracle):
    """Worlds above 1024 instances take the device-wide build; 1024 and below the workgroup build. Both must give
    the reference algorithm's keys, permutation and trees, and the same images."""
    counts = [3000, 17, 1025, 1024, 0, 2]
    _say(f"large-world build parity: counts {counts}")
    cfg = workloads.scattered(len(counts), 3000, counts=counts, spread=30.0)
    r = workloads.make_renderer(cfg)
    r.keep_build_intermediates(True)
    r.set_instance_transforms(cfg["matrices"])
    ref = forward(cfg, oracle)
    _check_build(r, ref, ref["world_off"])
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "large worlds")
    _say(rep)
    assert rep["hits"] > 1000


@pytest.mark.parametrize("worlds,per", [(1, 20000), (300, 40)])
def test_morton_trees_see_what_the_sah_trees_see(gpu, oracle, worlds, per):
    """tests/test_batch.cpp:335-371: 20,000 instances in one world and 300 worlds of 40, rendered through this
    backend's Morton-code trees and through the reference CPU backend's SAH trees (the oracle), agree: at most
    1 pixel in 2000 differs in object ID, and where IDs agree depths are within 0.01."""
    _say(f"LBVH vs SAH: {worlds} x {per}")
    cfg = workloads.scattered(worlds, per)
    r = workloads.make_renderer(cfg)
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, object_ids=True)
    ref = forward(cfg, oracle, tlas="sah", outputs=("depth", "id"))
    a_id, b_id = np.asarray(out["object_ids"]), ref["object_id"]
    a_d, b_d = np.asarray(out["depth"]), ref["depth"]
    diff = int((a_id != b_id).sum())
    hits = int((a_id != 0).sum())
    same = (a_id == b_id) & np.isfinite(a_d) & np.isfinite(b_d)
    worst = float(np.abs(a_d[same] - b_d[same]).max()) if same.any() else 0.0
    _say(f"{hits} of {a_id.size} pixels hit, {diff} differ in ID, worst depth difference {worst:.6f}")
    assert hits > 1000 and diff <= a_id.size // 2000 and worst < 0.01


@pytest.mark.parametrize("walk", list(WALKS))
@pytest.mark.parametrize("name,worlds", [("one_world_icospheres", None), ("instanced_worlds", 32), ("body_worlds", 16),
                                         ("posed_bodies", 16)])
def test_depth_normal_id_match_the_oracle(gpu, oracle, name, worlds, walk):
    _say(f"render parity: {name}, worlds {worlds}, {walk} walk")
    cfg = workloads.make(name, worlds=worlds)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS[walk])
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, name)
    _say(rep)
    assert rep["hits"] > 0.05 * rep["pixels"]
    # without the normal plane the kernel takes the object ID alone: the same depth and IDs, bit for bit
    lean = r.render(cfg["cameras"], cfg["fov"], depth=True, object_ids=True)
    assert np.array_equal(np.asarray(lean["depth"]).view(np.uint32), np.asarray(out["depth"]).view(np.uint32))
    assert np.array_equal(np.asarray(lean["object_ids"]), np.asarray(out["object_ids"]))


@pytest.mark.parametrize("name,worlds", [("body_worlds", 16), ("lidar_sweeps", 4)])
def test_lanes_own_stacks_give_the_oracle_results(gpu, oracle, name, worlds):
    """With the shared walk's LDS stack cut to one entry, nearly every far subtree a lane meets goes to the lane's own
    stack, the path a full LDS stack takes: depth, normal and ID images, and LiDAR ray results, still match the oracle, and
    no walk reports an overflow (a lane's own entries lie on one path down a tree). Capacities outside 1..128 are
    refused."""
    from neraulgraphics_amd import _lib
    _say(f"own-stack parity: {name}, {worlds} worlds, an LDS share stack of 1 entry")
    cfg = workloads.make(name, worlds=worlds, rays_per_world=8192)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS["hybrid"])
    for bad in (0, 129):
        with pytest.raises(ValueError, match="entries must be 1 to 128"):
            _lib.check(r._lib.ngx_debug_set_share_capacity(r._h, bad))
    _lib.check(r._lib.ngx_debug_set_share_capacity(r._h, 1))
    if name == "lidar_sweeps":
        out = r.cast_rays(cfg["origins"], cfg["directions"], cfg["max_distance"])
        st = forward(cfg, oracle, outputs=())
        ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                               st["meshes"]["attrs"], st["world_off"], cfg["origins"], cfg["directions"], cfg["max_distance"])
        rep = compare_rays(out, ref, name)
    else:
        out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
        ref = forward(cfg, oracle)
        rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, name)
    r.synchronize()
    _say(rep)
    assert rep["hits"] > 0


@pytest.mark.parametrize("tile_w,pair_limit,persistent", [(16, 16, True), (32, 1, False), (8, 64, True)])
def test_trace_configurations_render_the_same_images(gpu, oracle, tile_w, pair_limit, persistent):
    _say(f"render parity: body_worlds, 16 worlds, tile_w {tile_w}, pair_limit {pair_limit}, persistent {persistent}")
    cfg = workloads.make("body_worlds", worlds=16)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(tile_w=tile_w, pair_limit=pair_limit, persistent=persistent)
    assert r.trace_config == dict(tile_w=tile_w, xcd_remap=1, packet=1, pair_limit=pair_limit, persistent=int(persistent),
                                  lane_min_triangles=300, sort_rays=1, render_blocks=0, rgb_blocks=0, ray_blocks=0,
                                  backward_blocks=0, rgb_two_pass=1)
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    ref = forward(cfg, oracle)
    _say(compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "body_worlds"))


def test_every_kernel_variant_gives_the_same_results(gpu):
    """The kernels built for 4 to 7 workgroups per compute unit differ only in how many registers they may hold (and so
    in what they spill): depth, normal, ID and RGB images, ray-query results and pose gradients are bit-identical
    across every variant of both packet walks, and a variant that is not built is refused."""
    import torch
    import neraulgraphics_amd as ng
    _say("kernel variants: body_worlds 12 worlds, rays, backward")
    cfg = workloads.make("body_worlds", worlds=12)
    r = workloads.make_renderer(cfg, floor=ng.Plane(y=-6.0, half_extent_x=40.0, z_min=-60.0, z_max=20.0),
                                lights=[ng.Light((0.0, 12.0, -6.0), intensity=1.2, radius=0.5)])
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    rng = np.random.default_rng(17)
    o = torch.from_numpy(np.repeat(cfg["cameras"].reshape(12, 3, 4)[:, None, :, 3], 4096, axis=1).astype(np.float32)).cuda()
    d = torch.from_numpy(rng.normal(size=(12, 4096, 3)).astype(np.float32)).cuda()
    poses = torch.from_numpy(cfg["body_poses"]).cuda()

    def results(**variant):
        out = {}
        for walk in ("packet", "hybrid"):
            r.set_trace_config(**WALKS[walk], **variant)
            img = r.render(cfg["cameras"], cfg["fov"], 1, rgb=True, depth=True, normals=True, object_ids=True, frame=2)
            rays = r.cast_rays(o, d, 50.0)
            depth = torch.from_numpy(img["depth"]).cuda()
            g = torch.where(torch.isfinite(depth), torch.ones_like(depth), torch.zeros_like(depth))
            grad = r.pose_backward_depth(poses, cams, cfg["fov"], depth, g)
            out[walk] = [img["rgb"], img["depth"], img["normals"], img["object_ids"], rays["distance"].cpu().numpy(),
                         rays["object_id"].cpu().numpy(), grad.cpu().numpy()]
        return out

    base = results()
    for key, values in (("render_blocks", (5, 6, 7)), ("rgb_blocks", (4, 5, 6, 7)), ("ray_blocks", (5, 6, 7)),
                        ("backward_blocks", (4, 5, 6))):
        for v in values:
            got = results(**{key: v})
            for walk in base:
                for a, b in zip(base[walk], got[walk]):
                    assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), (key, v, walk)
    r.set_trace_config(render_blocks=6, ray_blocks=5)
    before = r.trace_config
    with pytest.raises(ValueError, match="render_blocks"):
        r.set_trace_config(render_blocks=8)
    with pytest.raises(ValueError, match="backward_blocks"):
        r.set_trace_config(render_blocks=5, backward_blocks=7)
    assert r.trace_config == before  # a refused configuration changes nothing


def test_ray_and_backward_autotuners_pick_built_variants_and_keep_results(gpu, oracle):
    """ngx_autotune_rays and ngx_autotune_backward: each returns a configuration of built variants with a time, leaves
    it set, and the results under it are the oracle's rays and the untuned gradient, bit for bit."""
    import torch
    _say("autotune: rays on lidar_sweeps 8 worlds x 4096 rays; backward on posed_bodies 16 worlds")
    cfg = workloads.make("lidar_sweeps", worlds=8, rays_per_world=4096)
    r = workloads.make_renderer(cfg)
    o, d = torch.from_numpy(cfg["origins"]).cuda(), torch.from_numpy(cfg["directions"]).cuda()
    got = r.autotune_rays(o, d, cfg["max_distance"], reps=1)
    assert got["ms"] > 0 and got["ray_blocks"] in (5, 6, 7) and got["sort_rays"] in (0, 1)
    assert {k: got[k] for k in r.trace_config} == r.trace_config
    out = r.cast_rays(cfg["origins"], cfg["directions"], cfg["max_distance"])
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], cfg["origins"], cfg["directions"], cfg["max_distance"])
    compare_rays(out, ref, "rays after autotune_rays")
    pc = workloads.make("posed_bodies", worlds=16)
    rb = workloads.make_renderer(pc)
    poses = torch.from_numpy(pc["body_poses"]).cuda()
    cams = torch.from_numpy(pc["cameras"]).cuda()
    depth = rb.render(cams, pc["fov"], depth=True, zero_copy=True)["depth"].clone()
    g = torch.where(torch.isfinite(depth), torch.full_like(depth, 0.5), torch.zeros_like(depth))
    want = rb.pose_backward_depth(poses, cams, pc["fov"], depth, g).cpu().numpy()
    tuned = rb.autotune_backward(poses, cams, pc["fov"], depth, g, reps=1)
    assert tuned["ms"] > 0 and tuned["backward_blocks"] in (4, 5, 6)
    again = rb.pose_backward_depth(poses, cams, pc["fov"], depth, g).cpu().numpy()
    assert np.array_equal(want.view(np.uint32), again.view(np.uint32))


def test_autotune_picks_a_configuration_and_keeps_parity(gpu, oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("NGX_TUNE_CACHE", str(tmp_path / "tune.json"))
    cfg = workloads.make("instanced_worlds", worlds=64)
    r = workloads.make_renderer(cfg)
    _say("autotune: instanced_worlds, 64 worlds")
    got = r.autotune(cfg["cameras"], cfg["fov"], reps=2)
    _say(got)
    assert got["ms"] > 0 and got["tile_w"] in (8, 16, 32, 64)
    assert {k: got[k] for k in r.trace_config} == r.trace_config
    again = r.autotune(cfg["cameras"], cfg["fov"])
    assert again["ms"] is None and {k: again[k] for k in r.trace_config} == r.trace_config
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    ref = forward(cfg, oracle)
    compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "instanced_worlds tuned")


@pytest.mark.parametrize("walk,sort_rays", [("packet", True), ("packet", False), ("hybrid", True), ("lanes", False)])
def test_ray_queries_match_the_oracle(gpu, oracle, walk, sort_rays):
    _say(f"ray parity: lidar_sweeps, 8 worlds x 4096 rays, {walk} walk, direction order {sort_rays}")
    cfg = workloads.make("lidar_sweeps", worlds=8, rays_per_world=4096)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS[walk], sort_rays=sort_rays)
    out = r.cast_rays(cfg["origins"], cfg["directions"], cfg["max_distance"])
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], cfg["origins"], cfg["directions"], cfg["max_distance"])
    rep = compare_rays(out, ref, "lidar_sweeps")
    _say(rep)
    assert rep["hits"] > 0


def test_ray_order_of_many_rays_a_world(gpu, oracle):
    """Above 65,536 rays a world the direction order sorts with 4-bit digits in three passes (below, 6-bit digits in
    two): 2 worlds of 70,000 rays in direction order against the oracle."""
    _say("ray parity: 2 worlds x 70,000 rays, hybrid walk, direction order")
    cfg = workloads.make("lidar_sweeps", worlds=2, rays_per_world=70016)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS["hybrid"], sort_rays=True)
    o, d = cfg["origins"][:, :70000], cfg["directions"][:, :70000]
    out = r.cast_rays(o, d, cfg["max_distance"])
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], o, d, cfg["max_distance"])
    rep = compare_rays(out, ref, "70,000 rays a world")
    assert rep["hits"] > 0


@pytest.mark.parametrize("walk", list(WALKS))
def test_rays_of_uneven_worlds_per_wave(gpu, oracle, walk):
    _say(f"ray parity: 37 rays per world, so waves straddle worlds, {walk} walk")
    cfg = workloads.make("lidar_sweeps", worlds=8, rays_per_world=37 * 64)
    o = cfg["origins"][:, :37]
    d = cfg["directions"][:, ::64][:, :37]
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS[walk])
    out = r.cast_rays(o, d, 100.0)
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], o, d, 100.0)
    compare_rays(out, ref, "37 rays per world")


@pytest.mark.parametrize("name", ["body_worlds", "instanced_worlds"])
def test_full_batch_on_a_sample_of_worlds(gpu, oracle, name):
    """A benchmark batch at full size (body_worlds: 4096 worlds at 128x128, 534,919 instances from body poses;
    instanced_worlds: 4096 worlds at 64x64, 131,072 instances from matrices) rendered in one launch with the default
    configuration; 16 worlds spread over the batch, including the first and last, are checked against the oracle
    run on those worlds alone. A world's tree and its pixels depend only on its own instances, so the sample is
    checked as the whole batch would be."""
    import torch
    _say(f"full-size parity: {name}, 4096 worlds, 16 sampled worlds")
    cfg = workloads.make(name)
    r = workloads.make_renderer(cfg)
    out = r.render(torch.from_numpy(cfg["cameras"]).cuda(), cfg["fov"], depth=True, normals=True, object_ids=True, zero_copy=True)
    torch.cuda.synchronize()
    sample = np.unique(np.concatenate([[0, cfg["worlds"] - 1], np.random.default_rng(5).choice(cfg["worlds"], 14, replace=False)]))
    got = {k: out[k][torch.from_numpy(sample).cuda()].cpu().numpy() for k in ("depth", "normals", "object_ids")}
    ref = forward(world_subset(cfg, sample), oracle)
    rep = compare_images(dict(depth=got["depth"], object_id=got["object_ids"], normal=got["normals"]), ref, f"{name} sample")
    _say(rep)
    assert rep["hits"] > 0.3 * rep["pixels"]


def test_full_lidar_batch_on_a_sample_of_worlds(gpu, oracle):
    """The LiDAR query at full size (1024 worlds x 65,536 rays over 100,000 instances of the 138,000-triangle
    mesh) in one call, rays in direction order; 3 worlds (first, middle, last) checked against the oracle."""
    import torch
    _say("full-size parity: lidar_sweeps, 1024 worlds x 65536 rays, 3 sampled worlds")
    cfg = workloads.make("lidar_sweeps")
    r = workloads.make_renderer(cfg)
    o = torch.from_numpy(cfg["origins"]).cuda()
    d = torch.from_numpy(cfg["directions"]).cuda()
    out = r.cast_rays(o, d, cfg["max_distance"])
    torch.cuda.synchronize()
    sample = np.array([0, cfg["worlds"] // 2, cfg["worlds"] - 1])
    idx = torch.from_numpy(sample).cuda()
    got = {k: out[k][idx].cpu().numpy() for k in ("distance", "normal", "object_id")}
    sub = world_subset(cfg, sample)
    st = forward(sub, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], sub["origins"], sub["directions"], cfg["max_distance"])
    rep = compare_rays(got, ref, "lidar_sweeps sample")
    _say(rep)
    assert rep["hits"] > 0.1 * rep["pixels"]


@pytest.mark.parametrize("walk", list(WALKS))
def test_a_mesh_of_coincident_triangles(gpu, oracle, walk):
    """A mesh of 150 copies of one triangle: every ray that meets it meets 150 surfaces at one distance, spread
    over many leaves that all overlap; the depth is exact and the ID the instance's whichever copy is kept."""
    import neraulgraphics_amd as ng
    _say(f"coincident triangles, {walk} walk")
    v = np.array([[-3.0, -3.0, 0.0], [3.0, -3.0, 0.0], [0.0, 3.0, 0.0]], np.float32)
    mesh = ng.Mesh(np.tile(v, (150, 1)), np.arange(450, dtype=np.int32))
    cfg = workloads.scattered(4, 6, width=32, height=32, spread=0.5)
    cfg["meshes"] = [mesh, cfg["meshes"][1]]
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS[walk])
    if walk == "hybrid":
        r.set_trace_config(lane_min_triangles=0)
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "coincident")
    _say(rep)
    assert rep["hits"] > 100


@pytest.mark.parametrize("walk", list(WALKS))
def test_several_cameras_per_world_and_a_ragged_frame(gpu, oracle, walk):
    """Three cameras per world (view v sees world v // 3), on a 45 x 29 frame whose last tiles are partly outside
    it; depth, normal and ID, and the RGB picture."""
    _say(f"3 cameras per world, 45x29, {walk} walk")
    cfg = workloads.make("instanced_worlds", worlds=6)
    rng = np.random.default_rng(21)
    cams = []
    for w in range(cfg["worlds"]):
        target = cfg["positions"][w].mean(axis=0)
        for k in range(3):
            cams.append(ng.look_at(target + rng.uniform(-6, 6, size=3) + np.array([0.0, 0.0, 8.0]), target).reshape(12))
    cfg.update(cameras_per_world=3, cameras=np.array(cams, np.float32), width=45, height=29)
    r = workloads.make_renderer(cfg)
    r.set_trace_config(**WALKS[walk])
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    assert out["depth"].shape == (18, 29, 45)
    ref = forward(cfg, oracle)
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "3 cameras")
    _say(rep)
    assert rep["hits"] > 0.1 * rep["pixels"]
    c = dict(cfg)
    c["shading"] = ng.shading_arrays(cfg["instances"], [], ng.Plane(), [ng.Light((0.0, 10.0, -5.0), radius=0.3)], 0.1, None)
    r2 = workloads.make_renderer(cfg, lights=[ng.Light((0.0, 10.0, -5.0), radius=0.3)])
    r2.set_trace_config(**WALKS[walk])
    r2.set_ambient(0.1)
    got = r2.render(cfg["cameras"], cfg["fov"], 1, rgb=True, frame=2)["rgb"].astype(np.int32)
    c["spp"], c["frame"] = 1, 2
    want = forward(c, oracle, outputs=("rgb",))["rgb"].astype(np.int32)
    assert np.abs(got - want).mean() < 0.05


def test_worlds_do_not_see_one_another(gpu):
    """tests/test_batch.cpp:185-214: boxes at z=-8 and z=-3 in two worlds, seen from the origin."""
    import neraulgraphics_amd as ng
    W, H = 96, 54
    b = ng.BatchRenderer(W, H, 2, meshes=[ng.make_box((1, 1, 1))], instances=[ng.Instance(0)])
    poses = np.stack([ng.make_transform((0, 0, z)).matrix for z in (-8.0, -3.0)]).reshape(2, 1, 3, 4)
    b.set_instance_transforms(poses)
    cam = ng.look_at((0, 0, 0), (0, 0, -1)).reshape(12)
    out = b.render(np.stack([cam, cam]), 1.5708, depth=True, object_ids=True)
    centre = out["depth"][:, H // 2, W // 2]
    assert np.allclose(centre, [7.0, 2.0], atol=0.02), centre
    hits = b.cast_rays(np.zeros((2, 1, 3), np.float32), np.tile(np.array([0, 0, -1], np.float32), (2, 1, 1)), 100.0)
    assert np.allclose(hits["distance"][:, 0], [7.0, 2.0], atol=1e-3) and (hits["object_id"] == OBJECT_INSTANCE).all()


def test_single_instance_worlds_and_a_pile_of_coincident_boxes(gpu):
    """tests/test_batch.cpp:311-331: one-leaf trees at 3 + w; 1,000 identical boxes whose front face is 4.8 away."""
    import neraulgraphics_amd as ng
    W, H = 96, 54
    cam = ng.look_at((0, 0, 0), (0, 0, -8)).reshape(12)
    single = ng.BatchRenderer(W, H, 4, meshes=[ng.make_box((1, 1, 1))], instances=[ng.Instance(0)])
    single.set_instance_transforms(np.stack([ng.make_transform((0, 0, -4.0 - w)).matrix for w in range(4)]).reshape(4, 1, 3, 4))
    d = single.render(np.tile(cam, (4, 1)), 1.5708, depth=True)["depth"][:, H // 2, W // 2]
    assert np.allclose(d, 3.0 + np.arange(4), atol=0.02), d
    pile = ng.BatchRenderer(W, H, 1, meshes=[ng.make_box((0.2, 0.2, 0.2))], instances=[ng.Instance(0)] * 1000)
    pile.set_instance_transforms(np.tile(ng.make_transform((0, 0, -5)).matrix, (1, 1000, 1, 1)))
    d = pile.render(cam[None], 1.5708, depth=True)["depth"][0, H // 2, W // 2]
    assert abs(d - 4.8) < 0.02, d


def test_spheres_and_floor(gpu, oracle):
    """Analytic spheres and the floor before the meshes (render_core.h:78-112); a mesh must be strictly closer."""
    import neraulgraphics_amd as ng
    cfg = workloads.make("instanced_worlds", worlds=4)
    spheres = [ng.Sphere((0.5, -1.0, -9.0), 1.0), ng.Sphere((-2.0, 1.0, -7.0), 0.7)]
    floor = ng.Plane(y=-2.0, half_extent_x=30.0, z_min=-60.0, z_max=5.0)
    r = workloads.make_renderer(cfg, spheres=spheres, floor=floor)
    out = r.render(cfg["cameras"], cfg["fov"], depth=True, normals=True, object_ids=True)
    st = forward(cfg, oracle, outputs=())
    ref = oracle.render(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                        st["meshes"]["attrs"], st["world_off"], cfg["cameras"], cfg["fov"], cfg["width"], cfg["height"],
                        spheres=np.array([[0.5, -1.0, -9.0, 1.0], [-2.0, 1.0, -7.0, 0.7]]),
                        floor=np.array([-2.0, 30.0, -60.0, 5.0]))
    rep = compare_images(dict(depth=out["depth"], object_id=out["object_ids"], normal=out["normals"]), ref, "spheres+floor")
    _say(rep)
    ids = np.unique(out["object_ids"])
    assert 33 in ids or 1 in ids or 2 in ids


def test_zero_copy_views_are_overwritten_by_the_next_render(gpu):
    import torch
    cfg = workloads.make("instanced_worlds", worlds=4)
    r = workloads.make_renderer(cfg)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    first = r.render(cams, cfg["fov"], depth=True, zero_copy=True)["depth"]
    kept = first.clone()
    moved = cams.clone()
    moved[:, 3] += 0.5
    second = r.render(moved, cfg["fov"], depth=True, zero_copy=True)["depth"]
    torch.cuda.synchronize()
    assert first.data_ptr() == second.data_ptr() and not torch.equal(first, kept)


def test_threads_sharing_a_renderer_get_their_own_images(gpu):
    """python/bindings.cpp:242-251: one renderer used from several threads, each call serialized; every thread's
    host images are those of its own cameras."""
    import threading
    cfg = workloads.make("instanced_worlds", worlds=16)
    r = workloads.make_renderer(cfg)
    cams = [cfg["cameras"].copy() for _ in range(4)]
    for k, c in enumerate(cams):
        c[:, 3] += 0.25 * k
    want = [r.render(c, cfg["fov"], depth=True, object_ids=True) for c in cams]
    got, errors = [[None] * 6 for _ in cams], []

    def work(k):
        try:
            for i in range(6):
                got[k][i] = r.render(cams[k], cfg["fov"], depth=True, object_ids=True)
        except Exception as e:  # reported below: a thread's exception does not fail the test by itself
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(cams))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not errors and not any(t.is_alive() for t in threads)
    for k in range(len(cams)):
        for g in got[k]:
            assert np.array_equal(g["depth"].view(np.uint32), want[k]["depth"].view(np.uint32))
            assert np.array_equal(g["object_ids"], want[k]["object_ids"])


def test_graph_replays_equal_direct_launches(gpu):
    """A step (new body poses, then depth + ID, persistent launch with its counter reset) captured in HIP graphs,
    one graph per pose set, replayed in turn with direct renders and host copies between the replays: every
    replay gives the images the same step gives launched directly."""
    import torch
    cfg = workloads.make("body_worlds", worlds=32)
    r = workloads.make_renderer(cfg)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    rng = np.random.default_rng(11)
    poses = [torch.from_numpy(cfg["body_poses"]).cuda()]
    for _ in range(2):
        p = cfg["body_poses"].copy()
        p[..., :3] += rng.normal(0, 0.3, p[..., :3].shape).astype(np.float32)
        poses.append(torch.from_numpy(p).cuda())

    def step(k):
        r.set_body_poses(poses[k])
        return r.render(cams, cfg["fov"], depth=True, object_ids=True, zero_copy=True)

    want = []
    for k in range(3):
        out = step(k)
        torch.cuda.synchronize()
        want.append({n: out[n].clone() for n in ("depth", "object_ids")})
    assert not torch.equal(want[0]["depth"], want[1]["depth"])
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(0)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graphs = []
    for k in range(3):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = step(k)
        graphs.append((g, out))
    torch.cuda.synchronize()
    for rep in range(3):
        for k in (2, 0, 1):
            g, out = graphs[k]
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out["depth"], want[k]["depth"]), (rep, k)
            assert torch.equal(out["object_ids"], want[k]["object_ids"]), (rep, k)
            step((k + 1) % 3)  # direct work between replays
            torch.cuda.synchronize()


def test_device_poses_equal_host_poses(gpu):
    import torch
    cfg = workloads.make("body_worlds", worlds=8)
    r = workloads.make_renderer(cfg)
    host = r.render(cfg["cameras"], cfg["fov"], depth=True, object_ids=True)
    r.set_body_poses(torch.from_numpy(cfg["body_poses"]).cuda())
    dev = r.render(torch.from_numpy(cfg["cameras"]).cuda(), cfg["fov"], depth=True, object_ids=True, zero_copy=True)
    torch.cuda.synchronize()
    assert np.array_equal(dev["depth"].cpu().numpy().view(np.uint32), host["depth"].view(np.uint32))
    assert np.array_equal(dev["object_ids"].cpu().numpy(), host["object_ids"])


def test_bad_input_is_rejected_and_changes_nothing(gpu):
    cfg = workloads.make("instanced_worlds", worlds=2)
    r = workloads.make_renderer(cfg)
    before = r.render(cfg["cameras"], cfg["fov"], depth=True)["depth"]
    bad = cfg["matrices"].copy()
    bad[1, 4] = 0.0
    with pytest.raises(ValueError, match="singular"):
        r.set_instance_transforms(bad)
    bad = cfg["matrices"].copy()
    bad[0, 0, 7] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        r.set_instance_transforms(bad)
    with pytest.raises(ValueError, match="at least one output"):
        r.render(cfg["cameras"], cfg["fov"])
    with pytest.raises(ValueError, match="fov"):
        r.render(cfg["cameras"], 4.0, depth=True)
    with pytest.raises(ValueError, match="samples_per_pixel"):
        r.render(cfg["cameras"], cfg["fov"], 0, depth=True)
    cams = cfg["cameras"].copy()
    cams[0, 3] = np.nan
    with pytest.raises(ValueError, match="camera"):
        r.render(cams, cfg["fov"], depth=True)
    with pytest.raises(ValueError, match="bind_bodies"):
        r.set_body_poses(np.zeros((2, 1, 7), np.float32))
    with pytest.raises(ValueError, match="body"):
        r.bind_bodies(np.array([0] * 31 + [5], np.int32), np.tile([0, 0, 0, 0, 0, 0, 1], (32, 1)), np.ones((32, 3)), 1)
    after = r.render(cfg["cameras"], cfg["fov"], depth=True)["depth"]
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))


def test_pose_gradient_matches_fp64_closed_form_and_differences(gpu, oracle):
    import torch
    from neraulgraphics_amd.autograd import render_depth
    _say("backward: posed_bodies, 8 worlds")
    cfg = workloads.make("posed_bodies", worlds=8)
    r = workloads.make_renderer(cfg)
    target = torch.from_numpy(render_target(r, cfg)).cuda()
    poses = torch.from_numpy(cfg["body_poses"]).cuda().requires_grad_(True)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    depth = render_depth(r, poses, cams, cfg["fov"])
    both = torch.isfinite(depth) & torch.isfinite(target)
    diff = torch.where(both, depth - target, torch.zeros_like(depth))
    loss = 0.5 * (diff * diff).sum()
    loss.backward()
    g_gpu = poses.grad.cpu().numpy().astype(np.float64)
    # FP64 references on the oracle: the closed form and central differences, step 1e-4, per hit pixel
    weight = diff.detach().cpu().numpy().reshape(-1).astype(np.float64)
    ref = forward(cfg, oracle)
    hit_inst = ref["hit_inst"].reshape(-1)
    hit_tri = ref["hit_tri"].reshape(-1)
    V, H, W = ref["depth"].shape
    origins = np.zeros((V * H * W, 3))
    dirs = np.zeros((V * H * W, 3))
    cosv = np.zeros(V * H * W)
    for v in range(V):
        for y in range(H):
            for x in range(W):
                o, d, f = oracle.camera_ray(cfg["cameras"][v], cfg["fov"], W, H, x, y)
                k = (v * H + y) * W + x
                origins[k], dirs[k], cosv[k] = o, d, float(np.dot(d.astype(np.float64), f.astype(np.float64)))
    ray_world = np.repeat(np.arange(V), H * W).astype(np.int32)
    args = (origins, dirs, weight, cosv, hit_inst, hit_tri, ray_world, ref["world_off"], cfg["bodies"], cfg["body_of_instance"],
            cfg["local_pose"], cfg["scale"], cfg["body_poses"], ref["meshes"]["tris"])
    g_cf = oracle.pose_grad(0, 0.0, *args)
    g_fd = oracle.pose_grad(1, 1e-4, *args)
    g_fine = oracle.pose_grad(1, 1e-6, *args)
    pin = np.linalg.norm(g_cf - g_fine) / np.linalg.norm(g_fine)
    rel = np.linalg.norm(g_gpu - g_fd) / np.linalg.norm(g_fd)
    rel_cf = np.linalg.norm(g_gpu - g_cf) / np.linalg.norm(g_cf)
    _say(f"|closed form - FD(1e-6)| / |FD| = {pin:.2e}; |GPU - FD(1e-4)| / |FD| = {rel:.2e}; |GPU - closed form| = {rel_cf:.2e}")
    assert pin < 1e-7  # the FP64 closed form is the derivative (a wrong sign or term would be O(1))
    assert rel < 1e-3  # the bar: FP32 GPU gradient within 1e-3 of FP64 central differences, step 1e-4
    assert rel_cf < 1e-4


def test_range_gradient_matches_fp64_closed_form_and_differences(gpu, oracle):
    """cast_range's backward (range_backward_kernel): dL/d(body poses) of L = 0.5 sum (range - target)^2 over rays
    that hit in both, against the oracle's FP64 closed form and central differences through the same hit triangles."""
    import torch
    from neraulgraphics_amd.autograd import cast_range
    _say("range backward: posed_bodies, 8 worlds x 2048 rays")
    cfg = workloads.make("posed_bodies", worlds=8)
    r = workloads.make_renderer(cfg)
    W, R = cfg["worlds"], 2048
    rng = np.random.default_rng(21)
    eye = cfg["cameras"].reshape(W, 3, 4)[:, :, 3]
    aim = cfg["body_poses"][np.arange(W)[:, None], rng.integers(0, cfg["bodies"], (W, R)), :3]
    origins = np.repeat(eye[:, None, :], R, axis=1).astype(np.float32)
    dirs = (aim + rng.normal(0, 0.3, aim.shape) - origins).astype(np.float32)
    r.set_body_poses(cfg["target_poses"])
    target = r.cast_rays(origins, dirs, 100.0)["distance"].copy()
    poses = torch.from_numpy(cfg["body_poses"]).cuda().requires_grad_(True)
    dist = cast_range(r, poses, torch.from_numpy(origins).cuda(), torch.from_numpy(dirs).cuda(), 100.0)
    tgt = torch.from_numpy(target).cuda()
    both = torch.isfinite(dist) & torch.isfinite(tgt)
    diff = torch.where(both, dist - tgt, torch.zeros_like(dist))
    (0.5 * (diff * diff).sum()).backward()
    g_gpu = poses.grad.cpu().numpy().astype(np.float64)
    assert int(both.sum()) > W * R // 4
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], origins, dirs, 100.0)
    d32 = dirs.reshape(-1, 3)
    unit = (d32 * (np.float32(1.0) / np.sqrt((d32 * d32).sum(axis=1, dtype=np.float32)))[:, None]).astype(np.float64)
    args = (origins.reshape(-1, 3).astype(np.float64), unit, diff.detach().cpu().numpy().reshape(-1).astype(np.float64),
            np.ones(W * R), ref["hit_inst"].reshape(-1), ref["hit_tri"].reshape(-1), np.repeat(np.arange(W), R).astype(np.int32),
            st["world_off"], cfg["bodies"], cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["body_poses"],
            st["meshes"]["tris"])
    g_cf = oracle.pose_grad(0, 0.0, *args)
    g_fd = oracle.pose_grad(1, 1e-4, *args)
    rel = np.linalg.norm(g_gpu - g_fd) / np.linalg.norm(g_fd)
    rel_cf = np.linalg.norm(g_gpu - g_cf) / np.linalg.norm(g_cf)
    _say(f"|GPU - FD(1e-4)| / |FD| = {rel:.2e}; |GPU - closed form| / |closed form| = {rel_cf:.2e}")
    assert rel < 1e-3
    assert rel_cf < 1e-4


def _depth_gradient_refs(oracle, cfg, weight, sample):
    """The oracle's FP64 closed-form gradient and central differences (step 1e-4) of sum(weight x depth) for the
    worlds in `sample` (weight: (len(sample) x cameras_per_world, H, W) over those worlds' views), through the hit
    triangles of the oracle's own render of those worlds."""
    sub = world_subset(cfg, sample)
    ref = forward(sub, oracle)
    V, H, W = ref["depth"].shape
    origins, dirs, cosv = np.zeros((V * H * W, 3)), np.zeros((V * H * W, 3)), np.zeros(V * H * W)
    for v in range(V):
        for y in range(H):
            for x in range(W):
                o, d, f = oracle.camera_ray(sub["cameras"][v], sub["fov"], W, H, x, y)
                k = (v * H + y) * W + x
                origins[k], dirs[k], cosv[k] = o, d, float(np.dot(d.astype(np.float64), f.astype(np.float64)))
    ray_world = np.repeat(np.arange(V) // sub["cameras_per_world"], H * W).astype(np.int32)
    args = (origins, dirs, np.asarray(weight, np.float64).reshape(-1), cosv, ref["hit_inst"].reshape(-1), ref["hit_tri"].reshape(-1),
            ray_world, ref["world_off"], sub["bodies"], sub["body_of_instance"], sub["local_pose"], sub["scale"], sub["body_poses"],
            ref["meshes"]["tris"])
    return oracle.pose_grad(0, 0.0, *args), oracle.pose_grad(1, 1e-4, *args)


def _depth_loss_gradient(r, cfg, cams):
    """GPU gradient of 0.5 |depth - target|^2 (pixels that hit in both) and the per-pixel weight depth - target."""
    import torch
    from neraulgraphics_amd.autograd import render_depth
    c = torch.from_numpy(np.ascontiguousarray(cams, np.float32)).cuda()
    r.set_body_poses(cfg["target_poses"])
    target = r.render(c, cfg["fov"], depth=True, zero_copy=True)["depth"].clone()
    poses = torch.from_numpy(cfg["body_poses"]).cuda().requires_grad_(True)
    depth = render_depth(r, poses, c, cfg["fov"])
    both = torch.isfinite(depth) & torch.isfinite(target)
    diff = torch.where(both, depth - target, torch.zeros_like(depth))
    (0.5 * (diff * diff).sum()).backward()
    return poses.grad.cpu().numpy().astype(np.float64), diff.detach().cpu().numpy()


def test_pose_gradient_with_three_cameras_a_world(gpu, oracle):
    """The depth backward with 3 cameras a world: each world's gradient sums its three views' partial sums (the
    cameras_per_world x parts reduction of the finish kernel), against the oracle's FP64 closed form and central
    differences over all 3 x 8 views."""
    from neraulgraphics_amd.batch import BatchRenderer
    _say("backward: posed_bodies, 8 worlds x 3 cameras")
    cfg = workloads.make("posed_bodies", worlds=8)
    cpw = 3
    rng = np.random.default_rng(41)
    cams = np.repeat(cfg["cameras"].reshape(8, 1, 12), cpw, axis=1).copy()
    cams[:, 1:, [3, 7, 11]] += rng.normal(0.0, 0.4, (8, cpw - 1, 3)).astype(np.float32)
    cfg = dict(cfg, cameras_per_world=cpw, cameras=cams.reshape(-1, 12))
    r = BatchRenderer(64, 64, 8, cpw, meshes=cfg["meshes"], instances=cfg["instances"])
    r.bind_bodies(cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], cfg["bodies"])
    g_gpu, weight = _depth_loss_gradient(r, cfg, cfg["cameras"])
    g_cf, g_fd = _depth_gradient_refs(oracle, cfg, weight, np.arange(8))
    rel = np.linalg.norm(g_gpu - g_fd) / np.linalg.norm(g_fd)
    rel_cf = np.linalg.norm(g_gpu - g_cf) / np.linalg.norm(g_cf)
    _say(f"|GPU - FD(1e-4)| / |FD| = {rel:.2e}; |GPU - closed form| / |closed form| = {rel_cf:.2e}")
    assert rel < 1e-3 and rel_cf < 1e-4


def test_pose_gradient_of_the_full_posed_bodies_batch_on_a_sample(gpu, oracle):
    """The fwd+bwd configuration at full size (1024 worlds at 64x64, 32 bodies each): one forward and one backward
    over the whole batch; 12 worlds spread over it (first and last included) against the oracle's FP64 closed form
    and central differences on those worlds alone. A world's gradient depends on its own views only."""
    _say("backward: posed_bodies, 1024 worlds, 12 sampled worlds")
    cfg = workloads.make("posed_bodies")
    r = workloads.make_renderer(cfg)
    g_gpu, weight = _depth_loss_gradient(r, cfg, cfg["cameras"])
    sample = np.unique(np.concatenate([[0, cfg["worlds"] - 1], np.random.default_rng(9).choice(cfg["worlds"], 10, replace=False)]))
    g_cf, g_fd = _depth_gradient_refs(oracle, cfg, weight[sample], sample)
    g = g_gpu[sample]
    rel = np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd)
    rel_cf = np.linalg.norm(g - g_cf) / np.linalg.norm(g_cf)
    _say(f"|GPU - FD(1e-4)| / |FD| = {rel:.2e}; |GPU - closed form| / |closed form| = {rel_cf:.2e}")
    assert rel < 1e-3 and rel_cf < 1e-4
    assert np.all(np.isfinite(g_gpu))


def test_range_gradient_on_a_lidar_sweeps_sample(gpu, oracle):
    """The range backward at the LiDAR configuration's per-world size: 3 worlds of 97 instances of the 138,000-triangle
    mesh, each on its own body, 65,536 rays a world (64 beams x 1024 azimuths); L = 0.5 sum (range - target)^2 against
    a target from poses moved by 0.02; the oracle's FP64 closed form and central differences (step 1e-4)."""
    import torch
    from neraulgraphics_amd.autograd import cast_range
    from neraulgraphics_amd.batch import BatchRenderer
    _say("range backward: lidar_sweeps, 3 worlds x 65536 rays")
    lc = workloads.make("lidar_sweeps", worlds=3)
    W, P, R = 3, len(lc["instances"]), lc["origins"].shape[1]
    poses = np.concatenate([lc["positions"], lc["quaternions"]], axis=-1).astype(np.float32)
    local = np.tile(np.array([0, 0, 0, 0, 0, 0, 1], np.float32), (P, 1))
    sc = np.repeat(lc["scales"][0][:, None], 3, axis=1).astype(np.float32)
    cfg = dict(lc, mode="bodies", bodies=P, body_of_instance=np.arange(P, dtype=np.int32), local_pose=local, scale=sc,
               body_poses=poses)
    r = BatchRenderer(64, 64, W, 1, meshes=lc["meshes"], instances=lc["instances"], instances_in_world=lc["counts"])
    r.bind_bodies(cfg["body_of_instance"], local, sc, P)
    moved = poses.copy()
    moved[..., :3] += np.random.default_rng(43).normal(0.0, 0.02, moved[..., :3].shape).astype(np.float32)
    r.set_body_poses(moved)
    target = r.cast_rays(lc["origins"], lc["directions"], lc["max_distance"])["distance"].copy()
    # the oracle's hits, and the rays whose hit is a tie (two triangles met at one distance, through their shared
    # edge: the same distance, another normal), where the two walks may keep different triangles; their loss weight is
    # set to zero on both sides, so both gradients go through the same triangles
    st = forward(cfg, oracle, outputs=())
    ref = oracle.cast_rays(st["nodes"], st["roots"], st["inv"], st["blas_root"], st["meshes"]["nodes"], st["meshes"]["tris"],
                           st["meshes"]["attrs"], st["world_off"], lc["origins"], lc["directions"], lc["max_distance"])
    r.set_body_poses(poses)
    fwd = r.cast_rays(lc["origins"], lc["directions"], lc["max_distance"])
    rep = compare_rays(fwd, ref, "lidar_sweeps sample, 3 worlds")
    tie = np.any(np.abs(fwd["normal"] - ref["normal"]) > 1e-5, axis=-1) | (fwd["object_id"] != ref["object_id"])
    _say(f"{rep}; tie rays left out of the loss: {int(tie.sum())}")
    p = torch.from_numpy(poses).cuda().requires_grad_(True)
    dist = cast_range(r, p, torch.from_numpy(lc["origins"]).cuda(), torch.from_numpy(lc["directions"]).cuda(), lc["max_distance"])
    tgt = torch.from_numpy(target).cuda()
    both = torch.isfinite(dist) & torch.isfinite(tgt) & ~torch.from_numpy(tie).cuda()
    diff = torch.where(both, dist - tgt, torch.zeros_like(dist))
    (0.5 * (diff * diff).sum()).backward()
    g_gpu = p.grad.cpu().numpy().astype(np.float64)
    assert int(both.sum()) > W * R // 10
    d32 = lc["directions"].reshape(-1, 3)
    unit = (d32 * (np.float32(1.0) / np.sqrt((d32 * d32).sum(axis=1, dtype=np.float32)))[:, None]).astype(np.float64)
    args = (lc["origins"].reshape(-1, 3).astype(np.float64), unit, diff.detach().cpu().numpy().reshape(-1).astype(np.float64),
            np.ones(W * R), ref["hit_inst"].reshape(-1), ref["hit_tri"].reshape(-1), np.repeat(np.arange(W), R).astype(np.int32),
            st["world_off"], P, cfg["body_of_instance"], local, sc, poses, st["meshes"]["tris"])
    g_cf = oracle.pose_grad(0, 0.0, *args)
    g_fd = oracle.pose_grad(1, 1e-4, *args)
    rel = np.linalg.norm(g_gpu - g_fd) / np.linalg.norm(g_fd)
    rel_cf = np.linalg.norm(g_gpu - g_cf) / np.linalg.norm(g_cf)
    _say(f"|GPU - FD(1e-4)| / |FD| = {rel:.2e}; |GPU - closed form| / |closed form| = {rel_cf:.2e}")
    assert rel < 1e-3 and rel_cf < 1e-4


def render_target(r, cfg):
    r.set_body_poses(cfg["target_poses"])
    d = r.render(cfg["cameras"], cfg["fov"], depth=True)["depth"]
    r.set_body_poses(cfg["body_poses"])
    return d


# ------------------------------------------------------------------------------------------ RGB picture --

def _rgb_pair(oracle, cfg, spheres=(), floor=None, lights=(), ambient=0.0, env=None, spp=1, frame=0):
    """The GPU's RGB picture and the oracle's for one scene."""
    import neraulgraphics_amd as ng
    floor = floor if floor is not None else ng.Plane()
    r = workloads.make_renderer(cfg, spheres=list(spheres), floor=floor, lights=list(lights))
    if ambient:
        r.set_ambient(ambient)
    if env is not None:
        r.set_environment(env)
    got = r.render(cfg["cameras"], cfg["fov"], spp, rgb=True, frame=frame)["rgb"]
    c = dict(cfg)
    c["spheres"] = np.array([[*s.center, s.radius] for s in spheres], np.float32) if spheres else None
    c["floor"] = np.array([floor.y, floor.half_extent_x, floor.z_min, floor.z_max], np.float32)
    c["shading"] = ng.shading_arrays(cfg["instances"], list(spheres), floor, list(lights), ambient, env)
    c["spp"], c["frame"] = spp, frame
    want = forward(c, oracle, outputs=("rgb",))["rgb"]
    return got.astype(np.int32), want.astype(np.int32)


def test_rgb_picture_matches_the_oracle(gpu, oracle):
    """tests/test_batch.cpp:149-181's bar for the picture: mean colour difference below 0.05 of a level per channel.
    Eight worlds of 32 instanced meshes, with materials of every kind (matte, glossy, mirror-like), two spheres, a
    checker floor, a soft sphere light, a spotlight with falloff and an ambient term, two samples a pixel."""
    import neraulgraphics_amd as ng
    _say("rgb parity: 8 worlds x 32 instances, 2 lights, spheres, floor, 2 spp")
    cfg = workloads.make("instanced_worlds", worlds=8)
    mats = [ng.Material((0.9, 0.2, 0.2), 0.9, 0.1, 0.0, 50.0), ng.Material((0.2, 0.8, 0.3), 0.7, 0.5, 0.3, 200.0),
            ng.Material((0.7, 0.7, 0.9), 0.2, 0.3, 0.8, 20.0)]
    cfg["instances"] = [ng.Instance(i.mesh, material=mats[k % 3]) for k, i in enumerate(cfg["instances"])]
    spheres = [ng.Sphere((0.5, -1.0, -9.0), 1.0, ng.Material((0.9, 0.9, 0.2), 0.8, 0.2, 0.5, 80.0)),
               ng.Sphere((-2.0, 1.0, -7.0), 0.7)]
    floor = ng.Plane(y=-4.0, half_extent_x=30.0, z_min=-60.0, z_max=5.0, checker_scale=0.5)
    lights = [ng.Light((2.0, 6.0, -4.0), intensity=1.2, radius=0.8),
              ng.Light((-3.0, 4.0, -2.0), intensity=2.0, falloff=8.0, direction=(0.3, -1.0, -0.8), spot_angle=0.6, color=(1.0, 0.8, 0.6))]
    got, want = _rgb_pair(oracle, cfg, spheres, floor, lights, ambient=0.15, spp=2, frame=3)
    diff = np.abs(got - want)
    mean = float(diff.mean())
    _say(f"mean colour difference {mean:.4f} of a level; {int((diff.max(axis=-1) > 2).sum())} of {diff.shape[0] * diff.shape[1] * diff.shape[2]} "
         f"pixels differ by more than 2 levels; bit-identical {float(np.mean(np.all(diff == 0, axis=-1))):.4f}")
    assert mean < 0.05
    assert len(np.unique(want.reshape(-1, 3), axis=0)) > 1000  # a picture, not a flat colour


def test_rgb_depth_id_of_the_full_body_worlds_batch_on_a_sample_of_worlds(gpu, oracle):
    """The RGB + depth + object-ID configuration at full size: 4096 worlds at 128x128, U{8..256} instances a world from
    64 body poses, a checker floor, one soft sphere light and ambient 0.1 (the bench's scene), one launch each; 16
    worlds spread over the batch against the oracle run on those worlds alone, to tests/test_batch.cpp:177's bar (mean
    difference below 0.05 of a level per channel). A pixel's random stream depends on its position and the frame only,
    so the sample sees what the whole batch sees."""
    import torch
    import neraulgraphics_amd as ng
    _say("full-size RGB parity: body_worlds, 4096 worlds, 16 sampled worlds")
    cfg = workloads.make("body_worlds")
    floor = ng.Plane(y=-6.0, half_extent_x=40.0, z_min=-60.0, z_max=20.0)
    lights = [ng.Light((0.0, 12.0, -6.0), intensity=1.2, radius=0.5)]
    r = workloads.make_renderer(cfg, floor=floor, lights=lights)
    r.set_ambient(0.1)
    out = r.render(torch.from_numpy(cfg["cameras"]).cuda(), cfg["fov"], 1, rgb=True, depth=True, object_ids=True, frame=3,
                   zero_copy=True)
    torch.cuda.synchronize()
    sample = np.unique(np.concatenate([[0, cfg["worlds"] - 1], np.random.default_rng(6).choice(cfg["worlds"], 14, replace=False)]))
    idx = torch.from_numpy(sample).cuda()
    got = {k: out[k][idx].cpu().numpy() for k in ("rgb", "depth", "object_ids")}
    sub = world_subset(cfg, sample)
    sub["floor"] = np.array([floor.y, floor.half_extent_x, floor.z_min, floor.z_max], np.float32)
    sub["spheres"] = None
    sub["shading"] = ng.shading_arrays(cfg["instances"], [], floor, lights, 0.1, None)
    sub["spp"], sub["frame"] = 1, 3
    ref = forward(sub, oracle, outputs=("rgb", "depth", "id"))
    rep = compare_images(dict(depth=got["depth"], object_id=got["object_ids"]), ref, "body_worlds RGB sample, depth and ID")
    diff = np.abs(got["rgb"].astype(np.int32) - ref["rgb"].astype(np.int32))
    mean = diff.reshape(-1, 3).mean(axis=0)
    _say(f"{rep}; mean colour difference per channel {mean}; bit-identical pixels {float(np.mean(np.all(diff == 0, axis=-1))):.5f}")
    assert np.all(mean < 0.05)
    assert rep["hits"] > 0.3 * rep["pixels"]
    # a picture: grey bodies in light and shadow over a checker floor under the sky gradient, not a flat colour
    assert len(np.unique(ref["rgb"].reshape(-1, 3), axis=0)) > 200


def test_two_pass_picture_gives_the_one_pass_bytes(gpu):
    """The picture drawn in two passes (the camera rays' hits first, then their shading and shadow rays: one sample a
    pixel, no surface reflective enough to bounce a ray) against the one-pass kernel, for the packet and hybrid walks:
    the body_worlds scene with its floor and soft light (24 worlds), 8 worlds of meshes with vertex normals, glossy
    materials, two spheres, a soft light and a spotlight, and the same meshes under the spotlight alone (scenes of one
    light take a shading kernel of their own). The hit pass walks the trees in another order, so a ray
    through an edge two triangles share may keep the other triangle (the parity bar's one exception): all but a
    handful of pixels are the same bytes."""
    import ctypes
    import neraulgraphics_amd as ng
    from neraulgraphics_amd import _lib
    from .test_gpu_normals import _smooth_meshes
    _say("two-pass picture against the one-pass kernel")
    bw = workloads.make("body_worlds", worlds=24)
    r1 = workloads.make_renderer(bw, floor=ng.Plane(y=-6.0, half_extent_x=40.0, z_min=-60.0, z_max=20.0),
                                 lights=[ng.Light((0.0, 12.0, -6.0), intensity=1.2, radius=0.5)])
    r1.set_ambient(0.1)
    iw = workloads.make("instanced_worlds", worlds=8)
    mats = [ng.Material((0.9, 0.2, 0.2), 0.9, 0.1, 0.0, 50.0), ng.Material((0.2, 0.8, 0.3), 0.7, 0.5, 0.0, 200.0)]
    iw["meshes"] = _smooth_meshes()
    iw["instances"] = [ng.Instance(k % 4, material=mats[k % 2]) for k in range(len(iw["instances"]))]
    r2 = workloads.make_renderer(iw, spheres=[ng.Sphere((0.5, -1.0, -9.0), 1.0), ng.Sphere((-2.0, 1.0, -7.0), 0.7)],
                                 floor=ng.Plane(y=-4.0, half_extent_x=30.0, z_min=-60.0, z_max=5.0, checker_scale=0.5),
                                 lights=[ng.Light((2.0, 6.0, -4.0), intensity=1.2, radius=0.8),
                                         ng.Light((-3.0, 4.0, -2.0), intensity=2.0, falloff=8.0, direction=(0.3, -1.0, -0.8),
                                                  spot_angle=0.6, color=(1.0, 0.8, 0.6))])
    r2.set_ambient(0.15)
    # one light again (the shading pass's one-light build), here a coloured spotlight with a distance falloff
    r4 = workloads.make_renderer(iw, spheres=[ng.Sphere((0.5, -1.0, -9.0), 1.0)],
                                 floor=ng.Plane(y=-4.0, half_extent_x=30.0, z_min=-60.0, z_max=5.0, checker_scale=0.5),
                                 lights=[ng.Light((-3.0, 4.0, -2.0), intensity=2.0, falloff=8.0, direction=(0.3, -1.0, -0.8),
                                                  spot_angle=0.6, color=(1.0, 0.8, 0.6))])
    r4.set_ambient(0.15)
    for name, r, cfg in (("body_worlds", r1, bw), ("vertex normals and spheres", r2, iw),
                         ("vertex normals, one spotlight", r4, iw)):
        for walk in ("packet", "hybrid"):
            out = {}
            for two in (True, False):
                r.set_trace_config(**WALKS[walk], rgb_two_pass=two)
                out[two] = r.render(cfg["cameras"], cfg["fov"], 1, rgb=True, frame=5)["rgb"].astype(np.int32)
                passes = ctypes.c_int32()
                _lib.check(r._lib.ngx_debug_rgb_passes(r._h, ctypes.byref(passes)))
                assert passes.value == (2 if two else 1), (name, walk, two, passes.value)
            differ = np.any(out[True] != out[False], axis=-1)
            mean = float(np.abs(out[True] - out[False]).mean())
            _say(f"{name}, {walk} walk: {int(differ.sum())} of {differ.size} pixels differ; mean difference {mean:.6f}")
            assert int(differ.sum()) <= max(2, differ.size // 20000) and mean < 0.05, (name, walk)
            assert len(np.unique(out[True].reshape(-1, 3), axis=0)) > 200
    # a surface that reflects (reflectivity 0.3) makes paths longer than one hit: one pass, whatever the setting
    mirror = workloads.make("instanced_worlds", worlds=2)
    mirror["instances"] = [ng.Instance(i.mesh, material=ng.Material((0.7, 0.7, 0.9), 0.5, 0.3, 0.3, 20.0)) if k == 0 else i
                           for k, i in enumerate(mirror["instances"])]
    r3 = workloads.make_renderer(mirror, lights=[ng.Light((2.0, 6.0, -4.0), intensity=1.2)])
    r3.render(mirror["cameras"], mirror["fov"], 1, rgb=True)
    passes = ctypes.c_int32()
    _lib.check(r3._lib.ngx_debug_rgb_passes(r3._h, ctypes.byref(passes)))
    assert passes.value == 1


def test_rgb_known_answers_on_the_device(gpu, oracle):
    """The oracle's closed-form cases (tests/test_oracle_rgb.py) through the device: a sky of one colour, a floor lit
    from straight above, a shadowed point, a mirror floor."""
    import neraulgraphics_amd as ng
    from .test_oracle_rgb import scene, srgb_byte
    _say("rgb known answers on the device")

    def gpu_rgb(cfg, spheres=(), floor=None, lights=(), ambient=0.0, env=None):
        got, want = _rgb_pair(oracle, cfg, spheres, floor, lights, ambient, env, spp=cfg["spp"])
        assert np.array_equal(got, want) or np.abs(got - want).mean() < 0.05
        return got[0]

    c = (0.2, 0.5, 0.8)
    env = ng.Environment(color_up=c, color_down=c, intensity=0.75)
    sky = gpu_rgb(scene(env=env), env=env)
    assert np.all(sky.reshape(-1, 3) == srgb_byte(np.array(c, np.float32) * np.float32(0.75)))
    fm = ng.Material((1.0, 1.0, 1.0), 0.8, 0.0, 0.0, 1.0)
    floor = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0, material=fm)
    light = ng.Light((0.0, 3.0, 0.0), intensity=1.5)
    down = ng.look_at((0, 0, 0), (0, -1, 0), up=(0, 0, -1))
    lit = gpu_rgb(scene(lights=[light], floor=floor, cam=down, size=33), floor=floor, lights=[light])
    want = srgb_byte(np.float32(0.3) * np.array([1.0, 0.7, 0.3], np.float32) * np.float32(1.5 * 0.8))
    assert np.all(np.abs(lit[16, 16] - want) <= 1)
    blocker = ng.Sphere((0.0, 1.5, 0.0), 0.5)
    dark = gpu_rgb(scene(lights=[light], spheres=[blocker], floor=floor, cam=down, size=33), [blocker], floor, [light])
    assert np.all(dark[16, 16] == 0)
    mirror = ng.Plane(y=-1.0, half_extent_x=50.0, z_min=-50.0, z_max=50.0, checker_scale=0.0,
                      material=ng.Material((1.0, 1.0, 1.0), 0.0, 0.0, 1.0, 1.0))
    env2 = ng.Environment(color_up=(0.3, 0.6, 0.1), color_down=(0.3, 0.6, 0.1))
    m = gpu_rgb(scene(floor=mirror, cam=down, env=env2), floor=mirror, env=env2)
    assert np.all(m.reshape(-1, 3) == srgb_byte(np.array((0.3, 0.6, 0.1), np.float32)))
