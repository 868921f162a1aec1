"""Throughput of the batched sensor-rendering path on one MI355X.

A step is one pass of the hot path over one batch: new body poses for every world (one kernel composes them
in FP64 and rebuilds every world's top-level BVH), then depth + object-ID images of every view (one trace
kernel). The workload is 4096 worlds at 128x128 with U{8..256} instances per world out of a 256-instance
template carried by 64 bodies (neraulgraphics_amd.workloads, body_worlds); four resident pose sets are
cycled so every step has new poses. Inputs are resident in HBM before timing; the step is replayed from a
HIP graph. Before timing, the trace configuration is autotuned on the benchmark's own cameras (--no-tune keeps
the defaults).

Also reported, on the same JSON line: the same batch with the RGB picture as well (one soft sphere light, a floor);
LiDAR ray queries (lidar_sweeps: 1024 worlds x 65,536 rays,
rays and results resident on the device); the forward + pose-backward pass (posed_bodies: 1024 worlds,
64x64 depth, 32 bodies) and its ratio to the forward alone; the render kernel against the HBM roofline;
and the reference CPU algorithm (this repository's C restatement of the reference CPU backend: FP64 body
poses, per-world SAH top level with refit, the same traversal, rows of views in chunks of 4 over threads)
timed on this host on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-graph] [--no-tune] [--no-lidar] [--no-rgb] [--no-bwd]
                [--only LEG] [--trace-config JSON]"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "images/sec @4096 worlds 128x128 depth+ID; fwd+bwd images/sec; % HBM BW"
HBM_PEAK_GBS = 8000.0


def render_bytes(views, width, height, instances, outputs_bytes_per_pixel=8):
    """Algorithmic bytes of one render launch: the outputs (4 B depth + 4 B ID per pixel), the cameras
    (48 B per view), and each world's top-level tree and instance records read at least once
    (64 B wide node + 64 B record per instance)."""
    return outputs_bytes_per_pixel * views * width * height + 48 * views + 128 * instances


def picture_bytes(views, width, height, instances):
    """Algorithmic bytes of the RGB picture drawn in two passes: the hit pass writes a 16 B record a pixel and the
    shading pass reads it back and writes 3 B of colour a pixel; each pass reads the cameras (48 B per view) and each
    world's top-level tree and instance records once (64 + 64 B per instance)."""
    return 35 * views * width * height + 2 * (48 * views + 128 * instances)


def build_bytes(worlds, bodies, instances):
    """Algorithmic bytes of one top-level build launch: body poses in (28 B per body), per instance the
    placed matrix, the 64 B instance record, the reference-layout leaf and interior node (2 x 32 B) and the
    64 B wide node out."""
    return 28 * worlds * bodies + instances * (48 + 64 + 64 + 64)


def timed_steps(step, steps, warmup, torch):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def reference_blocks(step, torch, warm_s=1.5, blocks=30, block_s=0.05):
    """The reference's timing protocol (bench/timing.py:43-72 of the reference): steps for at least warm_s seconds
    untimed, then `blocks` blocks of about block_s seconds each, every block between two device syncs; per-step
    times of the blocks' median, spread and extremes."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < warm_s or n < 3:
        step()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / n
    k = int(min(2000, max(1, round(block_s / per))))
    ms = sorted(timed_steps(step, k, 0, torch) * 1e3 for _ in range(blocks))
    q = np.percentile(ms, [10, 50, 90])
    return dict(median_ms=float(q[1]), p10_ms=float(q[0]), p90_ms=float(q[2]), min_ms=ms[0], max_ms=ms[-1], blocks=blocks,
                steps_per_block=k, warmup_s=warm_s)


def backward_bytes(views, width, height, instances, worlds, bodies, parts):
    """Algorithmic bytes of one depth-backward launch: the forward's depth and dL/d depth read (4 + 4 B per pixel),
    the cameras (48 B per view), each world's top-level tree and instance records read once (64 + 64 B per instance),
    the body poses (28 B per body) and the per-workgroup partial sums written (28 B per body and part)."""
    return 8 * views * width * height + 48 * views + 128 * instances + 28 * worlds * bodies + 28 * views * parts * bodies


def fixed_config(r, args):
    """Applies --trace-config (a JSON object of BatchRenderer.set_trace_config's arguments, such as the trace_config
    a bench line reports) in place of autotuning; True when it did."""
    if not args.trace_config:
        return False
    c = {k: v for k, v in json.loads(args.trace_config).items() if k != "ms"}
    r.set_trace_config(**c)
    print(f"# trace configuration given: {r.trace_config}", file=sys.stderr, flush=True)
    return True


def bench_body_worlds(args, torch):
    from neraulgraphics_amd import workloads
    cfg = workloads.make("body_worlds")
    r = workloads.make_renderer(cfg)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    tuned = None
    if args.tile_w:
        r.set_trace_config(args.tile_w, not args.no_xcd)
    elif fixed_config(r, args):
        pass
    elif not args.no_tune:
        # the autotuner times its candidates on these cameras (outside the timed region) and keeps the fastest
        tuned = r.autotune(cams, cfg["fov"], depth=True, object_ids=True, reps=2, cache=False)
        print(f"# autotuned: {tuned}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    pose_sets = [torch.from_numpy(cfg["body_poses"]).cuda()]
    for _ in range(3):
        p = cfg["body_poses"].copy()
        p[..., :3] += rng.normal(0, 0.05, p[..., :3].shape).astype(np.float32)
        pose_sets.append(torch.from_numpy(p).cuda())
    fov = cfg["fov"]
    k = [0]

    def step_direct():
        r.set_body_poses(pose_sets[k[0] % 4])
        r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
        k[0] += 1

    # per-kernel times with events on the stream the kernels are launched on
    for _ in range(3):
        step_direct()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    for i in range(args.steps):
        ev[i][0].record()
        r.set_body_poses(pose_sets[i % 4])
        ev[i][1].record()
        r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
        ev[i][2].record()
    torch.cuda.synchronize()
    build_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    render_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))

    graphs = None
    if not args.no_graph:
        try:
            graphs = []
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                for _ in range(2):
                    step_direct()
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            for i in range(4):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    r.set_body_poses(pose_sets[i])
                    r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
                graphs.append(g)
            torch.cuda.synchronize()
        except Exception as e:  # graph capture is an optimization; direct launches are the same work
            print(f"# hip graph capture unavailable ({e}); timing direct launches", file=sys.stderr, flush=True)
            graphs = None

    if graphs:
        def step():
            graphs[k[0] % 4].replay()
            k[0] += 1
    else:
        step = step_direct
    sec = timed_steps(step, args.steps, args.warmup, torch)
    # then the reference's protocol on the same step: 1.5 s of warm-up, 30 blocks of about 50 ms, the median block
    blocks = reference_blocks(step, torch)
    views = cfg["worlds"] * cfg["cameras_per_world"]
    inst = int(np.sum(cfg["counts"]))
    # correctness guard on the timed buffers: the last step's images are the ones of its poses
    out = r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
    torch.cuda.synchronize()
    hit_frac = float(torch.isfinite(out["depth"]).float().mean())
    return dict(cfg=cfg, sec=sec, build_ms=build_ms, render_ms=render_ms, views=views, instances=inst, graph=bool(graphs),
                hit_frac=hit_frac, renderer=r, trace_config=r.trace_config, blocks=blocks)


def bench_instanced_worlds(args, torch):
    """4096 worlds at 64x64, depth + object ID, 32 instances a world over three meshes, new instance matrices every
    step (device arrays), steps replayed from HIP graphs over four resident matrix sets; the trace configuration
    autotuned on these cameras first. The reference's README (README.md:21, :93-95) quotes "a million images a second"
    at 64x64 on an RTX 5090."""
    from neraulgraphics_amd import workloads
    cfg = workloads.make("instanced_worlds")
    r = workloads.make_renderer(cfg)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    if not fixed_config(r, args) and not args.no_tune:
        r.autotune(cams, cfg["fov"], depth=True, object_ids=True, reps=2, cache=False)
    rng = np.random.default_rng(8)
    sets = []
    for i in range(4):
        m = cfg["matrices"].copy()
        if i:
            m[..., 3::4] += rng.normal(0, 0.05, m[..., 3::4].shape).astype(np.float32)  # move every instance a little
        sets.append(torch.from_numpy(np.ascontiguousarray(m)).cuda())
    fov = cfg["fov"]
    k = [0]

    def step_direct():
        r.set_instance_transforms(sets[k[0] % 4])
        r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
        k[0] += 1

    for _ in range(3):
        step_direct()
    torch.cuda.synchronize()
    step = step_direct
    if not args.no_graph:
        try:
            graphs = []
            for i in range(4):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    r.set_instance_transforms(sets[i])
                    r.render(cams, fov, depth=True, object_ids=True, zero_copy=True)
                graphs.append(g)
            torch.cuda.synchronize()

            def step():
                graphs[k[0] % 4].replay()
                k[0] += 1
        except Exception as e:  # graph capture is an optimization; direct launches are the same work
            print(f"# hip graph capture unavailable ({e}); timing direct launches", file=sys.stderr, flush=True)
    sec = timed_steps(step, max(10, args.steps), args.warmup, torch)
    return dict(images_per_s=cfg["worlds"] / sec, ms=sec * 1e3, reference_claim_images_per_s=1.0e6, trace_config=r.trace_config)


def bench_lidar(args, torch):
    """LiDAR: 1024 worlds x 65,536 rays over 100,000 moving instances of a 138,000-triangle grid; a step places the
    instances (new matrices from the device, four resident sets cycled, every instance moved a little) and answers the
    query; rays resident on the device, results left there (zero-copy)."""
    from neraulgraphics_amd import workloads
    cfg = workloads.make("lidar_sweeps")
    r = workloads.make_renderer(cfg)
    o = torch.from_numpy(cfg["origins"]).cuda()
    d = torch.from_numpy(cfg["directions"]).cuda()
    if not fixed_config(r, args) and not args.no_tune:
        tuned = r.autotune_rays(o, d, cfg["max_distance"], reps=2)
        print(f"# LiDAR autotuned: {tuned}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(9)
    sets = []
    for i in range(4):
        m = cfg["matrices"].copy()
        if i:
            m[..., 3::4] += rng.normal(0, 0.05, m[..., 3::4].shape).astype(np.float32)
        sets.append(torch.from_numpy(np.ascontiguousarray(m)).cuda())
    k = [0]

    def step():
        r.set_instance_transforms(sets[k[0] % 4])
        r.cast_rays(o, d, cfg["max_distance"])
        k[0] += 1

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    steps = max(3, args.steps // 4)
    t = timed_steps(step, steps, 1, torch)
    n = o.shape[0] * o.shape[1]
    r.set_instance_transforms(sets[0])
    out = r.cast_rays(o, d, cfg["max_distance"])
    torch.cuda.synchronize()
    return dict(rays_per_s=n / t, ms=t * 1e3, rays=n, hit_frac=float(torch.isfinite(out["distance"]).float().mean()),
                trace_config=r.trace_config)


def bench_rgb(args, torch):
    """The same 4096-world batch with the RGB picture too (the reference's configuration is RGB + depth + object ID):
    a checker floor under the bodies, one sphere light of radius 0.5 (soft shadows), ambient 0.1, one sample a pixel,
    new body poses every step. Materials are the defaults (diffuse 0.8, a little specular)."""
    from neraulgraphics_amd import workloads
    from neraulgraphics_amd.geometry import Light, Plane
    cfg = workloads.make("body_worlds")
    floor = Plane(y=-6.0, half_extent_x=40.0, z_min=-60.0, z_max=20.0)
    r = workloads.make_renderer(cfg, floor=floor, lights=[Light((0.0, 12.0, -6.0), intensity=1.2, radius=0.5)])
    r.set_ambient(0.1)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    tuned = None
    if not fixed_config(r, args) and not args.no_tune:
        tuned = r.autotune(cams, cfg["fov"], depth=True, object_ids=True, rgb=True, reps=1, cache=False)
        print(f"# RGB autotuned: {tuned}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    pose_sets = [torch.from_numpy(cfg["body_poses"]).cuda()]
    for _ in range(3):
        p = cfg["body_poses"].copy()
        p[..., :3] += rng.normal(0, 0.05, p[..., :3].shape).astype(np.float32)
        pose_sets.append(torch.from_numpy(p).cuda())
    k = [0]

    def step():
        r.set_body_poses(pose_sets[k[0] % 4])
        r.render(cams, cfg["fov"], 1, rgb=True, depth=True, object_ids=True, frame=k[0], zero_copy=True)
        k[0] += 1

    steps = max(5, args.steps // 2)
    sec = timed_steps(step, steps, max(2, args.warmup // 2), torch)
    # the render call's kernels (depth + ID, then the picture's passes), timed by the library's events on their stream
    kernel_ms = []
    for _ in range(steps):
        step()
        torch.cuda.synchronize()
        kernel_ms.append(r.last_kernel_ms)
    return dict(images_per_s=cfg["worlds"] / sec, ms=sec * 1e3, trace_config=r.trace_config, kernel_ms=float(np.mean(kernel_ms)),
                views=cfg["worlds"], instances=int(np.sum(cfg["counts"])))


# The reference's own published batch-renderer timings on this scene (bench/results/run2/summary.txt:20-22, :41-43:
# neraulgraphics 0.5.0 on an RTX 5090, 4096 worlds, RGB only, ms per frame), keyed by (bodies per world, resolution).
REFERENCE_SCENE_MS = {(8, 64): 1.4055, (8, 128): 4.0413, (32, 64): 3.9754, (32, 128): 12.0280}
# The reference's README (README.md:93-95): 4096 worlds at 128x128, RGB, depth and object IDs, new poses every step,
# "about 6 ms" on an RTX 5090.
REFERENCE_README_RGB_DEPTH_ID_MS = 6.0


def bench_falling_bodies(args, torch):
    """The reference benchmark's own scene (bench/scene.py, bench/run_ours.py) through MJWarpBatchRenderer: 4096 worlds
    of 8 or 32 falling balls and boxes, per frame new geom frames (GPU arrays) composed into the top-level build, then
    the RGB picture (one hard-shadow light, one sample a pixel) into the renderer's buffer. Timed as the reference times
    it (bench/timing.py:43-72): warm-up, then blocks of frames of about 50 ms between two syncs; the median block."""
    from types import SimpleNamespace
    from neraulgraphics_amd import workloads
    from neraulgraphics_amd.geometry import Light
    from neraulgraphics_amd.simulators import MJWarpBatchRenderer
    out = {}
    for nbody in (8, 32):
        if args.scene_case and not args.scene_case.startswith(f"{nbody}_"):
            continue
        fb = workloads.falling_bodies(nbody, 4096)
        xpos = torch.from_numpy(fb["geom_xpos"]).cuda()
        xmat = torch.from_numpy(fb["geom_xmat"]).cuda()
        frames = [SimpleNamespace(geom_xpos=xpos[f], geom_xmat=xmat[f]) for f in range(xpos.shape[0])]
        cams = torch.from_numpy(np.tile(fb["camera"].reshape(1, 12), (4096, 1))).cuda()
        light = Light(fb["light"], intensity=1.0, radius=0.0)
        for res, aov in ((64, False), (128, False)) + (((128, True),) if nbody == 8 else ()):
            key = f"{nbody}_bodies_{res}" + ("_rgb_depth_id" if aov else "")
            if args.scene_case and key != args.scene_case:
                continue
            view = MJWarpBatchRenderer(fb["model"], 4096, res, res, lights=[light], fov=fb["fov"])
            k = [0]

            def frame():
                view.render(frames[k[0] % len(frames)], cams, rgb=True, depth=aov, object_ids=aov, zero_copy=True)
                k[0] += 1

            per = timed_steps(frame, 10, 10, torch)
            block = int(min(2000, max(1, np.ceil(0.05 / per))))
            reps = sorted(timed_steps(frame, block, 0, torch) for _ in range(args.scene_reps))
            ms = reps[len(reps) // 2] * 1e3
            ref = REFERENCE_README_RGB_DEPTH_ID_MS if aov else REFERENCE_SCENE_MS[(nbody, res)]
            out[key] = dict(ms=ms, images_per_s=4096 / ms * 1e3, reference_rtx5090_ms=ref, speedup_vs_reference=ref / ms)
            print(f"# falling_bodies {key}: {ms:.4f} ms/frame (reference on an RTX 5090: {ref} ms)", file=sys.stderr, flush=True)
            del view
    return out


def bench_posed_bodies(args, torch):
    from neraulgraphics_amd import workloads
    from neraulgraphics_amd.autograd import render_depth
    cfg = workloads.make("posed_bodies")
    r = workloads.make_renderer(cfg)
    cams = torch.from_numpy(cfg["cameras"]).cuda()
    r.set_body_poses(cfg["target_poses"])
    target = r.render(cams, cfg["fov"], depth=True, zero_copy=True)["depth"].clone()
    poses = torch.from_numpy(cfg["body_poses"]).cuda()
    ones = torch.ones_like(target)

    def fwd():
        with torch.no_grad():
            render_depth(r, poses, cams, cfg["fov"])

    def fwd_bwd():
        # L = 0.5 sum (depth - target)^2 over pixels that hit in both; its gradient is seeded per pixel (ones), so no
        # reduction runs (a reduction's memset nodes are kept out of the captured graph)
        p = poses.detach().requires_grad_(True)
        d = render_depth(r, p, cams, cfg["fov"])
        both = torch.isfinite(d) & torch.isfinite(target)
        diff = torch.where(both, d - target, torch.zeros_like(d))
        (0.5 * (diff * diff)).backward(ones)
        return p.grad

    if not fixed_config(r, args) and not args.no_tune:
        # the forward's walk and tiles on these cameras, then the backward kernel's variant on a real gradient
        r.autotune(cams, cfg["fov"], depth=True, object_ids=False, reps=2, cache=False)
        d = render_depth(r, poses, cams, cfg["fov"]).detach()
        both = torch.isfinite(d) & torch.isfinite(target)
        g = torch.where(both, d - target, torch.zeros_like(d)).contiguous()
        tuned = r.autotune_backward(poses, cams, cfg["fov"], d, g, reps=3)
        print(f"# backward autotuned: {tuned}", file=sys.stderr, flush=True)

    # both steps replayed from HIP graphs (one forward, or one forward, loss and backward), captured after eager warm-up
    # runs; the replayed gradient is checked against the eager one before timing
    graphs = {}
    if not args.no_graph:
        try:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(2):
                    fwd()
                    want = fwd_bwd().clone()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    out = fn()
                graphs[name] = (g, out)
            for _ in range(3):  # every replay, with eager work between, gives the eager gradient
                graphs["fwd_bwd"][0].replay()
                torch.cuda.synchronize()
                if not torch.equal(graphs["fwd_bwd"][1], want):
                    raise RuntimeError("the replayed gradient differs from the eager one")
                fwd()
                torch.cuda.synchronize()
        except Exception as e:  # graph capture is an optimization; the eager steps are the same work
            print(f"# fwd+bwd hip graph capture unavailable ({e}); timing eager steps", file=sys.stderr, flush=True)
            graphs = {}
    step_f = (lambda: graphs["fwd"][0].replay()) if graphs else fwd
    step_fb = (lambda: graphs["fwd_bwd"][0].replay()) if graphs else fwd_bwd
    steps = max(5, args.steps // 2)
    t_f = timed_steps(step_f, steps, max(2, args.warmup), torch)
    t_fb = timed_steps(step_fb, steps, max(2, args.warmup), torch)
    # the backward kernel alone, timed by the library's events on the stream it runs on
    kernel_ms = []
    for _ in range(steps):
        fwd_bwd()
        torch.cuda.synchronize()
        kernel_ms.append(r.last_kernel_ms)
    parts = max(1, min(16, ((cfg["width"] + 7) // 8) * ((cfg["height"] + 7) // 8) // 8))
    nbytes = backward_bytes(cfg["worlds"], cfg["width"], cfg["height"], cfg["worlds"] * len(cfg["instances"]), cfg["worlds"],
                            cfg["bodies"], parts)
    return dict(fwd_images_per_s=cfg["worlds"] / t_f, fwd_bwd_images_per_s=cfg["worlds"] / t_fb, ratio=t_f / t_fb,
                worlds=cfg["worlds"], backward_kernel_ms=float(np.mean(kernel_ms)), backward_bytes=nbytes,
                trace_config=r.trace_config, hip_graph=bool(graphs))


def cpu_threads():
    """The reference CPU backend runs on std::thread::hardware_concurrency() threads (src/cpu_backend.cpp:94-101 of the
    reference). Here: every CPU this process may run on (its affinity mask), at most 16, the host cores one GPU's
    share of a node gets."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(seconds=12.0):
    """The reference CPU algorithm on this host, on a sample of the body_worlds workload's worlds."""
    from neraulgraphics_amd import workloads
    from oracle.oracle import Oracle, SahTlas
    o = Oracle()
    threads = cpu_threads()
    sample = 64
    cfg = workloads.make("body_worlds", worlds=sample)
    counts = cfg["counts"]
    world_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    meshes = o.build_meshes(cfg["meshes"])
    tmpl_mesh = np.array([i.mesh for i in cfg["instances"]], np.int32)
    local_bounds = meshes["bounds"][tmpl_mesh]
    tmpl_of = np.concatenate([np.arange(c) for c in counts]).astype(np.int32)
    blas_root = meshes["roots"][tmpl_mesh][tmpl_of]
    tlas = SahTlas(o, world_off)
    rng = np.random.default_rng(3)
    images, t0 = 0, time.perf_counter()
    while True:
        poses = cfg["body_poses"].copy()
        poses[..., :3] += rng.normal(0, 0.05, poses[..., :3].shape).astype(np.float32)
        mats = o.body_poses_to_matrices(world_off, cfg["bodies"], cfg["body_of_instance"], cfg["local_pose"], cfg["scale"], poses)
        nodes, roots, inv = tlas.update(mats, local_bounds)
        o.render(nodes, roots, inv, blas_root, meshes["nodes"], meshes["tris"], meshes["attrs"], world_off, cfg["cameras"],
                 cfg["fov"], cfg["width"], cfg["height"], threads=threads, outputs=("depth", "id"))
        images += sample
        el = time.perf_counter() - t0
        if el > seconds:
            break
    out = dict(value=images / el, unit="images/s", cores=threads, kind="port",
               sample=f"{images} images of body_worlds ({sample} worlds a step, 128x128 depth+ID, body poses, "
                      f"per-world SAH top level with refit) in {el:.1f} s on {threads} threads "
                      f"(the host shows {os.cpu_count()} CPUs)")
    out.update(reference_core_calibration(o, cfg, nodes, roots, inv, blas_root, meshes))
    return out


def reference_core_calibration(o, cfg, nodes, roots, inv, blas_root, meshes, views=4):
    """The port's traversal against the reference's own compiled traversal core (src/mesh_dev.h of the reference, built
    into oracle/_ref where the reference's sources are present) on the same camera rays, one thread each: the ratio
    relates the CPU baseline (the port) to the reference CPU backend's own speed. {} where oracle/_ref is absent."""
    from oracle.oracle import Ref, ref_available
    if not ref_available():
        return {}
    ref = Ref()
    w, h = cfg["width"], cfg["height"]
    orig, dirs, _ = ref.camera_rays(cfg["cameras"][:views], cfg["fov"], w, h)
    orig, dirs = orig.reshape(-1, 3), dirs.reshape(-1, 3)
    ray_world = np.repeat(np.arange(views, dtype=np.int32), w * h)
    t0 = time.perf_counter()
    r = ref.mesh_intersect(nodes, roots, inv, blas_root, meshes["nodes"], meshes["tris"], ray_world, orig, dirs)
    t_ref = time.perf_counter() - t0
    # the port over the same rays: the first `views` worlds' rays as a ray query
    world_off = np.concatenate([[0], np.cumsum(cfg["counts"])]).astype(np.int32)[:views + 1]
    t0 = time.perf_counter()
    p = o.cast_rays(nodes, roots, inv, blas_root, meshes["nodes"], meshes["tris"], meshes["attrs"], world_off,
                    orig.reshape(views, -1, 3), dirs.reshape(views, -1, 3), 1e30, threads=1)
    t_port = time.perf_counter() - t0
    same = int(np.sum((r["inst"] >= 0) == np.isfinite(p["distance"].reshape(-1))))
    return dict(reference_core_ratio=t_port / t_ref,
                reference_core_note=f"{len(ray_world)} camera rays of {views} worlds, one thread each: the reference's compiled "
                                    f"traversal core {t_ref:.2f} s, the port {t_port:.2f} s ({same} of {len(ray_world)} rays "
                                    "agree on hit or miss)")


def _profile_summary(name):
    """A traffic summary committed under profiles/ (HBM bytes per launch from rocprofv3 counters), or {}."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


# The trace-configuration fields a profiled kernel's traffic depends on, and its variant fields, each with the variant
# a 0 stands for (the kernels' built-in defaults).
_TRAFFIC_FIELDS = {
    "render": (("tile_w", "xcd_remap", "packet", "pair_limit", "persistent", "lane_min_triangles"), (("render_blocks", 7),)),
    "backward": (("packet", "pair_limit", "lane_min_triangles"), (("backward_blocks", 5),)),
    # the RGB + depth + ID call: the depth kernel and the hit pass take render_blocks, the shading pass rgb_blocks
    # (5 for a scene of one light)
    "rgb": (("tile_w", "xcd_remap", "packet", "pair_limit", "persistent", "lane_min_triangles", "rgb_two_pass"),
            (("render_blocks", 7), ("rgb_blocks", 5))),
}


def _traffic(summary, kind, chosen):
    """HBM bytes per launch from the summary (or one of its `variants`) profiled under the configuration this run
    chose, compared on the fields the kernel depends on; else the summary's own figure. Returns (bytes, matched)."""
    fields, variants = _TRAFFIC_FIELDS[kind]

    def key(c):
        return tuple(int(c.get(f, 0)) for f in fields) + tuple(int(c.get(v, 0)) or d for v, d in variants)

    for entry in [summary] + list(summary.get("variants", [])):
        if entry.get("trace_config") is not None and key(entry["trace_config"]) == key(chosen):
            return entry.get("hbm_bytes_per_launch"), True
    return summary.get("hbm_bytes_per_launch"), False


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-bwd", action="store_true")
    ap.add_argument("--tile-w", type=int, default=0)
    ap.add_argument("--no-xcd", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default trace configuration instead of autotuning")
    ap.add_argument("--trace-config", default="", help="a JSON object of set_trace_config's arguments (e.g. a line's "
                                                        "trace_config) used in place of autotuning, for profiling it")
    ap.add_argument("--no-lidar", action="store_true")
    ap.add_argument("--no-instanced", action="store_true", help="skip the 64x64 instanced_worlds leg")
    ap.add_argument("--no-rgb", action="store_true")
    ap.add_argument("--no-scene", action="store_true", help="skip the reference benchmark's falling-bodies scene")
    ap.add_argument("--scene-reps", type=int, default=15)
    ap.add_argument("--scene-case", default="", help="only this case of the falling-bodies scene, e.g. 8_bodies_128")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 512 MiB three times with a plain copy kernel, to scale HBM counters")
    ap.add_argument("--only", default="", help="run only this leg (for profiling): body_worlds, instanced, lidar, bwd, rgb, "
                                               "scene or cpu; the line then holds that leg's fields")
    args = ap.parse_args()
    if args.only:
        legs = ("body_worlds", "instanced", "lidar", "bwd", "rgb", "scene", "cpu")
        if args.only not in legs:
            ap.error(f"--only takes one of {', '.join(legs)}")
        args.no_instanced = args.only != "instanced"
        args.no_lidar = args.only != "lidar"
        args.no_bwd = args.only != "bwd"
        args.no_rgb = args.only != "rgb"
        args.no_scene = args.only != "scene"
        args.no_cpu = args.only != "cpu"
    if args.gpus != 1:
        print(json.dumps({"error": f"this benchmark runs on one GPU; --gpus {args.gpus} is not supported: the path shards "
                                   "by world with no exchange, so N GPUs run N independent batches"}), flush=True)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}), flush=True)
        sys.exit(1)
    if args.copy_calibration:
        src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy calibration: 3 copies of 536870912 bytes", file=sys.stderr, flush=True)
        del src, dst
    c2 = None
    if not args.only or args.only == "body_worlds":
        print("# body_worlds: 4096 worlds 128x128 depth+ID", file=sys.stderr, flush=True)
        c2 = bench_body_worlds(args, torch)
        b = c2["blocks"]
        print(f"# step {c2['sec'] * 1e3:.3f} ms (build {c2['build_ms']:.3f} ms, render {c2['render_ms']:.3f} ms); "
              f"reference protocol: median {b['median_ms']:.3f} ms, p10-p90 {b['p10_ms']:.3f}-{b['p90_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    c1 = None
    if not args.no_instanced:
        print("# instanced_worlds: 4096 worlds 64x64 depth+ID", file=sys.stderr, flush=True)
        c1 = bench_instanced_worlds(args, torch)
        print(f"# instanced_worlds {c1['ms']:.3f} ms a step, {c1['images_per_s']:.0f} images/s", file=sys.stderr, flush=True)
    c3 = None
    if not args.no_lidar:
        print("# lidar_sweeps: LiDAR ray queries", file=sys.stderr, flush=True)
        c3 = bench_lidar(args, torch)
    c4 = None
    if not args.no_bwd:
        print("# posed_bodies: forward and forward+backward", file=sys.stderr, flush=True)
        c4 = bench_posed_bodies(args, torch)
    crgb = None
    if not args.no_rgb:
        print("# body_worlds with RGB + depth + ID", file=sys.stderr, flush=True)
        crgb = bench_rgb(args, torch)
    cscene = None
    if not args.no_scene:
        print("# falling_bodies: the reference benchmark's scene, RGB", file=sys.stderr, flush=True)
        cscene = bench_falling_bodies(args, torch)
    cpu = None
    if not args.no_cpu:
        print("# CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_seconds)
    line = {"metric": METRIC, "value": None, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic"}
    if c2:
        sec = c2["sec"]
        rb = render_bytes(c2["views"], 128, 128, c2["instances"])
        achieved = rb / (c2["render_ms"] * 1e-3) / 1e9
        kernel = "render_persistent_kernel" if c2["trace_config"].get("persistent") else "render_kernel"
        summary = _profile_summary(f"{kernel}_body_worlds_traffic.json")
        traffic, same = _traffic(summary, "render", c2["trace_config"])
        bb = build_bytes(4096, 64, c2["instances"])
        line.update({
            "value": c2["views"] / sec,
            "ms_per_step": sec * 1e3,
            "config": {"workload": "body_worlds: 4096 worlds, 128x128 depth+object-ID, U{8..256} of a 256-instance "
                                   "template per world, 64 body poses per world, new poses every step",
                       "worlds": 4096, "width": 128, "height": 128, "instances": c2["instances"], "outputs": "depth+ID",
                       "hip_graph": c2["graph"], "trace_config": c2["trace_config"]},
            "timing_blocks": c2["blocks"],
            "rays_per_s": c2["views"] * 128 * 128 / sec,
            "build_ms": c2["build_ms"],
            "render_ms": c2["render_ms"],
            "hit_fraction": c2["hit_frac"],
            "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         # whether the traffic was counted under the configuration this run's autotuner chose (spills,
                         # and so traffic, differ between builds)
                         "traffic_same_config": same,
                         "algorithmic_bytes_per_launch": rb,
                         # the bound that holds: the fraction of cycles every SIMD issues a vector instruction
                         # (rocprof, same kernel)
                         "valu_busy": summary.get("kernels", {}).get("render_kernel_sq", {}).get("valu_busy")},
            "step_hbm_frac": (rb + build_bytes(4096, 64, c2["instances"])) / sec / 1e9 / HBM_PEAK_GBS,
            # the top-level build alone, the step's one bandwidth-shaped phase: its algorithmic bytes over its event time
            "build_roofline": {"bound": "hbm", "kernel": "tlas_build_kernel",
                               "achieved": bb / (c2["build_ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": bb / (c2["build_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "traffic": summary.get("kernels", {}).get("tlas_build_kernel", {}).get("hbm_bytes_per_launch"),
                               "algorithmic_bytes_per_launch": bb},
        })
    if c1:
        line.update(instanced_worlds_64x64_images_per_s=c1["images_per_s"], instanced_worlds_64x64_ms=c1["ms"],
                    instanced_worlds_reference_claim_images_per_s=c1["reference_claim_images_per_s"],
                    instanced_worlds_trace_config=c1["trace_config"])
    if c3:
        line.update(lidar_rays_per_s=c3["rays_per_s"], lidar_ms=c3["ms"], lidar_hit_fraction=c3["hit_frac"],
                    lidar_trace_config=c3["trace_config"])
    if c4:
        line.update(fwd_bwd_images_per_s=c4["fwd_bwd_images_per_s"], fwd_images_per_s=c4["fwd_images_per_s"],
                    fwd_bwd_over_fwd=c4["ratio"])
        bw = c4["backward_bytes"] / (c4["backward_kernel_ms"] * 1e-3) / 1e9
        traffic, same = _traffic(_profile_summary("depth_backward_kernel_posed_bodies_traffic.json"), "backward",
                                 c4["trace_config"])
        line["fwd_bwd_roofline"] = {
            "bound": "hbm", "kernel": "depth_backward_kernel", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": bw / HBM_PEAK_GBS, "traffic": traffic, "traffic_same_config": same,
            "algorithmic_bytes_per_launch": c4["backward_bytes"], "kernel_ms": c4["backward_kernel_ms"]}
        line["fwd_bwd_trace_config"] = c4["trace_config"]
        line["fwd_bwd_hip_graph"] = c4["hip_graph"]
    if crgb:
        line.update(rgb_depth_id_images_per_s=crgb["images_per_s"], rgb_depth_id_ms=crgb["ms"],
                    rgb_depth_id_trace_config=crgb["trace_config"])
        rgbb = render_bytes(crgb["views"], 128, 128, crgb["instances"]) + picture_bytes(crgb["views"], 128, 128, crgb["instances"])
        summary = _profile_summary("rgb_two_pass_body_worlds_traffic.json")
        traffic, same = _traffic(summary, "rgb", crgb["trace_config"])
        if traffic is not None:
            traffic += summary["kernels"]["render_persistent_kernel"]["hbm_bytes_per_launch"]
        achieved = rgbb / (crgb["kernel_ms"] * 1e-3) / 1e9
        line["rgb_depth_id_roofline"] = {
            "bound": "hbm", "kernel": "render_persistent_kernel + rgb_hit_kernel + rgb_shade_kernel (one render call)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_same_config": same, "algorithmic_bytes_per_launch": rgbb, "kernel_ms": crgb["kernel_ms"]}
    if cscene:
        line["falling_bodies_rgb"] = cscene
    line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
