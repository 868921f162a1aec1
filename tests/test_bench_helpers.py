"""bench.py's own arithmetic, without a GPU: the refusal of more than one GPU, the algorithmic byte counts the roofline
objects divide by, and the choice of the committed traffic counts that match the configuration a run's autotuner took."""
import copy
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode != 0 and "one GPU" in out.stdout


def test_algorithmic_bytes(bench):
    """The counts the profiles quote: a body_worlds render (8 B a pixel of depth and ID, 48 B a view, 128 B an instance),
    its top-level build (28 B a body pose, 240 B an instance) and a posed_bodies backward (8 B a pixel, 48 B a view,
    128 B an instance, 28 B a body pose, 28 B a body and part of partial sums)."""
    assert bench.render_bytes(4096, 128, 128, 534919) == 8 * 4096 * 128 * 128 + 48 * 4096 + 128 * 534919 == 605537152
    assert bench.build_bytes(4096, 64, 534919) == 28 * 4096 * 64 + 240 * 534919
    assert bench.backward_bytes(1024, 64, 64, 32768, 1024, 32, 8) == 46055424


def test_traffic_is_taken_from_the_profile_of_the_configuration_run(bench):
    summary = bench._profile_summary("render_persistent_kernel_body_worlds_traffic.json")
    profiled = summary["trace_config"]
    # the same configuration with the default variant named (0 and 7 are one build) and fields the render kernel does
    # not read changed
    chosen = dict(profiled, render_blocks=7, rgb_blocks=5, ray_blocks=6, backward_blocks=4, sort_rays=0, rgb_two_pass=0)
    assert bench._traffic(summary, "render", chosen) == (summary["hbm_bytes_per_launch"], True)
    for v in summary.get("variants", []):
        assert bench._traffic(summary, "render", v["trace_config"]) == (v["hbm_bytes_per_launch"], True)
    # a configuration nobody profiled: the main figure, flagged
    other = dict(profiled, pair_limit=7)
    assert bench._traffic(summary, "render", other) == (summary["hbm_bytes_per_launch"], False)
    other = dict(profiled, render_blocks=5)
    assert bench._traffic(summary, "render", other)[1] is False
    # the backward's two builds
    bwd = bench._profile_summary("depth_backward_kernel_posed_bodies_traffic.json")
    six = copy.deepcopy(bwd["trace_config"])
    five = dict(six, backward_blocks=5)
    got6, got5 = bench._traffic(bwd, "backward", six), bench._traffic(bwd, "backward", five)
    assert got6[1] and got5[1] and got6[0] != got5[0]
    assert bench._traffic(bwd, "backward", dict(six, backward_blocks=0)) == got5  # 0 is the 5-workgroup build
    # the RGB + depth + ID call: the profiled configuration (its shading pass at 6) and not the default 5
    rgb = bench._profile_summary("rgb_two_pass_body_worlds_traffic.json")
    chosen = copy.deepcopy(rgb["trace_config"])
    assert bench._traffic(rgb, "rgb", chosen) == (rgb["hbm_bytes_per_launch"], True)
    assert bench._traffic(rgb, "rgb", dict(chosen, render_blocks=0)) == (rgb["hbm_bytes_per_launch"], True)
    assert bench._traffic(rgb, "rgb", dict(chosen, rgb_blocks=0))[1] is False
    assert bench._traffic(rgb, "rgb", dict(chosen, rgb_two_pass=0))[1] is False


def test_picture_bytes(bench):
    """The two-pass picture on body_worlds: the hit records written (16 B a pixel, 1,073.7 MB) and read back, 3 B a
    pixel of colour, and each pass's cameras and top-level trees: the hit pass's 1,142.4 MB and the shading pass's
    1,343.7 MB that the RGB profile quotes."""
    hit = 16 * 4096 * 128 * 128 + 48 * 4096 + 128 * 534919
    shade = (16 + 3) * 4096 * 128 * 128 + 48 * 4096 + 128 * 534919
    assert (hit, shade) == (1142408064, 1343734656)
    assert bench.picture_bytes(4096, 128, 128, 534919) == hit + shade
