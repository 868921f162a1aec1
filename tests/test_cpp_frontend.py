"""The C++ BatchRenderer over the C ABI (neraulgraphics_amd/cpp/ngx/batch.hpp), driven by tests/cpp/test_batch.cpp:
its scene checks need no GPU; the known answers and the refusals after creation run on the device."""
import os
import subprocess

import pytest

from neraulgraphics_amd import build as native


def _binary(build_if_missing):
    if not os.path.exists(native.CPP_TEST_BIN):
        if not build_if_missing:
            pytest.skip("tests/cpp/_build/test_batch was not built (neraulgraphics_amd/build.py builds it)")
        native.build_cpp_tests()
    return native.CPP_TEST_BIN


def test_cpp_front_end_refuses_bad_scenes_before_touching_a_device():
    r = subprocess.run([_binary(True), "cpu"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_front_end_known_answers(gpu):
    print("C++ front end: tests/cpp/test_batch.cpp gpu", flush=True)
    r = subprocess.run([_binary(False), "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
