"""The CPU code under sanitizers: the oracle (its C restatement and its thread pool) under AddressSanitizer +
UndefinedBehaviorSanitizer and under ThreadSanitizer, and the library's host bottom-level build (ngx_blas.cpp)
under AddressSanitizer + UndefinedBehaviorSanitizer with leak checking. Each leg first shows that its sanitizer is
live: a planted mistake in a few lines of C, built the same way, must be reported."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = os.path.join(ROOT, "oracle", "_san")


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip(f"no {name} in this toolchain")
    return path


def _run(cmd, env=None, timeout=600):
    return subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout, cwd=ROOT)


@pytest.fixture(scope="module")
def sanitized_oracle():
    r = _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    assert r.returncode == 0, r.stderr
    return os.path.join(SAN, "libngx_oracle_asan.so"), os.path.join(SAN, "libngx_oracle_tsan.so")


def _planted(tmp_path, source, flags, env=None):
    src = tmp_path / "planted.c"
    src.write_text(source)
    exe = tmp_path / "planted"
    r = _run(["gcc", "-O0", "-g", *flags, str(src), "-o", str(exe), "-lpthread"])
    assert r.returncode == 0, r.stderr
    return _run([str(exe)], env=env)


OUT_OF_BOUNDS = "#include <stdlib.h>\nint main(void) { int *a = malloc(8 * sizeof(int)); int s = a[8]; free(a); return s & 0; }\n"
RACE = ("#include <pthread.h>\nstatic int x;\nstatic void *bump(void *p) { (void)p; for (int i = 0; i < 100000; ++i) ++x; return 0; }\n"
        "int main(void) { pthread_t a, b; pthread_create(&a, 0, bump, 0); pthread_create(&b, 0, bump, 0);\n"
        "  pthread_join(a, 0); pthread_join(b, 0); return x & 0; }\n")


def test_the_oracle_under_address_and_undefined_behaviour_sanitizers(sanitized_oracle, tmp_path):
    planted = _planted(tmp_path, OUT_OF_BOUNDS, ["-fsanitize=address,undefined"])
    assert planted.returncode != 0 and "heap-buffer-overflow" in planted.stderr
    env = dict(os.environ, LD_PRELOAD=_runtime("libasan.so"), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", NGX_ORACLE_LIB=sanitized_oracle[0], PYTHONPATH=ROOT)
    tests = [os.path.join("tests", f) for f in ("test_oracle_pinning.py", "test_oracle_physics.py", "test_oracle_normals.py",
                                                "test_oracle_rgb.py")]
    r = _run([sys.executable, "-m", "pytest", *tests, "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"], env=env, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "Sanitizer" not in out and "runtime error" not in out, out[-4000:]


def test_the_oracle_thread_pool_under_thread_sanitizer(sanitized_oracle, tmp_path):
    planted = _planted(tmp_path, RACE, ["-fsanitize=thread"], env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert planted.returncode != 0 and "data race" in planted.stderr
    script = ("import numpy as np\n"
              "from neraulgraphics_amd import workloads\n"
              "from oracle.oracle import Oracle, forward\n"
              "o = Oracle()\n"
              "assert o.lib._name.endswith('libngx_oracle_tsan.so')\n"
              "cfg = workloads.make('body_worlds', worlds=6)\n"
              "out = forward(cfg, o, threads=8)\n"
              "lc = workloads.make('lidar_sweeps', worlds=2, rays_per_world=4096)\n"
              "st = forward(lc, o, outputs=())\n"
              "rays = o.cast_rays(st['nodes'], st['roots'], st['inv'], st['blas_root'], st['meshes']['nodes'], st['meshes']['tris'],\n"
              "                   st['meshes']['attrs'], st['world_off'], lc['origins'], lc['directions'], 100.0, threads=8)\n"
              "print('hits', int(np.isfinite(out['depth']).sum()), int(np.isfinite(rays['distance']).sum()))\n")
    env = dict(os.environ, LD_PRELOAD=_runtime("libtsan.so"), TSAN_OPTIONS="halt_on_error=1:report_signal_unsafe=0",
               NGX_ORACLE_LIB=sanitized_oracle[1], PYTHONPATH=ROOT)
    r = _run([sys.executable, "-c", script], env=env, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ThreadSanitizer" not in out and "hits" in r.stdout, out[-4000:]


def test_the_host_bottom_level_build_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("no hipcc")
    build = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(build, exist_ok=True)
    exe = os.path.join(build, "sanitize_blas")
    # host code only: no device code is built, and the sanitizers apply to the host pass
    flags = ["-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
             "-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-sanitize-recover=undefined"]
    planted_src = tmp_path / "planted.cpp"
    planted_src.write_text(OUT_OF_BOUNDS.replace("malloc(8 * sizeof(int))", "(int *)malloc(8 * sizeof(int))"))
    planted_exe = str(tmp_path / "planted")
    r = _run([HIPCC, *flags, str(planted_src), "-o", planted_exe])
    assert r.returncode == 0, r.stderr
    planted = _run([planted_exe])
    # clang's undefined-behaviour checks see the read past the allocation's size before the address check does
    assert planted.returncode != 0 and ("heap-buffer-overflow" in planted.stderr or "runtime error" in planted.stderr)
    csrc = os.path.join(ROOT, "neraulgraphics_amd", "csrc")
    r = _run([HIPCC, *flags, "-I" + os.path.join(ROOT, "include"), "-I" + csrc, os.path.join(ROOT, "tests", "cpp", "sanitize_blas.cpp"),
              os.path.join(csrc, "ngx_blas.cpp"), "-o", exe])
    assert r.returncode == 0, r.stderr
    r = _run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1"))
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "0 failed checks" in r.stdout and "Sanitizer" not in out, out[-4000:]
