"""The C ABI without a GPU: the library loads, exports every entry point include/ngx_backend.h declares, builds
bottom-level trees on the host exactly as the reference does, and rejects bad scenes before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import _lib
from neraulgraphics_amd.layouts import BLAS_NODE

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ngx_backend.h")
GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ngx_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _blas(mesh):
    lib = _lib.load()
    T = mesh.triangle_count
    nodes = np.zeros(2 * T, BLAS_NODE)
    tris = np.zeros((T, 12), np.float32)
    order = np.zeros(T, np.uint32)
    n = ctypes.c_int32()
    _lib.check(lib.ngx_build_blas_host(mesh.vertices.ctypes.data, len(mesh.vertices), mesh.indices.ctypes.data, T,
                                       nodes.ctypes.data, ctypes.byref(n), tris.ctypes.data, order.ctypes.data))
    return nodes[:n.value], tris, order


@pytest.mark.parametrize("mesh", [ng.make_icosphere(1.0, 2), ng.make_box((0.5, 0.5, 0.5)), ng.make_grid(40, 25, 2.0, 1.25, 0.05, 7)],
                         ids=["icosphere", "box", "grid"])
def test_host_bottom_level_build_is_the_references(oracle, mesh):
    """The product's host BLAS build (ngx_blas.cpp) against the oracle's restatement of bvh.cpp, which the
    golden vectors and oracle/_ref pin to the reference bit for bit."""
    nodes, tris, order = _blas(mesh)
    v, ix = mesh.vertices, mesh.indices.reshape(-1, 3)
    tb = np.concatenate([np.minimum(np.minimum(v[ix[:, 0]], v[ix[:, 1]]), v[ix[:, 2]]),
                         np.maximum(np.maximum(v[ix[:, 0]], v[ix[:, 1]]), v[ix[:, 2]])], axis=1)
    ref_nodes, ref_order = oracle.sah_build(tb, 4)
    assert nodes.tobytes() == ref_nodes.tobytes() and np.array_equal(order, ref_order)
    # the packing: v0, v1 - v0, v2 - v0 of each triangle in leaf order
    a, b, c = v[ix[order, 0]], v[ix[order, 1]], v[ix[order, 2]]
    assert np.array_equal(tris[:, 0:3], a) and np.array_equal(tris[:, 4:7], b - a) and np.array_equal(tris[:, 8:11], c - a)
    assert np.all(tris[:, [3, 7, 11]] == 0)


def test_host_bottom_level_build_matches_golden_sah():
    """The same builder on the golden bounds, through a mesh of degenerate 'triangles' spanning each box."""
    for k in range(3):
        b = GOLDEN[f"sah{k}_bounds"]
        if int(GOLDEN[f"sah{k}_leaf"]) != 4:
            continue
        verts = np.concatenate([b[:, :3], b[:, 3:], b[:, :3]], axis=0).astype(np.float32)
        n = len(b)
        idx = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], axis=1).reshape(-1).astype(np.uint32)
        nodes, _, order = _blas(ng.Mesh(verts, idx))
        assert nodes.tobytes() == GOLDEN[f"sah{k}_nodes"].tobytes() and np.array_equal(order, GOLDEN[f"sah{k}_order"])


def _scene(**kw):
    base = dict(width=32, height=24, world_count=2, meshes=[ng.make_box((1, 1, 1))], instances=[ng.Instance(0)])
    base.update(kw)
    return base


@pytest.mark.parametrize("kw,match", [
    (dict(world_count=0), "world_count"),
    (dict(cameras_per_world=0), "cameras_per_world"),
    (dict(width=0), "frame size"),
    (dict(instances=[ng.Instance(3)]), "uses mesh 3"),
    (dict(instances=[ng.Instance(0, ng.Transform(np.zeros((3, 4), np.float32)))]), "singular"),
    (dict(meshes=[ng.Mesh(np.zeros((3, 3), np.float32), np.array([0, 1, 7], np.uint32))]), "index"),
    (dict(meshes=[ng.Mesh(np.full((3, 3), np.nan, np.float32), np.array([0, 1, 2], np.uint32))]), "NaN"),
    (dict(instances_in_world=[1, 2]), "holds 2 instances"),
    (dict(spheres=[ng.Sphere((0, 0, 0), -1.0)]), "radius"),
])
def test_bad_scenes_are_rejected_before_any_device_work(kw, match):
    with pytest.raises(ValueError, match=match):
        ng.BatchRenderer(**_scene(**kw))


def test_more_triangles_than_the_walks_can_index_are_refused():
    """The batched walks pack a triangle or node index with a lane number in 32 bits; two meshes of 2^23 + 1 triangles
    (two more than 2^24 in all) are refused before any tree is built."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    big = ng.Mesh(tri, np.tile(np.arange(3, dtype=np.uint32), (1 << 23) + 1))
    with pytest.raises(NotImplementedError, match="16777218 triangles; at most 16777216"):
        ng.BatchRenderer(**_scene(meshes=[big, big], instances=[ng.Instance(0), ng.Instance(1)]))


def test_textures_are_refused_plainly():
    with pytest.raises(NotImplementedError, match="textured"):
        ng.BatchRenderer(**_scene(textures=[object()]))


@pytest.mark.parametrize("mesh", [ng.make_icosphere(1.0, 2), ng.make_box((0.5, 0.5, 0.5)), ng.make_grid(40, 25, 2.0, 1.25, 0.05, 7),
                                  ng.make_grid(3, 2, 1.0, 1.0, 0.0, 1)], ids=["icosphere", "box", "grid", "flat-grid"])
def test_skip_layout_is_an_ordered_walk_of_the_reference_tree(mesh):
    """The per-lane walk's layout (ngx_build_skip_host) against the tree it is made from: same boxes and children;
    for every octant, following first children into hits and escapes out of misses visits every node exactly once
    in pre-order, each interior node's children in the octant's order (the child lower along the node's axis
    first when the direction is positive along it), and a node's escape is the node right after its subtree."""
    from neraulgraphics_amd.layouts import SKIP_NODE
    lib = _lib.load()
    nodes, _, _ = _blas(mesh)
    T = mesh.triangle_count
    skip = np.zeros(2 * T, SKIP_NODE)
    n = ctypes.c_int32()
    _lib.check(lib.ngx_build_skip_host(mesh.vertices.ctypes.data, len(mesh.vertices), mesh.indices.ctypes.data, T,
                                       skip.ctypes.data, ctypes.byref(n)))
    skip = skip[:n.value]
    assert n.value == len(nodes)
    assert np.array_equal(skip["lo"], nodes["lo"]) and np.array_equal(skip["hi"], nodes["hi"]) and np.array_equal(skip["a"], nodes["a"])
    leaf = nodes["count"] > 0
    assert np.array_equal(skip["meta"][leaf], nodes["count"][leaf]) and np.all(skip["meta"][~leaf] < 0)
    for o in range(8):
        positive = [(o >> k) & 1 for k in range(3)]

        def children(i):
            code = -int(skip["meta"][i]) - 1
            axis, flip = code >> 1, code & 1
            a = int(skip["a"][i])
            centre = lambda c: nodes["lo"][c][axis] + nodes["hi"][c][axis]
            assert (centre(a) > centre(a + 1)) == bool(flip), "flip records which child is higher along the axis"
            lower_first = positive[axis] == 1
            return (a, a + 1) if lower_first != bool(flip) else (a + 1, a)

        # the pre-order the walk must produce when every box is entered
        want, stack = [], [0]
        while stack:
            i = stack.pop()
            want.append(i)
            if not leaf[i]:
                first, second = children(i)
                stack += [second, first]
        assert sorted(want) == list(range(len(nodes)))
        got, cur = [], 0
        while cur >= 0:
            got.append(cur)
            cur = children(cur)[0] if not leaf[cur] else int(skip["escape"][cur][o])
        assert got == want, f"octant {o}: the walk with every box entered is not the pre-order"
        # a node's escape is the first node after its subtree
        pos = {v: k for k, v in enumerate(want)}
        size = np.ones(len(nodes), np.int64)
        for i in reversed(want):
            if not leaf[i]:
                a = int(skip["a"][i])
                size[i] = 1 + size[a] + size[a + 1]
        for i in range(len(nodes)):
            after = pos[i] + size[i]
            expect = want[after] if after < len(want) else -1
            assert int(skip["escape"][i][o]) == expect, f"octant {o}: node {i}'s escape"


def test_a_missing_library_fails_loudly():
    """No CPU stand-in: without the HIP library every entry point raises, here BatchRenderer's constructor."""
    import os
    import subprocess
    import sys
    code = ("import neraulgraphics_amd as ng\n"
            "try:\n"
            "    ng.BatchRenderer(8, 8, 1, meshes=[ng.make_box()], instances=[ng.Instance(0)])\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    print('no error')\n")
    env = dict(os.environ, NGX_LIBRARY="/nonexistent/libngx.so")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert "raised:" in out.stdout and "is not built" in out.stdout, out.stdout + out.stderr
