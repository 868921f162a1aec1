"""Pins the oracle before it is trusted as the checker.

  * against the reference's own code: golden vectors written by tests/golden/make_golden.py from the
    reference compiled from its sources, and, where oracle/_ref is built, the live reference functions;
  * the top-level build, which exists in the reference only as CUDA, by what the mathematics fixes:
    bit interleaving by loop, a library stable sort, the binary radix tree's invariants, brute force on
    tiny inputs, and the per-world decomposition the GPU kernel relies on."""
import os

import numpy as np
import pytest

import neraulgraphics_amd as ng
from neraulgraphics_amd import workloads
from oracle.oracle import forward

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


# ------------------------------------------------------------------ against the reference's own code --

def test_meshes_are_the_references_bit_for_bit():
    ico, box = ng.make_icosphere(1.0, 2), ng.make_box((0.5, 0.5, 0.5))
    assert np.array_equal(ico.vertices.view(np.uint32), GOLDEN["ico_vertices"].view(np.uint32))
    assert np.array_equal(ico.indices, GOLDEN["ico_indices"])
    assert np.array_equal(box.vertices, GOLDEN["box_vertices"]) and np.array_equal(box.indices, GOLDEN["box_indices"])


def test_make_transform_is_the_references():
    for i in range(16):
        m = ng.make_transform(GOLDEN["transform_pos"][i], GOLDEN["transform_quat"][i], float(GOLDEN["transform_scale"][i]))
        assert np.array_equal(m.matrix.reshape(12).view(np.uint32), GOLDEN["transform_out"][i].view(np.uint32))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_sah_build_is_the_references(oracle, k):
    nodes, order = oracle.sah_build(GOLDEN[f"sah{k}_bounds"], int(GOLDEN[f"sah{k}_leaf"]))
    assert nodes.tobytes() == GOLDEN[f"sah{k}_nodes"].tobytes()
    assert np.array_equal(order, GOLDEN[f"sah{k}_order"])


def test_camera_rays_are_the_references(oracle):
    W, H = (int(x) for x in GOLDEN["rays_size"])
    cams, fov = GOLDEN["rays_cams"], float(GOLDEN["rays_fov"])
    for v in range(len(cams)):
        for y in range(0, H, 3):
            for x in range(0, W, 5):
                o, d, f = oracle.camera_ray(cams[v], fov, W, H, x, y)
                assert np.array_equal(d.view(np.uint32), GOLDEN["rays_dirs"][v, y, x].view(np.uint32))
                assert np.array_equal(o, GOLDEN["rays_origins"][v, y, x]) and np.array_equal(f, GOLDEN["rays_forward"][v])


def test_closest_hits_are_the_references(oracle):
    """Every pixel's instance, triangle and distance equal the reference's mesh_intersect."""
    cfg = workloads.make("instanced_worlds", worlds=2)
    cfg.update(width=32, height=32)
    assert np.array_equal(cfg["matrices"], GOLDEN["hits_matrices"])
    out = forward(cfg, oracle)
    assert np.array_equal(out["hit_inst"].reshape(-1), GOLDEN["hits_inst"])
    assert np.array_equal(out["hit_tri"].reshape(-1), GOLDEN["hits_tri"])
    hit = GOLDEN["hits_inst"] >= 0
    d, f = GOLDEN["rays_dirs"].reshape(2, -1, 3), GOLDEN["rays_forward"]
    cos = ((d[..., 0] * f[:, None, 0] + d[..., 1] * f[:, None, 1]) + d[..., 2] * f[:, None, 2]).reshape(-1)
    depth = np.where(hit, GOLDEN["hits_t"] * cos, np.float32(np.inf)).astype(np.float32)
    assert np.array_equal(out["depth"].reshape(-1).view(np.uint32), depth.view(np.uint32))
    assert hit.mean() > 0.2


def test_sah_build_and_refit_match_the_live_reference(oracle, ref):
    rng = np.random.default_rng(5)
    for n, leaf in [(1, 1), (2, 4), (3, 1), (17, 4), (1000, 1), (2000, 4)]:
        lo = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
        b = np.concatenate([lo, lo + rng.uniform(0, 0.5, (n, 3)).astype(np.float32)], axis=1)
        (a1, o1), (a2, o2) = oracle.sah_build(b, leaf), ref.sah_build(b, leaf)
        assert a1.tobytes() == a2.tobytes() and np.array_equal(o1, o2)
        moved = b + rng.normal(0, 0.2, (n, 1)).astype(np.float32)
        r1 = oracle.sah_refit(a1, o1, moved)
        r2, cost = ref.sah_refit_cost(a2, o2, moved)
        assert r1.tobytes() == r2.tobytes() and oracle.sah_cost(r1) == np.float32(cost)


def test_traversal_matches_the_live_reference_over_both_top_levels(oracle, ref):
    """The oracle's render equals the reference's mesh_intersect per pixel, over the LBVH and over the
    reference CPU backend's per-world SAH trees."""
    cfg = workloads.make("body_worlds", worlds=6)
    for tlas in ("lbvh", "sah"):
        out = forward(cfg, oracle, tlas=tlas)
        W, H = cfg["width"], cfg["height"]
        o, d, f = ref.camera_rays(cfg["cameras"], cfg["fov"], W, H)
        world = np.repeat(np.arange(cfg["worlds"]), W * H).astype(np.int32)
        hits = ref.mesh_intersect(out["nodes"], out["roots"], out["inv"], out["blas_root"], out["meshes"]["nodes"],
                                  out["meshes"]["tris"], world, o.reshape(-1, 3), d.reshape(-1, 3))
        assert np.array_equal(hits["inst"], out["hit_inst"].reshape(-1)), tlas
        assert np.array_equal(hits["tri"], out["hit_tri"].reshape(-1)), tlas


# --------------------------------------------------- the top-level build (no CPU reference exists) --

def test_spread_bits_interleaves():
    from oracle.oracle import Oracle
    o = Oracle()
    for v in range(1024):
        want = 0
        for b in range(10):
            want |= ((v >> b) & 1) << (3 * b)
        assert o.spread_bits(v) == want


def _lbvh(oracle, name="body_worlds", worlds=24, seed=None):
    cfg = workloads.make(name, worlds=worlds, seed=seed)
    return cfg, forward(cfg, oracle, outputs=())


def test_keys_quantize_each_worlds_centre_bounds(oracle):
    cfg, out = _lbvh(oracle)
    st, off = out["state"], out["world_off"]
    keys = st["keys"]
    world = (keys >> np.uint64(32)).astype(np.int64)
    assert np.array_equal(world, np.repeat(np.arange(cfg["worlds"]), np.diff(off)))
    morton = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    q = np.zeros((len(keys), 3), np.int64)
    for b in range(10):
        for a in range(3):
            q[:, a] |= ((morton >> (3 * b + 2 - a)) & 1) << b
    c = 0.5 * (st["aabb"][:, :3] + st["aabb"][:, 3:])
    for w in range(cfg["worlds"]):
        sl = slice(off[w], off[w + 1])
        assert np.array_equal(st["world_bounds"][w, :3], c[sl].min(axis=0))
        assert np.array_equal(st["world_bounds"][w, 3:], c[sl].max(axis=0))
        for a in range(3):
            # the lowest centre maps to 0, the highest to 1023, and the order of centres is kept
            assert q[sl, a][np.argmin(c[sl, a])] == 0 and q[sl, a][np.argmax(c[sl, a])] == 1023
            o = np.argsort(c[sl, a], kind="stable")
            assert np.all(np.diff(q[sl, a][o]) >= 0)


def test_sort_is_a_stable_sort(oracle):
    _, out = _lbvh(oracle)
    st = out["state"]
    assert np.array_equal(st["sorted_ids"], np.argsort(st["keys"], kind="stable"))
    assert np.array_equal(st["sorted_keys"], st["keys"][st["sorted_ids"]])


def _check_radix_tree(st, n):
    """Invariants of the binary radix tree (Karras 2012) over the sorted keys, with equal keys told apart
    by position, plus boxes and roots."""
    nodes, keys, parent = st["nodes"], st["sorted_keys"], st["parent"]

    def prefix(i, j):
        a, b = int(keys[i]), int(keys[j])
        if a == b:
            return 64 + (32 - (i ^ j).bit_length())
        return 64 - (a ^ b).bit_length()

    def span(v):
        if v >= n - 1:
            return v - (n - 1), v - (n - 1)
        return tuple(st["range"][v])

    seen = np.zeros(2 * n - 1, np.int64)
    for i in range(n - 1):
        lo, hi = span(i)
        assert i in (lo, hi)
        l, r = int(nodes[i]["left"]), int(nodes[i]["right"])
        seen[l] += 1
        seen[r] += 1
        assert parent[l] == i and parent[r] == i
        (l0, l1), (r0, r1) = span(l), span(r)
        assert l0 == lo and r1 == hi and l1 + 1 == r0
        node_p = prefix(lo, hi)
        assert (l0 == l1 or prefix(l0, l1) > node_p) and (r0 == r1 or prefix(r0, r1) > node_p)
        assert prefix(l1, r0) == node_p  # the split is where the highest differing bit changes
        for a in range(3):
            assert nodes[i]["lo"][a] == min(nodes[l]["lo"][a], nodes[r]["lo"][a])
            assert nodes[i]["hi"][a] == max(nodes[l]["hi"][a], nodes[r]["hi"][a])
    assert seen[0] == 0 and np.all(seen[1:] == 1)  # every node but the root has exactly one parent
    for k in range(n):
        leaf = nodes[n - 1 + k]
        assert ~int(leaf["left"]) == st["sorted_ids"][k] and leaf["right"] == -1


def test_radix_tree_invariants_boxes_and_roots(oracle):
    cfg, out = _lbvh(oracle)
    st, off = out["state"], out["world_off"]
    n = int(off[-1])
    _check_radix_tree(st, n)
    from tests.helpers import world_subtree
    for w in range(cfg["worlds"]):
        sub = world_subtree(st["nodes"], int(st["roots"][w]), n)
        leaves = sub[sub >= n - 1] - (n - 1)
        assert np.array_equal(np.sort(leaves), np.arange(off[w], off[w + 1]))


def _brute_tree(keys):
    """The radix tree by recursive splitting at the highest differing bit (equal keys by position): the
    set of (lo, hi) ranges of interior nodes."""
    n = len(keys)
    ext = [(int(k) << 32) | i for i, k in enumerate(keys)]
    out = set()

    def rec(lo, hi):
        if lo == hi:
            return
        out.add((lo, hi))
        bit = (ext[lo] ^ ext[hi]).bit_length() - 1
        split = next(k for k in range(lo, hi + 1) if (ext[k] >> bit) & 1)
        rec(lo, split - 1)
        rec(split, hi)

    rec(0, n - 1)
    return out


def test_tiny_trees_equal_brute_force(oracle):
    rng = np.random.default_rng(9)
    for trial in range(200):
        n = int(rng.integers(2, 9))
        m = rng.integers(0, 6, size=(n, 3)).astype(np.float32)
        mats = np.zeros((n, 12), np.float32)
        mats[:, [0, 5, 10]] = 1.0
        mats[:, [3, 7, 11]] = m  # boxes on a coarse lattice: many equal Morton codes
        st = oracle.lbvh_build(np.array([0, n], np.int32), mats, np.tile([-0.1, -0.1, -0.1, 0.1, 0.1, 0.1], (n, 1)))
        got = set()
        for i in range(n - 1):
            got.add(tuple(int(x) for x in st["range"][i]))
        assert got == _brute_tree(st["sorted_keys"]), (trial, st["sorted_keys"])


def test_each_worlds_subtree_is_its_own_radix_tree_with_global_tie_breaks(oracle):
    """What the one-workgroup-per-world GPU build relies on: a world's subtree in the global tree is the radix
    tree of that world's keys alone (equal keys split by global position), rooted at the world's first or
    last position by the rule of the host (root_at_end), and every other interior node keeps its index."""
    for name, worlds in [("body_worlds", 40), ("instanced_worlds", 33), ("lidar_sweeps", 9)]:
        cfg, out = _lbvh(oracle, name, worlds)
        st, off = out["state"], out["world_off"]
        n = int(off[-1])
        counts = np.diff(off)
        nonempty = np.flatnonzero(counts)
        for k, w in enumerate(nonempty):
            w = int(w)
            a, b = int(off[w]), int(off[w + 1]) - 1
            if a == b:
                assert st["roots"][w] == n - 1 + a
                continue
            prev = int(nonempty[k - 1]) if k > 0 else None
            nxt = int(nonempty[k + 1]) if k + 1 < len(nonempty) else None
            p_prev = -1 if prev is None else 32 - (w ^ prev).bit_length()
            p_next = -1 if nxt is None else 32 - (w ^ nxt).bit_length()
            root = b if p_next > p_prev else a
            assert st["roots"][w] == root, (name, w)
            assert tuple(st["range"][root]) == (a, b)
            for i in range(a + 1, b):
                lo, hi = st["range"][i]
                assert a <= lo and hi <= b  # interior indices strictly inside are the world's own nodes


def test_uneven_worlds_and_empty_worlds(oracle):
    cfg = workloads.make("body_worlds", worlds=12)
    cfg["counts"] = np.array([0, 5, 1, 0, 0, 256, 8, 1, 2, 0, 3, 0], np.int32)
    out = forward(cfg, oracle, outputs=())
    st, off = out["state"], out["world_off"]
    n = int(off[-1])
    _check_radix_tree(st, n)
    assert np.all(st["roots"][cfg["counts"] == 0] == -1)
