"""Comparison rules shared by the parity tests.

Object IDs and hit/miss are compared exactly, except at rays where two surfaces are hit at the same
distance (to 1e-5 relative): there the reference keeps whichever it met first, which depends on the order
a walk meets them. The same holds for the normal where two triangles of one object meet such a ray at one
distance (a ray through their shared edge). Ties of either kind must stay rare: a handful per image.
Depth, normal and range are compared to 1e-5 relative (the path's stated tolerance); in practice the
backend and the oracle compute the same IEEE singles, and the tests report how many values are
bit-identical."""
import numpy as np

REL = 1e-5


def compare_images(gpu, ref, what=""):
    """gpu, ref: dicts with depth (V,H,W), object_id (V,H,W), optionally normal (V,H,W,3)."""
    d, e = np.asarray(gpu["depth"]), np.asarray(ref["depth"])
    gi, ri = np.asarray(gpu["object_id"]), np.asarray(ref["object_id"])
    fin_g, fin_r = np.isfinite(d), np.isfinite(e)
    both = fin_g & fin_r
    close = np.zeros_like(both)
    close[both] = np.abs(d[both] - e[both]) <= REL * np.abs(e[both])
    id_diff = gi != ri
    # an ID may differ only where both saw a surface at the same distance
    bad_ids = id_diff & ~(both & close)
    bad_hit = fin_g != fin_r
    bad_depth = both & ~close
    report = dict(pixels=int(d.size), hits=int(fin_r.sum()), id_ties=int((id_diff & both & close).sum()),
                  bad_ids=int(bad_ids.sum()), bad_hit=int(bad_hit.sum()), bad_depth=int(bad_depth.sum()),
                  depth_bitexact=float(np.mean(d.view(np.uint32)[both] == e.view(np.uint32)[both])) if both.any() else 1.0)
    if "normal" in ref and ref["normal"] is not None and gpu.get("normal") is not None:
        n, m = np.asarray(gpu["normal"]), np.asarray(ref["normal"])
        # per pixel: a normal off by more than the tolerance in any component, on the same object
        off = ~id_diff & np.any(np.abs(n - m) > REL * np.maximum(np.abs(m), 1.0), axis=-1)
        # two triangles of one object met at the same distance (a ray through their shared edge) are a tie too
        report["normal_ties"] = int((off & both & close).sum())
        report["bad_normal"] = int((off & ~(both & close)).sum())
    assert report["bad_ids"] == 0 and report["bad_hit"] == 0 and report["bad_depth"] == 0 and report.get("bad_normal", 0) == 0, \
        f"{what}: {report}"
    # ties are rare: a handful per image at most
    assert report["id_ties"] + report.get("normal_ties", 0) <= max(2, report["pixels"] // 5000), f"{what}: too many ties {report}"
    return report


def compare_rays(gpu, ref, what=""):
    g = dict(depth=gpu["distance"], object_id=gpu["object_id"], normal=gpu["normal"])
    r = dict(depth=ref["distance"], object_id=ref["object_id"], normal=ref["normal"])
    return compare_images(g, r, what)


def world_subtree(nodes, root, n_total):
    """Every node index reachable from root (interior nodes index < n_total - 1; leaves >= n_total - 1)."""
    out, stack = [], [root]
    while stack:
        v = stack.pop()
        out.append(v)
        if v < n_total - 1:
            stack += [int(nodes[v]["left"]), int(nodes[v]["right"])]
    return np.array(sorted(out), np.int64)


def world_subset(cfg, worlds):
    """The workload restricted to the given worlds (in that order), for checking a full-size batch against the
    oracle on a sample: every per-world array is indexed, the template and meshes are kept."""
    worlds = np.asarray(worlds, np.int64)
    out = dict(cfg)
    out["worlds"] = len(worlds)
    cpw = cfg["cameras_per_world"]
    if cfg.get("counts") is not None:
        out["counts"] = np.asarray(cfg["counts"])[worlds]
    for k in ("matrices", "body_poses", "target_poses", "origins", "directions"):
        if cfg.get(k) is not None:
            out[k] = np.ascontiguousarray(np.asarray(cfg[k])[worlds])
    if cfg.get("cameras") is not None:
        views = (worlds[:, None] * cpw + np.arange(cpw)[None, :]).reshape(-1)
        out["cameras"] = np.ascontiguousarray(np.asarray(cfg["cameras"]).reshape(-1, 12)[views])
    return out
